// makeRandomBox keywordFile -- writes nCharge random charges in a box
// (app/makeRandomBox.cpp:39-212).  Keywords: nCharge (required), crdFile
// (default coord_<nCharge>_<chargeType>.txt), lx ly lz (default 10),
// chargeType real|complex, chargeScale (default 1), seed (default 1),
// printLog (default 1: <crdFile without extension>.log lists the keywords).
// Coordinates are uniform in [-l/2, l/2] per axis and charges uniform in
// [-chargeScale, chargeScale] (complex: re and im independently), as in the
// reference; the random stream is SplitMix64 (seed for the coordinates,
// seed + 1 for the charges), not the reference's MKL generator, so the
// same seed gives a different box.  Exit status 0 on success, 1 on error.
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "npme_amd.hpp"

using namespace NPME_Library;

// SplitMix64 -> double in [0, 1): element i (1-based) of stream seed
static void splitmix_uniform(uint64_t seed, size_t n, double *out)
{
  for (size_t i = 0; i < n; i++) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
}

int main(int argc, char **argv)
{
  if (argc != 2) {
    std::cout << "Error.\n" << argv[0] << " keywordFile\n";
    return 1;
  }
  const std::string kwFile(argv[1]);
  NPME_KeywordInput keyword0;
  if (!keyword0.ReadInputKeyword(kwFile)) {
    std::cout << "Error reading " << kwFile << ": " << npme_last_error() << "\n";
    return 1;
  }
  NPME_KeywordInput keyword = keyword0;
  size_t n = 0;
  if (!keyword.GetValue("nCharge", n) || n == 0) {
    std::cout << "Error reading " << kwFile << " . nCharge is not defined\n";
    return 1;
  }
  std::string chargeType = "real";
  if (!keyword.GetValue("chargeType", chargeType)) keyword.AddKeyword("chargeType", chargeType);
  if (chargeType != "real" && chargeType != "complex") {
    std::cout << "Error reading " << kwFile << ".  chargeType = " << chargeType
              << " is undefined.\n  chargeType must be 'real' or 'complex'\n";
    return 1;
  }
  const bool real = chargeType == "real";
  std::string crdFile = "coord_" + std::to_string(n) + "_" + chargeType + ".txt";
  if (!keyword.GetValue("crdFile", crdFile)) keyword.AddKeyword("crdFile", crdFile);
  bool printLog = true;
  if (!keyword.GetValue("printLog", printLog)) keyword.AddKeyword("printLog", printLog);
  double l[3] = {10.0, 10.0, 10.0}, scale = 1.0;
  const char *ln[3] = {"lx", "ly", "lz"};
  for (int d = 0; d < 3; d++)
    if (!keyword.GetValue(ln[d], l[d])) keyword.AddKeyword(ln[d], l[d]);
  if (!keyword.GetValue("chargeScale", scale)) keyword.AddKeyword("chargeScale", scale);
  if (l[0] <= 0 || l[1] <= 0 || l[2] <= 0) {
    std::printf("Error.  lx ly lz = %f %f %f must be all be positive\n", l[0], l[1], l[2]);
    return 1;
  }
  size_t seed = 1;
  if (!keyword.GetValue("seed", seed)) keyword.AddKeyword("seed", seed);

  if (printLog) {
    std::string stem;
    NPME_RemoveExtension(crdFile, stem);
    std::cout << "printing " << stem << ".log\n";
    std::ofstream log((stem + ".log").c_str());
    log << "Initial Keywords:\n";
    keyword0.PrintKeywords(log);
    log << "\n\nKeywords with defaults added:\n";
    keyword.PrintKeywords(log);
  }

  std::vector<double> coord(3 * n), q((real ? 1 : 2) * n);
  splitmix_uniform(seed, coord.size(), coord.data());
  for (size_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) coord[3 * i + d] = 0.5 * l[d] * (2.0 * coord[3 * i + d] - 1.0);
  splitmix_uniform(seed + 1, q.size(), q.data());
  for (double &v : q) v = scale * (2.0 * v - 1.0);

  std::cout << "printing " << crdFile << "\n";
  if (npme_write_coord_charge(crdFile.c_str(), n, real ? 1 : 0, coord.data(), q.data()) != NPME_OK) {
    std::cout << npme_last_error() << "\n";
    return 1;
  }
  return 0;
}
