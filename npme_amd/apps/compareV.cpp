// compareV Vfile1 Vfile2 -- the error of the potential file Vfile1 against
// the reference file Vfile2 (app/compareV.cpp:35-131): the compareV metric
// of src/SupportFunctions.cpp:88-305, absolute and relative to the average
// magnitudes of Vfile2.  Exit status 0 on success, 1 on an error.
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "npme_amd.hpp"

using namespace NPME_Library;

int main(int argc, char **argv)
{
  if (argc != 3) {
    std::cout << "Error.\n" << argv[0] << " Vfile1 Vfile2\n";
    return 1;
  }
  const std::string f1(argv[1]), f2(argv[2]);
  bool real1 = true, real2 = true;
  size_t n1 = 0, n2 = 0;
  std::vector<double> vr1, vr2;
  std::vector<_Complex double> vc1, vc2;
  if (!NPME_ReadSingleBox_V(f1, real1, n1, vr1, vc1)) {
    std::cout << "Error reading " << f1 << "\n";
    return 1;
  }
  if (!NPME_ReadSingleBox_V(f2, real2, n2, vr2, vc2)) {
    std::cout << "Error reading " << f2 << "\n";
    return 1;
  }
  if (real1 != real2 || n1 != n2 || n1 == 0) {
    std::cout << "Error reading " << f1 << " and " << f2 << "\n"
              << (real1 != real2 ? "  isChargeReal do not match\n" : "  nCharge do not match\n");
    return 1;
  }
  double eV, edV, mV, mdV;
  if (real1) {
    NPME_CompareV1Arrays(eV, edV, n1, "pme", vr1.data(), "ref", vr2.data(), false, std::cout);
    NPME_CalcV1AvgVecMag(mV, mdV, n1, vr2.data());
  } else {
    NPME_CompareV1Arrays(eV, edV, n1, "pme", vc1.data(), "ref", vc2.data(), false, std::cout);
    NPME_CalcV1AvgVecMag(mV, mdV, n1, vc2.data());
  }
  std::printf("error_V = %.3e error_dVrdr = %.3e (abs. error)\n", eV, edV);
  std::printf("error_V = %.3e error_dVrdr = %.3e (rel. error)\n", eV / mV, edV / mdV);
  return 0;
}
