// Shared driver of the calculation tools (npme, laplaceDM, RalphaDM,
// helmholtzDM): read a charge/coordinate file and a keyword file, add the
// missing keywords (defaults and the RSEM grid), set up, run CalcV and/or
// the exact sum as calcType asks (pme, exact, pme_exact), report the errors
// and write the potential files.
//
// Behaviour of app/npme.cpp:41-462 (and of app/laplaceDM.cpp,
// app/RalphaDM.cpp, app/helmholtzDM.cpp, which force funcType and
// EwaldSplit): the same arguments, output file names
// (<coord>_<keyword>.log, _V_pme.output, _V_exact.output) and log lines.
// Exit status: 0 on success, 1 on any error (the reference returns 1 on
// success and exits with 0 on errors).
#ifndef NPME_AMD_CALC_TOOL_HPP
#define NPME_AMD_CALC_TOOL_HPP

#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "npme_amd.hpp"

namespace npme_tools {

using namespace NPME_Library;

inline void log_time(std::ostream &os, double t, const char *what)
{
  char s[128];
  std::snprintf(s, sizeof(s), "  time = %6.2e (%s)\n", t, what);
  os << s;
}

inline void log_errors(std::ostream &os, double eV, double edV, double mV, double mdV)
{
  char s[256];
  std::snprintf(s, sizeof(s), "\n\nerrorV = %.2e errordVdr = %.2e (Absolute Error)\n", eV, edV);
  os << s;
  std::snprintf(s, sizeof(s), "errorV = %.2e errordVdr = %.2e (Relative Error)\n", eV / mV,
                edV / mdV);
  os << s;
  os.flush();
}

inline void log_mag(std::ostream &os, double mV, double mdV)
{
  char s[256];
  std::snprintf(s, sizeof(s), "V_mag  = %.2e dVdr_mag  = %.2e (Average magnitude)\n", mV, mdV);
  os << s;
  os.flush();
}

// the pme / exact / compare / print sequence for one charge type
template <typename Iface, typename Q>
int calc_and_report(Iface *npme, const std::string &calcType, bool printV, bool printLog,
                    size_t n, std::vector<Q> &q, std::vector<double> &coord, std::ostream &log,
                    const std::string &Vpme_file, const std::string &Vref_file)
{
  std::vector<Q> Vpme, Vref;
  const bool doPme = calcType == "pme" || calcType == "pme_exact";
  const bool doExact = calcType == "exact" || calcType == "pme_exact";
  if (doPme) {
    Vpme.assign(4 * n, Q(0));
    if (!npme->CalcV(Vpme.data(), n, q.data(), coord.data(), printLog, log)) {
      std::printf("Error.  npme->CalcV failed\n");
      return 1;
    }
    double mV, mdV;
    NPME_CalcV1AvgVecMag(mV, mdV, n, Vpme.data());
    if (printLog) log_mag(log, mV, mdV);
  }
  if (doExact) {
    Vref.assign(4 * n, Q(0));
    const double t0 = NPME_GetTime();
    if (!npme->CalcV_exact(Vref.data(), n, q.data(), coord.data(), false, log)) {
      std::printf("Error.  npme->CalcV_exact failed\n");
      return 1;
    }
    char s[128];
    std::snprintf(s, sizeof(s), " time_exact = %.2e\n", NPME_GetTime() - t0);
    log << s;
    log.flush();
  }
  if (calcType == "pme_exact") {
    double eV, edV, mV, mdV;
    NPME_CompareV1Arrays(eV, edV, n, "pme", Vpme.data(), "ref", Vref.data(), false, log);
    NPME_CalcV1AvgVecMag(mV, mdV, n, Vref.data());
    if (printLog) log_errors(log, eV, edV, mV, mdV);
  }
  if (printV) {
    if (doPme) {
      std::cout << "printing " << Vpme_file << "\n";
      NPME_PrintSingleBox_V(Vpme_file, n, Vpme, false);
    }
    if (doExact) {
      std::cout << "printing " << Vref_file << "\n";
      NPME_PrintSingleBox_V(Vref_file, n, Vref, false);
    }
  }
  return 0;
}

// forceFuncType: nullptr for npme (funcType from the keyword file), else
// the kernel of laplaceDM / RalphaDM / helmholtzDM
inline int run_calc(int argc, char **argv, const char *forceFuncType)
{
  const double tStart = NPME_GetTime();
  if (argc != 3) {
    std::cout << "Error.\n" << argv[0] << " coordChargeFile keywordFile\n";
    return 1;
  }
  const std::string coordFile(argv[1]), keywordFile(argv[2]);
  std::string logFile, Vpme_file, Vref_file;
  NPME_SetFilenames(logFile, Vpme_file, Vref_file, coordFile, keywordFile);

  double t0 = NPME_GetTime();
  size_t n = 0;
  bool isReal = true;
  std::vector<double> coord, qr;
  std::vector<_Complex double> qc;
  if (!NPME_ReadSingleBoxChargeCoord(coordFile, n, isReal, coord, qr, qc)) {
    std::cout << "Error reading " << coordFile << "\n";
    return 1;
  }
  const double tReadCoord = NPME_GetTime() - t0;

  t0 = NPME_GetTime();
  NPME_KeywordInput keyword;
  if (!keyword.ReadInputKeyword(keywordFile)) {
    std::cout << "Error reading " << keywordFile << ": " << npme_last_error() << "\n";
    return 1;
  }
  if (forceFuncType) {
    keyword.AddKeyword("funcType", forceFuncType);
    keyword.AddKeyword("EwaldSplit", "DerivMatch");
    const bool wantReal = std::string(forceFuncType) != "Helmholtz";
    if (wantReal != isReal) {
      std::cout << "Error reading " << coordFile << ".  " << argv[0] << " needs "
                << (wantReal ? "@ChargeReal" : "@ChargeComplex") << " charges\n";
      return 1;
    }
  }
  const double tReadInput = NPME_GetTime() - t0;

  bool printLog = true;
  keyword.GetValue("printLog", printLog);
  std::ofstream log;
  if (printLog) {
    std::cout << "printing " << logFile << "\n";
    log.open(logFile.c_str());
    keyword.PrintKeywords(log);
    NPME_PrintBoxInfo(log, n, isReal);
  }

  t0 = NPME_GetTime();
  if (!NPME_AddMissingKeywords(keywordFile.c_str(), keyword, n, isReal, coord, qr, qc, printLog,
                               log)) {
    std::cout << "Error in NPME_AddMissingKeywords\n";
    return 1;
  }
  const double tProcess = NPME_GetTime() - t0;
  if (printLog) {
    log << "\n\nKeywords with defaults added:\n";
    keyword.PrintKeywords(log);
    log_time(log, tReadCoord, "read charge+coord");
    log_time(log, tReadInput, "read input");
    log_time(log, tProcess, "process input");
  }

  std::string calcType;
  bool printV = true;
  if (!keyword.GetValue("calcType", calcType)) {
    std::cout << "Error.  'calcType' is missing from keyword\n";
    return 1;
  }
  if (calcType != "pme" && calcType != "exact" && calcType != "pme_exact") {
    std::cout << "Error.  calcType = '" << calcType << "' must be pme, exact or pme_exact\n";
    return 1;
  }
  keyword.GetValue("printV", printV);

  NPME_KernelList kernelList;
  int rc = 1;
  t0 = NPME_GetTime();
  if (isReal) {
    NPME_KfuncReal *f = nullptr, *fLR = nullptr, *fSR = nullptr;
    NPME_InterfaceReal *npme = nullptr;
    if (!NPME_Interface_SelectKernelPtr(f, fLR, fSR, kernelList, keyword, printLog, log)) {
      std::cout << "Error in NPME_Interface_SelectKernelPtr\n";
      return 1;
    }
    if (!NPME_Interface_SetUpRealKernel(npme, f, fLR, fSR, n, coord.data(), keyword, false,
                                        printLog, log)) {
      std::cout << "Error in NPME_Interface_SetUp: " << npme_last_error() << "\n";
      return 1;
    }
    if (printLog) {
      log_time(log, NPME_GetTime() - t0, "setup");
      npme->Print(log);
    }
    rc = calc_and_report(npme, calcType, printV, printLog, n, qr, coord, log, Vpme_file, Vref_file);
    delete npme;
  } else {
    NPME_KfuncComplex *f = nullptr, *fLR = nullptr, *fSR = nullptr;
    NPME_InterfaceComplex *npme = nullptr;
    if (!NPME_Interface_SelectKernelPtr(f, fLR, fSR, kernelList, keyword, printLog, log)) {
      std::cout << "Error in NPME_Interface_SelectKernelPtr\n";
      return 1;
    }
    if (!NPME_Interface_SetUpComplexKernel(npme, f, fLR, fSR, n, coord.data(), keyword, false,
                                           printLog, log)) {
      std::cout << "Error in NPME_Interface_SetUp: " << npme_last_error() << "\n";
      return 1;
    }
    if (printLog) {
      log_time(log, NPME_GetTime() - t0, "setup");
      npme->Print(log);
    }
    rc = calc_and_report(npme, calcType, printV, printLog, n, qc, coord, log, Vpme_file, Vref_file);
    delete npme;
  }
  if (printLog && rc == 0) {
    char s[128];
    std::snprintf(s, sizeof(s), "\n\nTotal_Run_Time = %.2f\n", NPME_GetTime() - tStart);
    log << s;
  }
  return rc;
}

}  // namespace npme_tools

#endif
