// Host-only use of the file and metric helpers of the C++ drop-in header
// (the src/ReadPrint.h and src/SupportFunctions.h names) and of the
// parameterised kernel descriptors, as the reference's command-line tools
// call them.  Exit status 0 when every check holds; no GPU call.
#include <cmath>
#include <cstdio>
#include <fstream>

#include "npme_amd.hpp"

using namespace NPME_Library;

int main(int argc, char **argv)
{
  if (argc != 2) return 64;
  const std::string dir(argv[1]);
  const size_t n = 6;
  std::vector<double> coord(3 * n), q(n);
  std::vector<_Complex double> qc;
  for (size_t i = 0; i < 3 * n; i++) coord[i] = 0.37 * (double)i - 2.0;
  for (size_t i = 0; i < n; i++) q[i] = 0.25 * (double)i - 0.5;
  {
    std::ofstream os((dir + "/crd.txt").c_str());
    NPME_PrintSingleBoxChargeCoord(os, n, true, coord, q, qc);
  }
  size_t n2 = 0;
  std::vector<double> c2, q2;
  if (!NPME_ReadSingleBoxChargeCoordReal(dir + "/crd.txt", n2, c2, q2)) return 1;
  if (n2 != n) return 2;
  for (size_t i = 0; i < 3 * n; i++)
    if (std::fabs(c2[i] - coord[i]) > 5e-7) return 3;
  for (size_t i = 0; i < n; i++)
    if (std::fabs(q2[i] - q[i]) > 5e-7) return 4;
  std::vector<double> V(4 * n), W(4 * n);
  for (size_t i = 0; i < 4 * n; i++) V[i] = std::sin((double)i) * 10.0;
  NPME_PrintSingleBox_V(dir + "/v.out", n, V, true);
  if (!NPME_ReadSingleBox_V(dir + "/v.out", n2, W) || n2 != n) return 5;
  double eV, edV, mV, mdV;
  NPME_CompareV1Arrays(eV, edV, n, "a", W.data(), "b", V.data(), false, std::cout);
  NPME_CalcV1AvgVecMag(mV, mdV, n, V.data());
  if (eV / mV > 1e-15 || edV / mdV > 1e-15) return 6;
  std::string logF, vp, vr;
  NPME_SetFilenames(logF, vp, vr, "crdChg.txt", "laplaceDM_04.txt");
  if (logF != "crdChg_laplaceDM_04.log" || vp != "crdChg_laplaceDM_04_V_pme.output" ||
      vr != "crdChg_laplaceDM_04_V_exact.output")
    return 7;
  // descriptor parameters override the keywords
  NPME_Kfunc_Ralpha f(-3.0);
  NPME_Kfunc_Ralpha_LR_DM fLR(-3.0, 7, 6.5);
  NPME_Kfunc_Ralpha_SR_DM fSR(-3.0, 7, 6.5);
  NPME_KeywordInput kw;
  kw.AddKeyword("alpha", 1.0);
  kw.AddKeyword("Rdir", 8.0);
  const NPME_KfuncReal *fs[3] = {&f, &fLR, &fSR};
  NPME_KeywordInput k2 = npme_apply_kernel_params(kw, fs, 3);
  double a = 0, r = 0;
  int nd = 0;
  if (!k2.GetValue("alpha", a) || !k2.GetValue("Rdir", r) || !k2.GetValue("nDeriv", nd)) return 8;
  if (a != -3.0 || r != 6.5 || nd != 7) return 9;
  _Complex double k0;
  __real__ k0 = 1.1;
  __imag__ k0 = 0.2;
  NPME_Kfunc_Helmholtz_SR_DM h(k0, 8, 4.0);
  const NPME_KfuncComplex *hs[1] = {&h};
  NPME_KeywordInput k3 = npme_apply_kernel_params(kw, hs, 1);
  double kr = 0, ki = 0;
  if (!k3.GetValue("k0_r", kr) || !k3.GetValue("k0_i", ki) || kr != 1.1 || ki != 0.2) return 10;
  std::ostringstream ks;
  k3.PrintKeywords(ks);
  if (ks.str().rfind("@Keyword\n      Rdir            4", 0) != 0) return 11;
  std::printf("ok\n");
  return 0;
}
