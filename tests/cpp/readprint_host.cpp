// Host-only use of the file and metric helpers of the C++ drop-in header
// (the src/ReadPrint.h and src/SupportFunctions.h names) and of the
// built-in kernel classes (host Calc, NPME_KernelFuncCheck, the kernel list),
// as the reference's command-line tools call them.  Exit status 0 when every check holds; no GPU call.
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
  // built-in kernels evaluate on the host: LR + SR = the kernel, the split
  // is continuous at Rdir, and the gradients match central differences
  NPME_Kfunc_Ralpha f(-3.0);
  NPME_Kfunc_Ralpha_LR_DM fLR(-3.0, 7, 6.5);
  NPME_Kfunc_Ralpha_SR_DM fSR(-3.0, 7, 6.5);
  {
    const size_t m = 5;
    double X[3][m], Y[3][m], Z[3][m];
    const double px[m] = {0.7, 2.0, 4.1, 6.49, 6.51}, py[m] = {0.1, -1.0, 0.5, 0.0, 0.0};
    for (int c = 0; c < 3; c++)
      for (size_t i = 0; i < m; i++) {
        X[c][i] = px[i];
        Y[c][i] = py[i];
        Z[c][i] = 0.3;
      }
    f.Calc(m, X[0], Y[0], Z[0]);
    fLR.Calc(m, X[1], Y[1], Z[1]);
    fSR.Calc(m, X[2], Y[2], Z[2]);
    for (size_t i = 0; i < m; i++)
      if (std::fabs(X[1][i] + X[2][i] - X[0][i]) > 1e-12 * std::fabs(X[0][i])) return 8;
  }
  if (!NPME_KernelFuncCheck(fLR, 64, "Ralpha_LR_DM", -6.0, 6.0, 0, true, false, std::cout)) return 9;
  if (!NPME_KernelFuncCheck(fSR, 64, "Ralpha_SR_DM", -6.0, 6.0, 0, true, false, std::cout)) return 9;
  _Complex double k0;
  __real__ k0 = 1.1;
  __imag__ k0 = 0.2;
  NPME_Kfunc_Helmholtz h(k0);
  NPME_Kfunc_Helmholtz_LR_DM hLR(k0, 8, 4.0);
  NPME_Kfunc_Helmholtz_SR_DM hSR(k0, 8, 4.0);
  if (!NPME_KernelFuncCheck(h, 64, "Helmholtz", -4.0, 4.0, 0, true, false, std::cout)) return 10;
  if (!NPME_KernelFuncCheck(hLR, 64, "Helmholtz_LR_DM", -4.0, 4.0, 0, true, false, std::cout)) return 10;
  {
    double xr[2] = {1.5, 3.9}, xi[2], y[2] = {0.2, 0.1}, z[2] = {0.0, 0.4};
    double ar[2] = {1.5, 3.9}, ai[2], br[2] = {1.5, 3.9}, bi[2];
    h.Calc(2, xr, xi, y, z);
    hLR.Calc(2, ar, ai, y, z);
    hSR.Calc(2, br, bi, y, z);
    for (int i = 0; i < 2; i++)
      if (std::hypot(ar[i] + br[i] - xr[i], ai[i] + bi[i] - xi[i]) > 1e-12 * std::hypot(xr[i], xi[i]))
        return 11;
  }
  // the built-in objects carry their family and parameters
  const NPME_BuiltinKernel *d = fSR.npme_builtin();
  if (!d || d->family != NPME_BuiltinKernel::RALPHA || d->part != NPME_BuiltinKernel::SR ||
      d->alpha != -3.0 || d->Nder != 7 || d->Rdir != 6.5)
    return 12;
  NPME_KernelList kl;
  NPME_KfuncReal *a1, *a2, *a3;
  NPME_KeywordInput kw;
  kw.AddKeyword("funcType", "Laplace");
  kw.AddKeyword("EwaldSplit", "LaplaceOrig");
  kw.AddKeyword("Rdir", 8.0);
  kw.AddKeyword("tol", 1e-6);
  if (!NPME_Interface_SelectKernelPtr(a1, a2, a3, kl, kw, true, std::cout)) return 13;
  if (a2 != &kl.funcLaplace_LR_Orig || a3 != &kl.funcLaplace_SR_Orig) return 13;
  {
    double x1 = 2.5, y1 = 0.0, z1 = 0.0, x2 = 2.5, x3 = 2.5;
    a1->Calc(1, &x1, &y1, &z1);
    a2->Calc(1, &x2, &y1, &z1);
    a3->Calc(1, &x3, &y1, &z1);
    if (std::fabs(x2 + x3 - x1) > 1e-14) return 14;
  }
  std::ostringstream ks;
  kw.PrintKeywords(ks);
  if (ks.str().rfind("@Keyword\n      EwaldSplit      LaplaceOrig", 0) != 0) return 15;
  std::printf("ok\n");
  return 0;
}
