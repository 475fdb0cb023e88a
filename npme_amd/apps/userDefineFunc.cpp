// userDefineFunc -- the user-defined kernel example of app/userDefineFunc.cpp
// (f(x, y, z) = r, with the smooth long-range part r erf(beta r) and the
// short-range part r erfc(beta r)) written against the C++ drop-in header.
//
//   userDefineFunc                 Calc of rFunc on 8 random points, then
//                                  NPME_KernelFuncCheck, as the reference's
//                                  app does (app/userDefineFunc.cpp:102-173)
//   userDefineFunc calcv N [prec]  additionally: NPME_InterfaceReal_GenFunc
//                                  with rFunc and its two parts on N random
//                                  charges in a 20^3 box (beta 0.5, Rdir 8,
//                                  BnOrder 8, 96^3 grid), CalcV on the GPU
//                                  against CalcV_exact (the user's kernel on
//                                  the host); prints "relV <e> reldV <e>"
//
// Exit status 0 on success, 1 on a failed check (the reference's app returns
// 1 on success).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "npme_amd.hpp"

using namespace NPME_Library;

// f(r) = r and its gradient x/r (app/userDefineFunc.cpp:180-215)
class rFunc : public NPME_KfuncReal {
 public:
  void Print(std::ostream &) const override {}
  void Calc(const size_t N, double *x_f0, const double *y, const double *z) const override
  {
    for (size_t i = 0; i < N; i++) x_f0[i] = std::sqrt(std::fabs(x_f0[i] * x_f0[i] + y[i] * y[i] + z[i] * z[i]));
  }
  void Calc(const size_t N, double *f0, double *x_fX, double *y_fY, double *z_fZ) const override
  {
    for (size_t i = 0; i < N; i++) {
      const double r = std::sqrt(std::fabs(x_fX[i] * x_fX[i] + y_fY[i] * y_fY[i] + z_fZ[i] * z_fZ[i]));
      f0[i] = r;
      const double f1 = 1.0 / r;
      x_fX[i] *= f1;
      y_fY[i] *= f1;
      z_fZ[i] *= f1;
    }
  }
};

// r erf(beta r) (long range) or r erfc(beta r) (short range): f' = erf(x) +
// r g, or erfc(x) - r g, with x = beta r and g = 2 beta / sqrt(pi) e^-x^2
class rFuncSplit : public NPME_KfuncReal {
 public:
  rFuncSplit(double beta, bool longRange) : _beta(beta), _lr(longRange) {}
  void Print(std::ostream &os) const override
  {
    os << "r " << (_lr ? "erf" : "erfc") << "(beta r), beta = " << _beta << "\n";
  }
  void Calc(const size_t N, double *x_f0, const double *y, const double *z) const override
  {
    for (size_t i = 0; i < N; i++) {
      double f, d;
      eval(std::sqrt(x_f0[i] * x_f0[i] + y[i] * y[i] + z[i] * z[i]), f, d);
      x_f0[i] = f;
    }
  }
  void Calc(const size_t N, double *f0, double *x_fX, double *y_fY, double *z_fZ) const override
  {
    for (size_t i = 0; i < N; i++) {
      const double r = std::sqrt(x_fX[i] * x_fX[i] + y_fY[i] * y_fY[i] + z_fZ[i] * z_fZ[i]);
      double f, d;
      eval(r, f, d);
      f0[i] = f;
      const double u = r > 0 ? d / r : 0.0;  // the gradient vanishes at r = 0
      x_fX[i] *= u;
      y_fY[i] *= u;
      z_fZ[i] *= u;
    }
  }

 private:
  void eval(double r, double &f, double &d) const
  {
    const double x = _beta * r, g = 2.0 / std::sqrt(M_PI) * _beta * std::exp(-x * x);
    if (_lr) {
      f = r * std::erf(x);
      d = std::erf(x) + r * g;
    } else {
      f = r * std::erfc(x);
      d = std::erfc(x) - r * g;
    }
  }
  double _beta;
  bool _lr;
};

static double urand(double a, double b) { return a + (b - a) * (std::rand() / (double)RAND_MAX); }

static int calcv_demo(size_t n, const std::string &prec)
{
  std::vector<double> coord(3 * n), q(n);
  double sum = 0;
  for (size_t i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) coord[3 * i + d] = urand(0.0, 20.0);
    q[i] = urand(-1.0, 1.0);
    sum += q[i];
  }
  for (size_t i = 0; i < n; i++) q[i] -= sum / (double)n;
  NPME_KeywordInput kw;
  kw.AddKeyword("calcType", "pme");
  kw.AddKeyword("funcType", "Laplace");
  kw.AddKeyword("Rdir", 8.0);
  kw.AddKeyword("BnOrder", 8);
  kw.AddKeyword("tol", 1e-6);
  kw.AddKeyword("precision", prec);
  for (const char *k : {"N1", "N2", "N3"}) kw.AddKeyword(k, 96);
  for (const char *k : {"n1", "n2", "n3"}) kw.AddKeyword(k, 8);
  if (!NPME_AddMissingKeywords("", kw, n, coord, q, true, std::cout)) return 1;
  rFunc f;
  rFuncSplit fLR(0.5, true), fSR(0.5, false);
  NPME_InterfaceReal_GenFunc npme;
  if (!npme.SetUp(kw, n, coord.data(), &f, &fLR, &fSR, true, std::cout)) return 1;
  std::vector<double> V(4 * n), Vx(4 * n);
  if (!npme.CalcV(V.data(), n, q.data(), coord.data(), false, std::cout)) return 1;
  if (!npme.CalcV_exact(Vx.data(), n, q.data(), coord.data(), false, std::cout)) return 1;
  double eV, edV, mV, mdV, te[2];
  NPME_CompareV1Arrays(eV, edV, n, "exact", Vx.data(), "pme", V.data(), false, std::cout);
  NPME_CalcV1AvgVecMag(mV, mdV, n, Vx.data());
  npme.TableError(te);
  std::printf("table error %.2e %.2e\nrelV %.3e reldV %.3e\n", te[0], te[1], eV / mV, edV / mdV);
  return 0;
}

int main(int argc, char *argv[])
{
  char str[1000];
  const size_t N = 8;
  double x[N], y[N], z[N], f0[N], xSave[N];
  for (size_t i = 0; i < N; i++) {
    x[i] = urand(-5.0, 5.0);
    y[i] = urand(-5.0, 5.0);
    z[i] = urand(-5.0, 5.0);
    xSave[i] = x[i];
    std::snprintf(str, sizeof(str), "xyz = %8.4f %8.4f %8.4f\n", x[i], y[i], z[i]);
    std::cout << str;
  }
  rFunc func;
  func.Calc(N, &x[0], &y[0], &z[0]);
  for (size_t i = 0; i < N; i++) {
    std::snprintf(str, sizeof(str), "r = %8.4f\n", x[i]);
    std::cout << str;
  }
  for (size_t i = 0; i < N; i++) x[i] = xSave[i];
  func.Calc(N, &f0[0], &x[0], &y[0], &z[0]);
  for (size_t i = 0; i < N; i++) {
    std::snprintf(str, sizeof(str), "r = %8.4f  drdx = %8.4f  drdy = %8.4f  drdz = %8.4f\n", f0[i],
                  x[i], y[i], z[i]);
    std::cout << str;
  }
  if (!NPME_KernelFuncCheck(func, N, "r", -5.0, 5.0, 1, true, true, std::cout)) return 1;
  std::cout << "NPME_KernelFuncCheck passed\n";
  if (argc >= 3 && std::string(argv[1]) == "calcv")
    return calcv_demo((size_t)std::atol(argv[2]), argc >= 4 ? argv[3] : "mixed");
  return 0;
}
