/*
 * npme_amd.hpp -- C++ drop-in for the NPME library interface
 * (src/Interface.h, src/ReadPrint.h) over the C ABI of npme_amd.h.
 *
 * Code written against the reference's NPME_Library classes compiles
 * unchanged against this header and links with -lnpme_amd:
 *
 *   NPME_Library::NPME_KeywordInput keyword(file);
 *   NPME_Library::NPME_AddMissingKeywords(file, keyword, n, coord, q, 1, std::cout);
 *   NPME_Library::NPME_KernelList kernelList;
 *   NPME_Library::NPME_KfuncReal *f, *fLR, *fSR;
 *   NPME_Library::NPME_Interface_SelectKernelPtr(f, fLR, fSR, kernelList, keyword, 1, std::cout);
 *   NPME_Library::NPME_InterfaceReal *npme = nullptr;
 *   NPME_Library::NPME_Interface_SetUpRealKernel(npme, f, fLR, fSR, n, &coord[0], keyword,
 *                                                0, 1, std::cout);
 *   npme->CalcV(&V[0], n, &q[0], &coord[0], 1, std::cout);
 *
 * The built-in kernels (Laplace, r^alpha, Helmholtz, each with its DM
 * split) run natively on the GPU.  A user-defined NPME_KfuncReal /
 * NPME_KfuncComplex subclass (the reference's plugin interface, pure
 * virtual Calc and Print) runs through GenFunc::SetUp: its long-range part
 * is sampled on the host, its radial short-range part tabulated for the
 * device direct sum.  Failures return false and print the library's
 * message on the log stream, as the reference does.
 */
#ifndef NPME_AMD_HPP
#define NPME_AMD_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "npme_amd.h"

namespace NPME_Library {

// ------------------------------------------------------------ keywords
// NPME_KeywordInput (src/ReadPrint.h:28-102)
class NPME_KeywordInput {
 public:
  NPME_KeywordInput() : _kw(npme_keywords_create()) {}
  explicit NPME_KeywordInput(const std::string &inputFile) : _kw(npme_keywords_create())
  {
    ReadInputKeyword(inputFile);
  }
  NPME_KeywordInput(const NPME_KeywordInput &rhs) : _kw(npme_keywords_create()) { copy_from(rhs); }
  NPME_KeywordInput &operator=(const NPME_KeywordInput &rhs)
  {
    if (this != &rhs) {
      ClearKeyword();
      copy_from(rhs);
    }
    return *this;
  }
  virtual ~NPME_KeywordInput() { npme_keywords_destroy(_kw); }

  bool ReadInputKeyword(const std::string &inputFile)
  {
    return npme_keywords_read_file(_kw, inputFile.c_str()) == NPME_OK;
  }
  void ClearKeyword()
  {
    npme_keywords_destroy(_kw);
    _kw = npme_keywords_create();
  }
  // "@Keyword", then one line per keyword (name and value left-aligned
  // in 15 columns) in name order, then two blank lines
  // (src/ReadPrint.cpp:68-85)
  void PrintKeywords(std::ostream &os) const
  {
    const int n = npme_keywords_count(_kw);
    std::vector<std::string> keys;
    for (int i = 0; i < n; i++) keys.push_back(key(i));
    std::sort(keys.begin(), keys.end());
    os << "@Keyword\n";
    for (const std::string &k : keys) {
      std::vector<char> buf(1024);
      npme_keywords_get(_kw, k.c_str(), buf.data(), buf.size());
      char line[2200];
      std::snprintf(line, sizeof(line), "      %-15s %-15s\n", k.c_str(), buf.data());
      os << line;
    }
    os << "\n\n";
  }
  void PrintKeywords(const std::string &filename) const
  {
    std::ofstream ofs(filename.c_str());
    PrintKeywords(ofs);
  }

  void AddKeyword(const std::string &keyword, const std::string &value)
  {
    npme_keywords_set(_kw, keyword.c_str(), value.c_str());
  }
  template <class T1, class T2> void AddKeyword(const T1 &keyword, const T2 &value)
  {
    std::stringstream s1, s2;
    s1 << keyword;
    s2.precision(17);
    s2 << value;
    AddKeyword(s1.str(), s2.str());
  }

  // value of keyword converted with operator>>; false if absent
  template <class T> bool GetValue(const std::string &keyword, T &value) const
  {
    std::vector<char> buf(1024);
    if (!npme_keywords_get(_kw, keyword.c_str(), buf.data(), buf.size())) return false;
    std::stringstream ss(buf.data());
    ss >> value;
    return true;
  }
  template <class T> bool GetValue(const char *keyword, T &value) const
  {
    return GetValue(std::string(keyword), value);
  }

  npme_keywords *c_handle() const { return _kw; }

 private:
  std::string key(int i) const
  {
    char buf[512];
    npme_keywords_key(_kw, i, buf, sizeof(buf));
    return buf;
  }
  void copy_from(const NPME_KeywordInput &rhs)
  {
    const int n = npme_keywords_count(rhs._kw);
    for (int i = 0; i < n; i++) {
      const std::string k = rhs.key(i);
      std::vector<char> buf(1024);
      npme_keywords_get(rhs._kw, k.c_str(), buf.data(), buf.size());
      npme_keywords_set(_kw, k.c_str(), buf.data());
    }
  }
  npme_keywords *_kw;
};

inline bool npme_report(int rc, bool printLog, std::ostream &os)
{
  if (rc != NPME_OK && printLog) os << "NPME error: " << npme_last_error() << "\n";
  return rc == NPME_OK;
}

// ------------------------------------------------------------ NPME_AddMissingKeywords
// (src/Interface.h:37-51, src/Interface.cpp:53-517)
inline bool NPME_AddMissingKeywords(const char *keywordFile, NPME_KeywordInput &keyword,
                                    const size_t nCharge, const bool isChargeReal,
                                    const std::vector<double> &coord,
                                    const std::vector<double> &chargeReal,
                                    const std::vector<_Complex double> &chargeComplex,
                                    bool printLog, std::ostream &ofs_log)
{
  (void)keywordFile;
  const int rc = npme_add_missing_keywords(
      keyword.c_handle(), nCharge, isChargeReal ? 1 : 0, coord.data(),
      isChargeReal ? chargeReal.data() : nullptr,
      isChargeReal ? nullptr : reinterpret_cast<const double *>(chargeComplex.data()));
  return npme_report(rc, printLog, ofs_log);
}
inline bool NPME_AddMissingKeywords(const char *keywordFile, NPME_KeywordInput &keyword,
                                    const size_t nCharge, const std::vector<double> &coord,
                                    const std::vector<double> &chargeReal, bool printLog,
                                    std::ostream &ofs_log)
{
  return NPME_AddMissingKeywords(keywordFile, keyword, nCharge, true, coord, chargeReal,
                                 std::vector<_Complex double>(), printLog, ofs_log);
}
inline bool NPME_AddMissingKeywords(const char *keywordFile, NPME_KeywordInput &keyword,
                                    const size_t nCharge, const std::vector<double> &coord,
                                    const std::vector<_Complex double> &chargeComplex,
                                    bool printLog, std::ostream &ofs_log)
{
  return NPME_AddMissingKeywords(keywordFile, keyword, nCharge, false, coord,
                                 std::vector<double>(), chargeComplex, printLog, ofs_log);
}

// ------------------------------------------------------------ kernel plugins
// NPME_KfuncReal / NPME_KfuncComplex (src/KernelFunction.h:25-116): the
// plugin bases with the reference's pure virtual Calc and Print.  A user
// subclass defines them; the library calls Calc on the host (possibly from
// several threads at once, as the reference's OpenMP loops do, so Calc must
// be safe to call concurrently and must not throw).
//
// The built-in kernels below (src/KernelFunction{Laplace,Ralpha,Helmholtz}.h)
// evaluate on the host too, with the reference's formulas and cutoff
// tests.  When GenFunc::SetUp receives the full, long-range and short-range
// objects of one built-in family, the GPU runs its native kernels for them;
// any other combination runs the user-kernel path (host-sampled long-range
// grid, tabulated short-range part).
struct NPME_BuiltinKernel {
  enum Family { LAPLACE, RALPHA, HELMHOLTZ } family = LAPLACE;
  enum Part { FULL, LR, SR } part = FULL;
  double alpha = 0, k0_r = 0, k0_i = 0, Rdir = 0;
  int Nder = 0;
};

class NPME_KfuncReal {
 public:
  NPME_KfuncReal() {}
  virtual ~NPME_KfuncReal() {}
  NPME_KfuncReal(const NPME_KfuncReal &) {}
  NPME_KfuncReal &operator=(const NPME_KfuncReal &) { return *this; }

  virtual void Print(std::ostream &os) const = 0;
  // x_f0[N] holds x on entry and f(x, y, z) on return
  virtual void Calc(const size_t N, double *x_f0, const double *y, const double *z) const = 0;
  // f0[N] receives f; x_fX, y_fY, z_fZ hold x, y, z on entry and the
  // gradient on return
  virtual void Calc(const size_t N, double *f0, double *x_fX, double *y_fY,
                    double *z_fZ) const = 0;
  // the built-in kernel this object is, or nullptr (a user kernel)
  virtual const NPME_BuiltinKernel *npme_builtin() const { return nullptr; }
};

class NPME_KfuncComplex {
 public:
  NPME_KfuncComplex() {}
  virtual ~NPME_KfuncComplex() {}
  NPME_KfuncComplex(const NPME_KfuncComplex &) {}
  NPME_KfuncComplex &operator=(const NPME_KfuncComplex &) { return *this; }

  virtual void Print(std::ostream &os) const = 0;
  virtual void Calc(const size_t N, double *x_f0_r, double *f0_i, const double *y,
                    const double *z) const = 0;
  virtual void Calc(const size_t N, double *f0_r, double *f0_i, double *x_fX_r, double *fX_i,
                    double *y_fY_r, double *fY_i, double *z_fZ_r, double *fZ_i) const = 0;
  virtual const NPME_BuiltinKernel *npme_builtin() const { return nullptr; }
};

namespace npme_detail {
// radial kernel f(r) with g(r) = (1/r) df/dr (complex: re, im) applied to
// the reference's in-place arrays
template <typename F>
inline void real_value(size_t N, double *x_f0, const double *y, const double *z, F &&fg)
{
  for (size_t i = 0; i < N; i++) {
    const double r2 = x_f0[i] * x_f0[i] + y[i] * y[i] + z[i] * z[i];
    double f, g;
    fg(r2, f, g);
    x_f0[i] = f;
  }
}
template <typename F>
inline void real_grad(size_t N, double *f0, double *x, double *y, double *z, F &&fg)
{
  for (size_t i = 0; i < N; i++) {
    const double r2 = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
    double f, g;
    fg(r2, f, g);
    f0[i] = f;
    x[i] *= g;
    y[i] *= g;
    z[i] *= g;
  }
}
template <typename F>
inline void cx_value(size_t N, double *x_f0_r, double *f0_i, const double *y, const double *z,
                     F &&fg)
{
  for (size_t i = 0; i < N; i++) {
    const double r2 = x_f0_r[i] * x_f0_r[i] + y[i] * y[i] + z[i] * z[i];
    double f[2], g[2];
    fg(r2, f, g);
    x_f0_r[i] = f[0];
    f0_i[i] = f[1];
  }
}
template <typename F>
inline void cx_grad(size_t N, double *f0r, double *f0i, double *x, double *xi, double *y,
                    double *yi, double *z, double *zi, F &&fg)
{
  for (size_t i = 0; i < N; i++) {
    const double r2 = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
    double f[2], g[2];
    fg(r2, f, g);
    f0r[i] = f[0];
    f0i[i] = f[1];
    const double X = x[i], Y = y[i], Z = z[i];
    x[i] = X * g[0];
    xi[i] = X * g[1];
    y[i] = Y * g[0];
    yi[i] = Y * g[1];
    z[i] = Z * g[0];
    zi[i] = Z * g[1];
  }
}
// sum_k c_k r2^k, k = 0..n (c as (re, im) pairs)
inline void horner2(const double *c, int n, double r2, double out[2])
{
  double re = c[2 * n], im = c[2 * n + 1];
  for (int k = n - 1; k >= 0; k--) {
    re = re * r2 + c[2 * k];
    im = im * r2 + c[2 * k + 1];
  }
  out[0] = re;
  out[1] = im;
}
// the DM coefficients of a built-in kernel through the library
inline bool dm_fit(const char *ft, double alpha, double k0r, double k0i, int Nder, double Rdir,
                   double *a, double *b, bool PRINT, std::ostream &os)
{
  const int rc = npme_dm_coefficients(ft, alpha, k0r, k0i, Nder, Rdir, a, b);
  if (rc != NPME_OK && PRINT) os << "NPME error: " << npme_last_error() << "\n";
  return rc == NPME_OK;
}
// 1/r, r^alpha, exp(i k0 r)/r with g = (1/r) df/dr
// (src/KernelFunctionLaplace.cpp:126-150, src/KernelFunctionRalpha.cpp:57-82,
// src/KernelFunctionHelmholtz.cpp:162-200)
inline void laplace_fg(double r2, double &f, double &g)
{
  const double r = std::sqrt(r2);
  f = 1.0 / r;
  g = -f / r2;
}
inline void ralpha_fg(double alpha, double r2, double &f, double &g)
{
  f = std::pow(r2, 0.5 * alpha);
  g = alpha * f / r2;
}
inline void helmholtz_fg(double k0r, double k0i, double r2, double f[2], double g[2])
{
  const double r = std::sqrt(r2), d = std::exp(-k0i * r) / r;
  f[0] = d * std::cos(k0r * r);
  f[1] = d * std::sin(k0r * r);
  // (i k0 - 1/r) f / r
  const double ar = -k0i - 1.0 / r, ai = k0r;
  g[0] = (ar * f[0] - ai * f[1]) / r;
  g[1] = (ar * f[1] + ai * f[0]) / r;
}
}  // namespace npme_detail

// exact kernels: NPME_Kfunc_Laplace (src/KernelFunctionLaplace.h:34-78),
// NPME_Kfunc_Ralpha(alpha) (src/KernelFunctionRalpha.h:34-90),
// NPME_Kfunc_Helmholtz(k0) (src/KernelFunctionHelmholtz.h:34-100)
class NPME_Kfunc_Laplace : public NPME_KfuncReal {
 public:
  NPME_Kfunc_Laplace() { _d.family = NPME_BuiltinKernel::LAPLACE; }
  void Print(std::ostream &os) const override { os << "NPME_Kfunc_Laplace: f = 1/r\n"; }
  void Calc(const size_t N, double *x_f0, const double *y, const double *z) const override
  {
    npme_detail::real_value(N, x_f0, y, z, npme_detail::laplace_fg);
  }
  void Calc(const size_t N, double *f0, double *x, double *y, double *z) const override
  {
    npme_detail::real_grad(N, f0, x, y, z, npme_detail::laplace_fg);
  }
  const NPME_BuiltinKernel *npme_builtin() const override { return &_d; }

 private:
  NPME_BuiltinKernel _d;
};

class NPME_Kfunc_Ralpha : public NPME_KfuncReal {
 public:
  NPME_Kfunc_Ralpha() { _d.family = NPME_BuiltinKernel::RALPHA; }
  explicit NPME_Kfunc_Ralpha(double a) : NPME_Kfunc_Ralpha() { SetParm(a); }
  bool SetParm(const double a)
  {
    _d.alpha = a;
    return true;
  }
  void Print(std::ostream &os) const override { os << "NPME_Kfunc_Ralpha: alpha = " << _d.alpha << "\n"; }
  void Calc(const size_t N, double *x_f0, const double *y, const double *z) const override
  {
    const double a = _d.alpha;
    npme_detail::real_value(N, x_f0, y, z,
                            [a](double r2, double &f, double &g) { npme_detail::ralpha_fg(a, r2, f, g); });
  }
  void Calc(const size_t N, double *f0, double *x, double *y, double *z) const override
  {
    const double a = _d.alpha;
    npme_detail::real_grad(N, f0, x, y, z,
                           [a](double r2, double &f, double &g) { npme_detail::ralpha_fg(a, r2, f, g); });
  }
  const NPME_BuiltinKernel *npme_builtin() const override { return &_d; }

 private:
  NPME_BuiltinKernel _d;
};

class NPME_Kfunc_Helmholtz : public NPME_KfuncComplex {
 public:
  NPME_Kfunc_Helmholtz() { _d.family = NPME_BuiltinKernel::HELMHOLTZ; }
  explicit NPME_Kfunc_Helmholtz(_Complex double k0) : NPME_Kfunc_Helmholtz() { SetParm(k0); }
  bool SetParm(const _Complex double k0)
  {
    _d.k0_r = __real__ k0;
    _d.k0_i = __imag__ k0;
    return true;
  }
  void Print(std::ostream &os) const override
  {
    os << "NPME_Kfunc_Helmholtz: k0 = " << _d.k0_r << " + " << _d.k0_i << "i\n";
  }
  void Calc(const size_t N, double *x_f0_r, double *f0_i, const double *y,
            const double *z) const override
  {
    const double kr = _d.k0_r, ki = _d.k0_i;
    npme_detail::cx_value(N, x_f0_r, f0_i, y, z, [kr, ki](double r2, double *f, double *g) {
      npme_detail::helmholtz_fg(kr, ki, r2, f, g);
    });
  }
  void Calc(const size_t N, double *f0_r, double *f0_i, double *x, double *xi, double *y, double *yi,
            double *z, double *zi) const override
  {
    const double kr = _d.k0_r, ki = _d.k0_i;
    npme_detail::cx_grad(N, f0_r, f0_i, x, xi, y, yi, z, zi, [kr, ki](double r2, double *f, double *g) {
      npme_detail::helmholtz_fg(kr, ki, r2, f, g);
    });
  }
  const NPME_BuiltinKernel *npme_builtin() const override { return &_d; }

 private:
  NPME_BuiltinKernel _d;
};

// the derivative-match split parts: inside Rdir the long-range part is the
// even polynomial sum a_k r^2k (gradient factor sum b_k r^2k) and the
// short-range part the kernel minus it; outside, LR is the kernel and SR is
// zero.  The inside tests are the reference's per family: Laplace r <= Rdir
// (src/KernelFunctionLaplace.cpp:394-437,762-803), r^alpha r2 <= Rdir^2
// (src/KernelFunctionRalpha.cpp:343-385,718-758), Helmholtz LR r <= Rdir and
// SR r < Rdir (src/KernelFunctionHelmholtz.cpp:506-560,962-1029).
class NPME_DMSplitBase {
 protected:
  NPME_BuiltinKernel _d;
  double _a[2 * 31] = {0}, _b[2 * 31] = {0};
  bool fit(const char *ft, bool PRINT, std::ostream &os)
  {
    return npme_detail::dm_fit(ft, _d.alpha, _d.k0_r, _d.k0_i, _d.Nder, _d.Rdir, _a, _b, PRINT, os);
  }
  bool inside(double r2) const
  {
    const double R = _d.Rdir;
    if (_d.family == NPME_BuiltinKernel::RALPHA) return !(r2 > R * R);
    if (_d.family == NPME_BuiltinKernel::HELMHOLTZ && _d.part == NPME_BuiltinKernel::SR)
      return std::sqrt(r2) < R;
    return !(std::sqrt(r2) > R);
  }
  void full(double r2, double f[2], double g[2]) const
  {
    f[1] = g[1] = 0;
    if (_d.family == NPME_BuiltinKernel::LAPLACE) npme_detail::laplace_fg(r2, f[0], g[0]);
    else if (_d.family == NPME_BuiltinKernel::RALPHA) npme_detail::ralpha_fg(_d.alpha, r2, f[0], g[0]);
    else npme_detail::helmholtz_fg(_d.k0_r, _d.k0_i, r2, f, g);
  }
  void eval(double r2, double f[2], double g[2]) const
  {
    const bool in = inside(r2);
    if (_d.part == NPME_BuiltinKernel::LR) {
      if (in) {
        npme_detail::horner2(_a, _d.Nder, r2, f);
        npme_detail::horner2(_b, _d.Nder, r2, g);
      } else {
        full(r2, f, g);
      }
    } else {
      f[0] = f[1] = g[0] = g[1] = 0;
      if (in) {
        double pa[2], pb[2];
        full(r2, f, g);
        npme_detail::horner2(_a, _d.Nder, r2, pa);
        npme_detail::horner2(_b, _d.Nder, r2, pb);
        f[0] -= pa[0];
        f[1] -= pa[1];
        g[0] -= pb[0];
        g[1] -= pb[1];
      }
    }
  }
};

#define NPME_AMD_REAL_SPLIT(name, FAM, PART)                                                   \
  class name : public NPME_KfuncReal, private NPME_DMSplitBase {                                \
   public:                                                                                      \
    name()                                                                                      \
    {                                                                                           \
      _d.family = NPME_BuiltinKernel::FAM;                                                      \
      _d.part = NPME_BuiltinKernel::PART;                                                       \
      _d.alpha = -1.0;                                                                          \
    }                                                                                           \
    void Print(std::ostream &os) const override                                                 \
    {                                                                                           \
      os << #name << ": Nder = " << _d.Nder << " Rdir = " << _d.Rdir << "\n";                   \
    }                                                                                           \
    void Calc(const size_t N, double *x_f0, const double *y, const double *z) const override    \
    {                                                                                           \
      npme_detail::real_value(N, x_f0, y, z, [this](double r2, double &f, double &g) {          \
        double F[2], G[2];                                                                      \
        eval(r2, F, G);                                                                         \
        f = F[0];                                                                               \
        g = G[0];                                                                               \
      });                                                                                       \
    }                                                                                           \
    void Calc(const size_t N, double *f0, double *x, double *y, double *z) const override      \
    {                                                                                           \
      npme_detail::real_grad(N, f0, x, y, z, [this](double r2, double &f, double &g) {          \
        double F[2], G[2];                                                                      \
        eval(r2, F, G);                                                                         \
        f = F[0];                                                                               \
        g = G[0];                                                                               \
      });                                                                                       \
    }                                                                                           \
    const NPME_BuiltinKernel *npme_builtin() const override { return &_d; }
// Laplace DM parts: SetParm(Nder, Rdir) (src/KernelFunctionLaplace.h:80-212)
#define NPME_AMD_LAPLACE_DM(name, PART)                                                         \
  NPME_AMD_REAL_SPLIT(name, LAPLACE, PART)                                                      \
    name(int nd, double rd) : name() { SetParm(nd, rd, false, std::cout); }                     \
    bool SetParm(const int nd, const double rd, bool PRINT, std::ostream &os)                   \
    {                                                                                           \
      _d.Nder = nd;                                                                             \
      _d.Rdir = rd;                                                                             \
      return fit("Laplace", PRINT, os);                                                         \
    }                                                                                           \
  };
// r^alpha DM parts: SetParm(alpha, Nder, Rdir) (src/KernelFunctionRalpha.h:94-200)
#define NPME_AMD_RALPHA_DM(name, PART)                                                          \
  NPME_AMD_REAL_SPLIT(name, RALPHA, PART)                                                       \
    name(double a, int nd, double rd) : name() { SetParm(a, nd, rd, false, std::cout); }        \
    bool SetParm(const double a, const int nd, const double rd, bool PRINT, std::ostream &os)   \
    {                                                                                           \
      _d.alpha = a;                                                                             \
      _d.Nder = nd;                                                                             \
      _d.Rdir = rd;                                                                             \
      return fit("Ralpha", PRINT, os);                                                          \
    }                                                                                           \
  };
NPME_AMD_LAPLACE_DM(NPME_Kfunc_Laplace_LR_DM, LR)
NPME_AMD_LAPLACE_DM(NPME_Kfunc_Laplace_SR_DM, SR)
NPME_AMD_RALPHA_DM(NPME_Kfunc_Ralpha_LR_DM, LR)
NPME_AMD_RALPHA_DM(NPME_Kfunc_Ralpha_SR_DM, SR)
#undef NPME_AMD_LAPLACE_DM
#undef NPME_AMD_RALPHA_DM
#undef NPME_AMD_REAL_SPLIT

// Helmholtz DM parts: SetParm(k0, Nder, Rdir)
// (src/KernelFunctionHelmholtz.h:102-212)
#define NPME_AMD_HELMHOLTZ_DM(name, PART)                                                       \
  class name : public NPME_KfuncComplex, private NPME_DMSplitBase {                             \
   public:                                                                                      \
    name()                                                                                      \
    {                                                                                           \
      _d.family = NPME_BuiltinKernel::HELMHOLTZ;                                                \
      _d.part = NPME_BuiltinKernel::PART;                                                       \
    }                                                                                           \
    name(_Complex double k0, int nd, double rd) : name() { SetParm(k0, nd, rd, false, std::cout); } \
    bool SetParm(const _Complex double k0, const int nd, const double rd, bool PRINT,           \
                 std::ostream &os)                                                              \
    {                                                                                           \
      _d.k0_r = __real__ k0;                                                                    \
      _d.k0_i = __imag__ k0;                                                                    \
      _d.Nder = nd;                                                                             \
      _d.Rdir = rd;                                                                             \
      return fit("Helmholtz", PRINT, os);                                                       \
    }                                                                                           \
    void Print(std::ostream &os) const override                                                 \
    {                                                                                           \
      os << #name << ": k0 = " << _d.k0_r << " + " << _d.k0_i << "i Nder = " << _d.Nder        \
         << " Rdir = " << _d.Rdir << "\n";                                                      \
    }                                                                                           \
    void Calc(const size_t N, double *x_f0_r, double *f0_i, const double *y,                    \
              const double *z) const override                                                   \
    {                                                                                           \
      npme_detail::cx_value(N, x_f0_r, f0_i, y, z,                                              \
                            [this](double r2, double *f, double *g) { eval(r2, f, g); });       \
    }                                                                                           \
    void Calc(const size_t N, double *f0_r, double *f0_i, double *x, double *xi, double *y,     \
              double *yi, double *z, double *zi) const override                                 \
    {                                                                                           \
      npme_detail::cx_grad(N, f0_r, f0_i, x, xi, y, yi, z, zi,                                  \
                           [this](double r2, double *f, double *g) { eval(r2, f, g); });        \
    }                                                                                           \
    const NPME_BuiltinKernel *npme_builtin() const override { return &_d; }                     \
  };
NPME_AMD_HELMHOLTZ_DM(NPME_Kfunc_Helmholtz_LR_DM, LR)
NPME_AMD_HELMHOLTZ_DM(NPME_Kfunc_Helmholtz_SR_DM, SR)
#undef NPME_AMD_HELMHOLTZ_DM

// the erf/erfc split of 1/r, SetParm(beta): LR = erf(beta r)/r and
// SR = erfc(beta r)/r with no cutoff of its own
// (src/KernelFunctionLaplace.h:214-320, src/KernelFunctionLaplace.cpp:
// 1087-1116, 1344-1373).  They run on the GPU as user kernels (tabulated
// short-range part) when passed to GenFunc::SetUp.
class NPME_Kfunc_Laplace_LR_Original : public NPME_KfuncReal {
 public:
  NPME_Kfunc_Laplace_LR_Original() {}
  explicit NPME_Kfunc_Laplace_LR_Original(double beta) { SetParm(beta); }
  void SetParm(const double beta) { _beta = beta; }
  void Print(std::ostream &os) const override { os << "NPME_Kfunc_Laplace_LR_Original: beta = " << _beta << "\n"; }
  void Calc(const size_t N, double *x_f0, const double *y, const double *z) const override
  {
    const double b = _beta;
    npme_detail::real_value(N, x_f0, y, z, [b](double r2, double &f, double &g) { fg(b, r2, f, g); });
  }
  void Calc(const size_t N, double *f0, double *x, double *y, double *z) const override
  {
    const double b = _beta;
    npme_detail::real_grad(N, f0, x, y, z, [b](double r2, double &f, double &g) { fg(b, r2, f, g); });
  }
  // erf(x)/r with x = beta r, and (1/r) d/dr, by series below x = 1e-4
  static void fg(double beta, double r2, double &f, double &g)
  {
    const double r = std::sqrt(r2), x = beta * r, c = 2.0 / std::sqrt(M_PI);
    if (x < 1e-4) {
      f = beta * c * (1.0 - x * x / 3.0 + x * x * x * x / 10.0);
      g = beta * beta * beta * c * (-2.0 / 3.0 + 0.4 * x * x);
      return;
    }
    f = std::erf(x) / r;
    g = (beta * c * std::exp(-x * x) - f) / r2;
  }

 private:
  double _beta = 0;
};
class NPME_Kfunc_Laplace_SR_Original : public NPME_KfuncReal {
 public:
  NPME_Kfunc_Laplace_SR_Original() {}
  explicit NPME_Kfunc_Laplace_SR_Original(double beta) { SetParm(beta); }
  void SetParm(const double beta) { _beta = beta; }
  void Print(std::ostream &os) const override { os << "NPME_Kfunc_Laplace_SR_Original: beta = " << _beta << "\n"; }
  void Calc(const size_t N, double *x_f0, const double *y, const double *z) const override
  {
    const double b = _beta;
    npme_detail::real_value(N, x_f0, y, z, [b](double r2, double &f, double &g) { fg(b, r2, f, g); });
  }
  void Calc(const size_t N, double *f0, double *x, double *y, double *z) const override
  {
    const double b = _beta;
    npme_detail::real_grad(N, f0, x, y, z, [b](double r2, double &f, double &g) { fg(b, r2, f, g); });
  }
  static void fg(double beta, double r2, double &f, double &g)
  {
    const double r = std::sqrt(r2), x = beta * r;
    f = std::erfc(x) / r;
    g = -(beta * 2.0 / std::sqrt(M_PI) * std::exp(-x * x) + f) / r2;
  }

 private:
  double _beta = 0;
};

// NPME_KernelFuncCheck (src/KernelFunction.h:118-131,
// src/KernelFunction.cpp:36-372): the two forms of Calc agree (rms 1e-12)
// and the analytic gradient matches central differences with h = 1e-5
// (rms 1e-6) at N random points in (Xmin, Xmax)^3.  vecOption is accepted
// and ignored (there are no AVX forms to compare).
inline void npme_check_rms(double &error, size_t N, const double *a, const double *b)
{
  error = 0;
  for (size_t i = 0; i < N; i++) error += (a[i] - b[i]) * (a[i] - b[i]);
  error = std::sqrt(error / (double)N);
}
inline bool NPME_KernelFuncCheck(NPME_KfuncReal &func, const size_t N, const char *funcName,
                                 const double Xmin, const double Xmax, int vecOption, bool PRINT,
                                 bool PRINT_ALL, std::ostream &os)
{
  (void)vecOption;
  const double tol = 1.0e-12, tolNume = 1.0e-6, h = 1.0e-5;
  std::vector<double> x(N), y(N), z(N), f0(N), fX(N), fY(N), fZ(N), fRef(N), fp(N), fm(N), num(N);
  for (size_t i = 0; i < N; i++) {
    x[i] = Xmin + (Xmax - Xmin) * (std::rand() / (double)RAND_MAX);
    y[i] = Xmin + (Xmax - Xmin) * (std::rand() / (double)RAND_MAX);
    z[i] = Xmin + (Xmax - Xmin) * (std::rand() / (double)RAND_MAX);
  }
  fRef = x;
  func.Calc(N, fRef.data(), y.data(), z.data());
  fX = x;
  fY = y;
  fZ = z;
  func.Calc(N, f0.data(), fX.data(), fY.data(), fZ.data());
  double err;
  npme_check_rms(err, N, fRef.data(), f0.data());
  if (PRINT_ALL) os << "NPME_KernelFuncCheck for " << funcName << ": f0 error " << err << "\n";
  if (err > tol) {
    if (PRINT) os << "Error in NPME_KernelFuncCheck: error in f0 = " << err << " > " << tol << "\n";
    return false;
  }
  const std::vector<double> *an[3] = {&fX, &fY, &fZ};
  for (int p = 0; p < 3; p++) {
    std::vector<double> xp = x, yp = y, zp = z, xm = x, ym = y, zm = z;
    std::vector<double> &vp = p == 0 ? xp : p == 1 ? yp : zp, &vm = p == 0 ? xm : p == 1 ? ym : zm;
    for (size_t i = 0; i < N; i++) {
      vp[i] += h;
      vm[i] -= h;
    }
    func.Calc(N, xp.data(), yp.data(), zp.data());
    func.Calc(N, xm.data(), ym.data(), zm.data());
    for (size_t i = 0; i < N; i++) num[i] = (xp[i] - xm[i]) / (2 * h);
    npme_check_rms(err, N, an[p]->data(), num.data());
    if (PRINT_ALL) os << "  d/d" << "xyz"[p] << " numerical derivative error " << err << "\n";
    if (err > tolNume) {
      if (PRINT)
        os << "Error in NPME_KernelFuncCheck: numerical derivative error in f" << "XYZ"[p] << " = "
           << err << " > " << tolNume << "\n";
      return false;
    }
  }
  return true;
}
inline bool NPME_KernelFuncCheck(NPME_KfuncComplex &func, const size_t N, const char *funcName,
                                 const double Xmin, const double Xmax, int vecOption, bool PRINT,
                                 bool PRINT_ALL, std::ostream &os)
{
  (void)vecOption;
  const double tol = 1.0e-12, tolNume = 1.0e-6, h = 1.0e-5;
  std::vector<double> x(N), y(N), z(N);
  for (size_t i = 0; i < N; i++) {
    x[i] = Xmin + (Xmax - Xmin) * (std::rand() / (double)RAND_MAX);
    y[i] = Xmin + (Xmax - Xmin) * (std::rand() / (double)RAND_MAX);
    z[i] = Xmin + (Xmax - Xmin) * (std::rand() / (double)RAND_MAX);
  }
  std::vector<double> vr = x, vi(N), f0r(N), f0i(N), g[6];
  func.Calc(N, vr.data(), vi.data(), y.data(), z.data());
  g[0] = x;
  g[2] = y;
  g[4] = z;
  g[1].resize(N);
  g[3].resize(N);
  g[5].resize(N);
  func.Calc(N, f0r.data(), f0i.data(), g[0].data(), g[1].data(), g[2].data(), g[3].data(),
            g[4].data(), g[5].data());
  double er, ei;
  npme_check_rms(er, N, vr.data(), f0r.data());
  npme_check_rms(ei, N, vi.data(), f0i.data());
  const double err = std::sqrt(er * er + ei * ei);
  if (PRINT_ALL) os << "NPME_KernelFuncCheck for " << funcName << ": f0 error " << err << "\n";
  if (err > tol) {
    if (PRINT) os << "Error in NPME_KernelFuncCheck: error in f0 = " << err << " > " << tol << "\n";
    return false;
  }
  for (int p = 0; p < 3; p++) {
    std::vector<double> xp = x, yp = y, zp = z, xm = x, ym = y, zm = z, ip(N), im(N);
    std::vector<double> &vp = p == 0 ? xp : p == 1 ? yp : zp, &vm = p == 0 ? xm : p == 1 ? ym : zm;
    for (size_t i = 0; i < N; i++) {
      vp[i] += h;
      vm[i] -= h;
    }
    func.Calc(N, xp.data(), ip.data(), yp.data(), zp.data());
    func.Calc(N, xm.data(), im.data(), ym.data(), zm.data());
    std::vector<double> nr(N), ni(N);
    for (size_t i = 0; i < N; i++) {
      nr[i] = (xp[i] - xm[i]) / (2 * h);
      ni[i] = (ip[i] - im[i]) / (2 * h);
    }
    double e1, e2;
    npme_check_rms(e1, N, g[2 * p].data(), nr.data());
    npme_check_rms(e2, N, g[2 * p + 1].data(), ni.data());
    const double e = std::sqrt(e1 * e1 + e2 * e2);
    if (PRINT_ALL) os << "  d/d" << "xyz"[p] << " numerical derivative error " << e << "\n";
    if (e > tolNume) {
      if (PRINT)
        os << "Error in NPME_KernelFuncCheck: numerical derivative error in f" << "XYZ"[p] << " = " << e
           << " > " << tolNume << "\n";
      return false;
    }
  }
  return true;
}

// the kernel objects of src/Interface.h:757-771
struct NPME_KernelList {
  NPME_Kfunc_Laplace funcLaplace;
  NPME_Kfunc_Laplace_LR_DM funcLaplace_LR_DM;
  NPME_Kfunc_Laplace_SR_DM funcLaplace_SR_DM;
  NPME_Kfunc_Laplace_LR_Original funcLaplace_LR_Orig;
  NPME_Kfunc_Laplace_SR_Original funcLaplace_SR_Orig;
  NPME_Kfunc_Ralpha funcRalpha;
  NPME_Kfunc_Ralpha_LR_DM funcRalpha_LR_DM;
  NPME_Kfunc_Ralpha_SR_DM funcRalpha_SR_DM;
  NPME_Kfunc_Helmholtz funcHelmholtz;
  NPME_Kfunc_Helmholtz_LR_DM funcHelmholtz_LR_DM;
  NPME_Kfunc_Helmholtz_SR_DM funcHelmholtz_SR_DM;
};

// NPME_Interface_SelectKernelPtr (src/Interface.cpp:545-733): picks the
// kernels of the keywords' funcType and EwaldSplit from the list and sets
// their parameters from the keywords (nDeriv and Rdir for DerivMatch, beta
// from Rdir and tol for LaplaceOrig, alpha, k0)
inline bool NPME_Interface_SelectKernelPtr(NPME_KfuncReal *&func, NPME_KfuncReal *&func_LR,
                                           NPME_KfuncReal *&func_SR, NPME_KernelList &kl,
                                           const NPME_KeywordInput &keyword, bool printLog,
                                           std::ostream &os)
{
  func = func_LR = func_SR = nullptr;
  std::string ft, split;
  int nd = 0;
  double Rdir = 0, tol = 0, beta = 0, alpha = 0;
  if (!keyword.GetValue("funcType", ft) || !keyword.GetValue("EwaldSplit", split)) {
    if (printLog) os << "NPME error: funcType and EwaldSplit must be defined\n";
    return false;
  }
  if (!keyword.GetValue("Rdir", Rdir)) {
    if (printLog) os << "NPME error: 'Rdir' keyword is missing\n";
    return false;
  }
  if (split == "DerivMatch") {
    if (!keyword.GetValue("nDeriv", nd)) {
      if (printLog) os << "NPME error: 'nDeriv' keyword is missing\n";
      return false;
    }
  } else if (split == "LaplaceOrig") {
    if (!keyword.GetValue("tol", tol) || npme_ewald_rdir2beta(Rdir, tol, &beta) != NPME_OK) {
      if (printLog) os << "NPME error: LaplaceOrig needs Rdir and tol\n";
      return false;
    }
  } else {
    if (printLog) os << "NPME error: EwaldSplit = " << split << " is not defined\n";
    return false;
  }
  if (ft == "Laplace") {
    func = &kl.funcLaplace;
    if (split == "DerivMatch") {
      if (!kl.funcLaplace_LR_DM.SetParm(nd, Rdir, printLog, os) ||
          !kl.funcLaplace_SR_DM.SetParm(nd, Rdir, printLog, os))
        return false;
      func_LR = &kl.funcLaplace_LR_DM;
      func_SR = &kl.funcLaplace_SR_DM;
    } else {
      kl.funcLaplace_LR_Orig.SetParm(beta);
      kl.funcLaplace_SR_Orig.SetParm(beta);
      func_LR = &kl.funcLaplace_LR_Orig;
      func_SR = &kl.funcLaplace_SR_Orig;
    }
    return true;
  }
  if (ft == "Ralpha" && split == "DerivMatch") {
    if (!keyword.GetValue("alpha", alpha)) {
      if (printLog) os << "NPME error: 'alpha' keyword is missing\n";
      return false;
    }
    kl.funcRalpha.SetParm(alpha);
    if (!kl.funcRalpha_LR_DM.SetParm(alpha, nd, Rdir, printLog, os) ||
        !kl.funcRalpha_SR_DM.SetParm(alpha, nd, Rdir, printLog, os))
      return false;
    func = &kl.funcRalpha;
    func_LR = &kl.funcRalpha_LR_DM;
    func_SR = &kl.funcRalpha_SR_DM;
    return true;
  }
  if (printLog) os << "NPME error: funcType '" << ft << "' with EwaldSplit '" << split
                   << "' is not a real kernel\n";
  return false;
}
inline bool NPME_Interface_SelectKernelPtr(NPME_KfuncComplex *&func, NPME_KfuncComplex *&func_LR,
                                           NPME_KfuncComplex *&func_SR, NPME_KernelList &kl,
                                           const NPME_KeywordInput &keyword, bool printLog,
                                           std::ostream &os)
{
  func = func_LR = func_SR = nullptr;
  std::string ft, split;
  int nd = 0;
  double Rdir = 0, k0r = 0, k0i = 0;
  keyword.GetValue("funcType", ft);
  keyword.GetValue("EwaldSplit", split);
  if (ft != "Helmholtz" || split != "DerivMatch") {
    if (printLog) os << "NPME error: funcType '" << ft << "' with EwaldSplit '" << split
                     << "' is not a complex kernel\n";
    return false;
  }
  if (!keyword.GetValue("nDeriv", nd) || !keyword.GetValue("Rdir", Rdir) ||
      !keyword.GetValue("k0_r", k0r) || !keyword.GetValue("k0_i", k0i)) {
    if (printLog) os << "NPME error: Helmholtz needs nDeriv, Rdir, k0_r, k0_i\n";
    return false;
  }
  _Complex double k0;
  __real__ k0 = k0r;
  __imag__ k0 = k0i;
  kl.funcHelmholtz.SetParm(k0);
  if (!kl.funcHelmholtz_LR_DM.SetParm(k0, nd, Rdir, printLog, os) ||
      !kl.funcHelmholtz_SR_DM.SetParm(k0, nd, Rdir, printLog, os))
    return false;
  func = &kl.funcHelmholtz;
  func_LR = &kl.funcHelmholtz_LR_DM;
  func_SR = &kl.funcHelmholtz_SR_DM;
  return true;
}

namespace npme_detail {
// C callbacks over the virtual Calc of a plugin object (ctx)
inline void tramp_value_r(void *ctx, size_t n, double *x_f0, const double *y, const double *z)
{
  static_cast<const NPME_KfuncReal *>(ctx)->Calc(n, x_f0, y, z);
}
inline void tramp_grad_r(void *ctx, size_t n, double *f0, double *x, double *y, double *z)
{
  static_cast<const NPME_KfuncReal *>(ctx)->Calc(n, f0, x, y, z);
}
inline void tramp_value_c(void *ctx, size_t n, double *x_f0_r, double *f0_i, const double *y,
                          const double *z)
{
  static_cast<const NPME_KfuncComplex *>(ctx)->Calc(n, x_f0_r, f0_i, y, z);
}
inline void tramp_grad_c(void *ctx, size_t n, double *f0r, double *f0i, double *x, double *xi,
                         double *y, double *yi, double *z, double *zi)
{
  static_cast<const NPME_KfuncComplex *>(ctx)->Calc(n, f0r, f0i, x, xi, y, yi, z, zi);
}
// the three objects are the full, long-range and short-range parts of one
// built-in family (then the GPU's native kernels run)
template <typename K> inline bool one_family(const K *f, const K *lr, const K *sr)
{
  const NPME_BuiltinKernel *a = f->npme_builtin(), *b = lr->npme_builtin(), *c = sr->npme_builtin();
  return a && b && c && a->family == b->family && b->family == c->family &&
         a->part == NPME_BuiltinKernel::FULL && b->part == NPME_BuiltinKernel::LR &&
         c->part == NPME_BuiltinKernel::SR;
}
}  // namespace npme_detail

// ------------------------------------------------------------ interfaces
// NPME_InterfaceBase getters (src/Interface.h:55-115)
class NPME_InterfaceBase {
 public:
  NPME_InterfaceBase() {}
  NPME_InterfaceBase(const NPME_InterfaceBase &) = delete;
  NPME_InterfaceBase &operator=(const NPME_InterfaceBase &) = delete;
  virtual ~NPME_InterfaceBase() { npme_destroy(_h); }

  size_t GetNumCharge() const { return _n; }
  double GetRdir() const { return param().Rdir; }
  long Get_BnOrder() const { return param().Bn; }
  long Get_N1() const { return param().N[0]; }
  long Get_N2() const { return param().N[1]; }
  long Get_N3() const { return param().N[2]; }
  long Get_n1() const { return param().nb[0]; }
  long Get_n2() const { return param().nb[1]; }
  long Get_n3() const { return param().nb[2]; }
  int Get_Nder() const { return param().Nder; }
  double DeviceGB() const { return param().gb; }
  npme_handle *c_handle() const { return _h; }

 protected:
  struct Params {
    long N[3] = {0, 0, 0}, nb[3] = {0, 0, 0}, Bn = 0;
    int Nder = 0;
    double Rdir = 0, gb = 0;
  };
  Params param() const
  {
    Params p;
    if (_h) npme_get_params(_h, p.N, p.nb, &p.Bn, &p.Nder, &p.Rdir, &p.gb);
    return p;
  }
  bool setup(const NPME_KeywordInput &keyword, size_t nCharge, const double *coord,
             bool printLog, std::ostream &os)
  {
    npme_destroy(_h);
    _h = nullptr;
    _n = 0;
    const int rc = npme_create(&_h, keyword.c_handle(), nCharge, coord);
    if (rc != NPME_OK) {
      _h = nullptr;
      return npme_report(rc, printLog, os);
    }
    _n = nCharge;
    return true;
  }
  bool reset(size_t nCharge, const double *coord, bool printLog, std::ostream &os)
  {
    const int rc = npme_reset_coord(_h, nCharge, coord);
    if (rc == NPME_OK) _n = nCharge;
    return npme_report(rc, printLog, os);
  }
  void print(std::ostream &os) const
  {
    std::vector<char> buf(8192);
    npme_print(_h, buf.data(), buf.size());
    os << buf.data();
  }
  // CalcV with PRINT: the device times of the direct sum (gather + pairs)
  // and of the reciprocal sum (spreading, FFT passes, interpolation), and
  // the wall time of the call, as the lines of src/Interface.cpp:2051-2095
  template <typename F> bool timed_calcv(const char *name, F &&call, std::ostream &os)
  {
    npme_set_profiling(_h, 1);
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = call();
    const double tot = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double ms[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int ns = rc == NPME_OK ? npme_get_stage_times(_h, ms, 9) : 0;
    npme_set_profiling(_h, 0);
    if (rc != NPME_OK) return npme_report(rc, true, os);
    double rec = 0;
    for (int k = 2; k < ns; k++) rec += ms[k];
    char str[256];
    os << "\n\n" << name << "::CalcV\n";
    std::snprintf(str, sizeof(str), " time = %6.2e (direct sum)\n", (ms[0] + ms[1]) * 1e-3);
    os << str;
    std::snprintf(str, sizeof(str), " time = %6.2e (rec sum)\n", rec * 1e-3);
    os << str;
    std::snprintf(str, sizeof(str), " time = %6.2e (total V1)\n\n\n", tot);
    os << str;
    return true;
  }
  npme_handle *_h = nullptr;
  size_t _n = 0;
};

// NPME_InterfaceReal (src/Interface.h:168-222)
class NPME_InterfaceReal : public NPME_InterfaceBase {
 public:
  virtual bool ResetCoord(const size_t nCharge, const double *coord, bool printLog,
                          std::ostream &ofs_log)
  {
    return reset(nCharge, coord, printLog, ofs_log);
  }
  // V1[nCharge][4] = (V, dV/dx, dV/dy, dV/dz) (src/Interface.h:203-205)
  virtual bool CalcV(double *V1, const size_t nCharge, double *charge, double *coord, bool PRINT,
                     std::ostream &ofs_log)
  {
    auto call = [&] { return npme_calcv_real(_h, V1, nCharge, charge, coord); };
    if (PRINT) return timed_calcv("NPME_InterfaceReal", call, ofs_log);
    return npme_report(call(), true, ofs_log);
  }
  virtual bool CalcV_exact(double *V1, const size_t nCharge, double *charge, double *coord,
                           bool PRINT, std::ostream &ofs_log)
  {
    (void)PRINT;
    return npme_report(npme_calcv_exact_real(_h, V1, nCharge, charge, coord), true, ofs_log);
  }
  virtual void Print(std::ostream &ofs_log) { print(ofs_log); }
};

// NPME_InterfaceReal_GenFunc::SetUp (src/Interface.h:291-296,
// src/Interface.cpp:1606-1671).  The objects are borrowed and must outlive
// this one.  The full, long-range and short-range objects of one built-in
// family run the GPU's native kernels with their parameters; any other
// objects run as user kernels: f_LR sampled on the host on the kernel grid,
// f_SR (which must be radial) tabulated for the device direct sum, and
// CalcV_exact evaluating func on the host.
class NPME_InterfaceReal_GenFunc : public NPME_InterfaceReal {
 public:
  bool SetUp(const NPME_KeywordInput &keyword, const size_t nCharge, const double *coord,
             NPME_KfuncReal *func, NPME_KfuncReal *funcLR, NPME_KfuncReal *funcSR,
             bool printLog, std::ostream &ofs_log)
  {
    const char *miss = !func ? "func" : !funcLR ? "funcLR" : !funcSR ? "funcSR" : nullptr;
    if (miss) {
      if (printLog) ofs_log << "Error in NPME_InterfaceReal_GenFunc::SetUp\n  " << miss << " == NULL\n";
      return false;
    }
    if (npme_detail::one_family<NPME_KfuncReal>(func, funcLR, funcSR)) {
      const NPME_BuiltinKernel &d = *funcLR->npme_builtin();
      NPME_KeywordInput k = keyword;
      k.AddKeyword("funcType", d.family == NPME_BuiltinKernel::RALPHA ? "Ralpha" : "Laplace");
      k.AddKeyword("EwaldSplit", "DerivMatch");
      if (d.family == NPME_BuiltinKernel::RALPHA) k.AddKeyword("alpha", d.alpha);
      k.AddKeyword("nDeriv", d.Nder);
      k.AddKeyword("Rdir", d.Rdir);
      return setup(k, nCharge, coord, printLog, ofs_log);
    }
    npme_kfunc_real c[3];
    NPME_KfuncReal *f[3] = {func, funcLR, funcSR};
    for (int i = 0; i < 3; i++) c[i] = {f[i], npme_detail::tramp_value_r, npme_detail::tramp_grad_r};
    npme_destroy(_h);
    _h = nullptr;
    _n = 0;
    const int rc = npme_create_genfunc_real(&_h, keyword.c_handle(), nCharge, coord, &c[0], &c[1], &c[2]);
    if (rc != NPME_OK) {
      _h = nullptr;
      return npme_report(rc, printLog, ofs_log);
    }
    _n = nCharge;
    return true;
  }
  // errors of the short-range table of user kernels (0 for built-in ones)
  void TableError(double err[2]) const { npme_get_table_error(_h, err); }
};

// NPME_InterfaceReal_Laplace_DM::SetUp (src/Interface.h:393-395)
class NPME_InterfaceReal_Laplace_DM : public NPME_InterfaceReal {
 public:
  bool SetUp(const NPME_KeywordInput &keyword, const size_t nCharge, const double *coord,
             bool printLog, std::ostream &ofs_log)
  {
    return setup(keyword, nCharge, coord, printLog, ofs_log);
  }
};

// NPME_InterfaceReal_Laplace_Original::SetUp (src/Interface.h:413-505,
// src/Interface.cpp:2152-2230): 1/r with the erf/erfc split, beta from Rdir
// and tol (EwaldSplit LaplaceOrig)
class NPME_InterfaceReal_Laplace_Original : public NPME_InterfaceReal {
 public:
  bool SetUp(const NPME_KeywordInput &keyword, const size_t nCharge, const double *coord,
             bool printLog, std::ostream &ofs_log)
  {
    NPME_KeywordInput k = keyword;
    k.AddKeyword("funcType", "Laplace");
    k.AddKeyword("EwaldSplit", "LaplaceOrig");
    return setup(k, nCharge, coord, printLog, ofs_log);
  }
};

// NPME_InterfaceComplex (src/Interface.h:506-560)
class NPME_InterfaceComplex : public NPME_InterfaceBase {
 public:
  virtual bool ResetCoord(const size_t nCharge, const double *coord, bool printLog,
                          std::ostream &ofs_log)
  {
    return reset(nCharge, coord, printLog, ofs_log);
  }
  // V1[nCharge][4] complex (src/Interface.h:542-544)
  virtual bool CalcV(_Complex double *V1, const size_t nCharge, _Complex double *charge,
                     double *coord, bool PRINT, std::ostream &ofs_log)
  {
    auto call = [&] {
      return npme_calcv_complex(_h, reinterpret_cast<double *>(V1), nCharge,
                                reinterpret_cast<const double *>(charge), coord);
    };
    if (PRINT) return timed_calcv("NPME_InterfaceComplex", call, ofs_log);
    return npme_report(call(), true, ofs_log);
  }
  virtual bool CalcV_exact(_Complex double *V1, const size_t nCharge, _Complex double *charge,
                           double *coord, bool PRINT, std::ostream &ofs_log)
  {
    (void)PRINT;
    return npme_report(npme_calcv_exact_complex(_h, reinterpret_cast<double *>(V1), nCharge,
                                                reinterpret_cast<const double *>(charge), coord),
                       true, ofs_log);
  }
  virtual void Print(std::ostream &ofs_log) { print(ofs_log); }
};

// NPME_InterfaceComplex_Helmholtz_DM::SetUp (src/Interface.h:733-735)
class NPME_InterfaceComplex_Helmholtz_DM : public NPME_InterfaceComplex {
 public:
  bool SetUp(const NPME_KeywordInput &keyword, const size_t nCharge, const double *coord,
             bool printLog, std::ostream &ofs_log)
  {
    return setup(keyword, nCharge, coord, printLog, ofs_log);
  }
};

// NPME_InterfaceComplex_GenFunc::SetUp (src/Interface.h:563-650): as the
// real GenFunc, for complex kernels (the built-in Helmholtz family runs
// natively)
class NPME_InterfaceComplex_GenFunc : public NPME_InterfaceComplex {
 public:
  bool SetUp(const NPME_KeywordInput &keyword, const size_t nCharge, const double *coord,
             NPME_KfuncComplex *func, NPME_KfuncComplex *funcLR, NPME_KfuncComplex *funcSR,
             bool printLog, std::ostream &ofs_log)
  {
    const char *miss = !func ? "func" : !funcLR ? "funcLR" : !funcSR ? "funcSR" : nullptr;
    if (miss) {
      if (printLog) ofs_log << "Error in NPME_InterfaceComplex_GenFunc::SetUp\n  " << miss << " == NULL\n";
      return false;
    }
    if (npme_detail::one_family<NPME_KfuncComplex>(func, funcLR, funcSR)) {
      const NPME_BuiltinKernel &d = *funcLR->npme_builtin();
      NPME_KeywordInput k = keyword;
      k.AddKeyword("funcType", "Helmholtz");
      k.AddKeyword("EwaldSplit", "DerivMatch");
      k.AddKeyword("k0_r", d.k0_r);
      k.AddKeyword("k0_i", d.k0_i);
      k.AddKeyword("nDeriv", d.Nder);
      k.AddKeyword("Rdir", d.Rdir);
      return setup(k, nCharge, coord, printLog, ofs_log);
    }
    npme_kfunc_complex c[3];
    NPME_KfuncComplex *f[3] = {func, funcLR, funcSR};
    for (int i = 0; i < 3; i++) c[i] = {f[i], npme_detail::tramp_value_c, npme_detail::tramp_grad_c};
    npme_destroy(_h);
    _h = nullptr;
    _n = 0;
    const int rc =
        npme_create_genfunc_complex(&_h, keyword.c_handle(), nCharge, coord, &c[0], &c[1], &c[2]);
    if (rc != NPME_OK) {
      _h = nullptr;
      return npme_report(rc, printLog, ofs_log);
    }
    _n = nCharge;
    return true;
  }
  void TableError(double err[2]) const { npme_get_table_error(_h, err); }
};

// NPME_Interface_SetUpRealKernel / SetUpComplexKernel
// (src/Interface.h:800-829): allocates with new; the caller deletes
inline bool NPME_Interface_SetUpRealKernel(NPME_InterfaceReal *&npme, NPME_KfuncReal *func,
                                           NPME_KfuncReal *funcLR, NPME_KfuncReal *funcSR,
                                           const size_t nCharge, const double *coord,
                                           const NPME_KeywordInput &keyword,
                                           bool useGenericKernel, bool printLog,
                                           std::ostream &ofs_log)
{
  // src/Interface.cpp:884-1060: Laplace + DerivMatch -> Laplace_DM,
  // Laplace + LaplaceOrig -> Laplace_Original, Ralpha + DerivMatch and
  // useGenericKernel -> GenFunc with the given kernel objects
  npme = nullptr;
  const char *miss = !func ? "func" : !funcLR ? "funcLR" : !funcSR ? "funcSR" : nullptr;
  if (miss) {
    if (printLog) ofs_log << "Error in NPME_Interface_SetUpRealKernel\n  " << miss << " == NULL\n";
    return false;
  }
  std::string ft, split;
  if (!keyword.GetValue("funcType", ft) || !keyword.GetValue("EwaldSplit", split)) {
    if (printLog) ofs_log << "Error in NPME_Interface_SetUpRealKernel\n  funcType and EwaldSplit must be defined\n";
    return false;
  }
  const bool known = (ft == "Laplace" && (split == "DerivMatch" || split == "LaplaceOrig")) ||
                     (ft == "Ralpha" && split == "DerivMatch");
  if (!known) {
    if (printLog) ofs_log << "Error in NPME_Interface_SetUpRealKernel\n  funcType = " << ft
                          << " and EwaldSplit = " << split << " is not provided\n";
    return false;
  }
  NPME_InterfaceReal *p = nullptr;
  bool ok;
  if (ft == "Laplace" && !useGenericKernel && split == "DerivMatch") {
    auto *q = new NPME_InterfaceReal_Laplace_DM;
    ok = q->SetUp(keyword, nCharge, coord, printLog, ofs_log);
    p = q;
  } else if (ft == "Laplace" && !useGenericKernel) {
    auto *q = new NPME_InterfaceReal_Laplace_Original;
    ok = q->SetUp(keyword, nCharge, coord, printLog, ofs_log);
    p = q;
  } else {
    auto *q = new NPME_InterfaceReal_GenFunc;
    ok = q->SetUp(keyword, nCharge, coord, func, funcLR, funcSR, printLog, ofs_log);
    p = q;
  }
  if (!ok) {
    delete p;
    return false;
  }
  npme = p;
  return true;
}
inline bool NPME_Interface_SetUpComplexKernel(NPME_InterfaceComplex *&npme,
                                              NPME_KfuncComplex *func, NPME_KfuncComplex *funcLR,
                                              NPME_KfuncComplex *funcSR, const size_t nCharge,
                                              const double *coord,
                                              const NPME_KeywordInput &keyword,
                                              bool useGenericKernel, bool printLog,
                                              std::ostream &ofs_log)
{
  // src/Interface.cpp:1061-1200: Helmholtz + DerivMatch -> Helmholtz_DM, or
  // GenFunc with the given kernel objects when useGenericKernel
  npme = nullptr;
  const char *miss = !func ? "func" : !funcLR ? "funcLR" : !funcSR ? "funcSR" : nullptr;
  if (miss) {
    if (printLog) ofs_log << "Error in NPME_Interface_SetUpComplexKernel\n  " << miss << " == NULL\n";
    return false;
  }
  std::string ft, split;
  keyword.GetValue("funcType", ft);
  keyword.GetValue("EwaldSplit", split);
  if (ft != "Helmholtz" || split != "DerivMatch") {
    if (printLog) ofs_log << "Error in NPME_Interface_SetUpComplexKernel\n  funcType = " << ft
                          << " and EwaldSplit = " << split << " is not provided\n";
    return false;
  }
  NPME_InterfaceComplex *p = nullptr;
  bool ok;
  if (!useGenericKernel) {
    auto *q = new NPME_InterfaceComplex_Helmholtz_DM;
    ok = q->SetUp(keyword, nCharge, coord, printLog, ofs_log);
    p = q;
  } else {
    auto *q = new NPME_InterfaceComplex_GenFunc;
    ok = q->SetUp(keyword, nCharge, coord, func, funcLR, funcSR, printLog, ofs_log);
    p = q;
  }
  if (!ok) {
    delete p;
    return false;
  }
  npme = p;
  return true;
}

// ------------------------------------------------------------ files and metric
// src/ReadPrint.h and the helpers of src/SupportFunctions.h that the
// reference's command-line tools use.  Files go through the library's
// readers and writers (npme_read_coord_charge ...).

inline double NPME_GetTime()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// "name.ext" -> "name" (src/ReadPrint.cpp:179-189: up to the first '.');
// the directory part of a path is kept whole
inline void NPME_RemoveExtension(const std::string &s1, std::string &s2)
{
  const size_t slash = s1.find_last_of('/');
  const size_t from = slash == std::string::npos ? 0 : slash + 1;
  const size_t dot = s1.find('.', from);
  s2 += s1.substr(0, dot == std::string::npos ? std::string::npos : dot);
}

// <coord>_<keyword>.log / _V_pme.output / _V_exact.output
// (src/ReadPrint.cpp:197-213)
inline void NPME_SetFilenames(std::string &logFile, std::string &V_pme_File, std::string &V_ref_File,
                              const std::string &coordFile, const std::string &keywordFile)
{
  std::string a, b;
  NPME_RemoveExtension(coordFile, a);
  const size_t slash = keywordFile.find_last_of('/');
  NPME_RemoveExtension(slash == std::string::npos ? keywordFile : keywordFile.substr(slash + 1), b);
  logFile = a + "_" + b + ".log";
  V_pme_File = a + "_" + b + "_V_pme.output";
  V_ref_File = a + "_" + b + "_V_exact.output";
}

inline void NPME_PrintBoxInfo(std::ostream &os, const size_t nCharge, const bool isChargeReal)
{
  os << "@BoxInfo\n  nCharge       = " << nCharge << "\n  isChargeReal  = " << (int)isChargeReal
     << "\n";
  os.flush();
}

// src/ReadPrint.cpp:282-440 (returns false and prints the message on failure)
inline bool NPME_ReadSingleBoxChargeCoord(const std::string &filename, size_t &nCharge,
                                          bool &isChargeReal, std::vector<double> &coord,
                                          std::vector<double> &chargeReal,
                                          std::vector<_Complex double> &chargeComplex)
{
  size_t n = 0;
  int real = 1;
  double *c = nullptr, *q = nullptr;
  if (npme_read_coord_charge(filename.c_str(), &n, &real, &c, &q) != NPME_OK) {
    std::cout << npme_last_error() << "\n";
    return false;
  }
  nCharge = n;
  isChargeReal = real != 0;
  coord.assign(c, c + 3 * n);
  if (real) {
    chargeReal.assign(q, q + n);
  } else {
    chargeComplex.resize(n);
    std::memcpy(static_cast<void *>(chargeComplex.data()), q, 2 * n * sizeof(double));
  }
  npme_free(c);
  npme_free(q);
  return true;
}
inline bool NPME_ReadSingleBoxChargeCoordReal(const std::string &filename, size_t &nCharge,
                                              std::vector<double> &coord, std::vector<double> &charge)
{
  bool real = true;
  std::vector<_Complex double> cc;
  if (!NPME_ReadSingleBoxChargeCoord(filename, nCharge, real, coord, charge, cc)) return false;
  if (!real) std::cout << "Error in NPME_ReadSingleBoxChargeCoordReal.  isChargeReal = 0\n";
  return real;
}
inline bool NPME_ReadSingleBoxChargeCoordComplex(const std::string &filename, size_t &nCharge,
                                                 std::vector<double> &coord,
                                                 std::vector<_Complex double> &charge)
{
  bool real = false;
  std::vector<double> cr;
  if (!NPME_ReadSingleBoxChargeCoord(filename, nCharge, real, coord, cr, charge)) return false;
  if (real) std::cout << "Error in NPME_ReadSingleBoxChargeCoordComplex.  isChargeReal = 1\n";
  return !real;
}

// src/ReadPrint.cpp:234-280, to a stream
inline void NPME_PrintSingleBoxChargeCoord(std::ostream &os, const size_t nCharge,
                                           const bool isChargeReal, const std::vector<double> &coord,
                                           const std::vector<double> &chargeReal,
                                           const std::vector<_Complex double> &chargeComplex)
{
  char str[128];
  os << "@Coord\n  nCharge " << nCharge << "\n";
  for (size_t i = 0; i < nCharge; i++) {
    std::snprintf(str, sizeof(str), "    %10.6f %10.6f %10.6f\n", coord[3 * i], coord[3 * i + 1],
                  coord[3 * i + 2]);
    os << str;
  }
  os << "\n\n" << (isChargeReal ? "@ChargeReal\n" : "@ChargeComplex\n") << "  nCharge " << nCharge;
  const double *q = isChargeReal ? chargeReal.data()
                                 : reinterpret_cast<const double *>(chargeComplex.data());
  const size_t tot = (isChargeReal ? 1 : 2) * nCharge;
  for (size_t k = 0; k < tot; k++) {
    if (k % 5 == 0) os << "\n    ";
    std::snprintf(str, sizeof(str), "%14.6e ", q[k]);
    os << str;
  }
  os << "\n\n";
}

// src/ReadPrint.cpp:487-560
inline void NPME_PrintSingleBox_V(const std::string &filename, const size_t nCharge,
                                  const std::vector<double> &Vreal, bool printHighPrecision)
{
  if (npme_write_v(filename.c_str(), nCharge, 1, Vreal.data(), printHighPrecision) != NPME_OK)
    std::cout << npme_last_error() << "\n";
}
inline void NPME_PrintSingleBox_V(const std::string &filename, const size_t nCharge,
                                  const std::vector<_Complex double> &Vcomplex,
                                  bool printHighPrecision)
{
  if (npme_write_v(filename.c_str(), nCharge, 0, reinterpret_cast<const double *>(Vcomplex.data()),
                   printHighPrecision) != NPME_OK)
    std::cout << npme_last_error() << "\n";
}

// src/ReadPrint.cpp:565-739
inline bool NPME_ReadSingleBox_V(const std::string &filename, bool &isChargeReal, size_t &nCharge,
                                 std::vector<double> &Vreal, std::vector<_Complex double> &Vcomplex)
{
  size_t n = 0;
  int real = 1;
  double *v = nullptr;
  if (npme_read_v(filename.c_str(), &n, &real, &v) != NPME_OK) {
    std::cout << npme_last_error() << "\n";
    return false;
  }
  nCharge = n;
  isChargeReal = real != 0;
  if (real) {
    Vreal.assign(v, v + 4 * n);
  } else {
    Vcomplex.resize(4 * n);
    std::memcpy(static_cast<void *>(Vcomplex.data()), v, 8 * n * sizeof(double));
  }
  npme_free(v);
  return true;
}
inline bool NPME_ReadSingleBox_V(const std::string &filename, size_t &nCharge,
                                 std::vector<double> &V)
{
  bool real = true;
  std::vector<_Complex double> vc;
  return NPME_ReadSingleBox_V(filename, real, nCharge, V, vc) && real;
}
inline bool NPME_ReadSingleBox_V(const std::string &filename, size_t &nCharge,
                                 std::vector<_Complex double> &V)
{
  bool real = false;
  std::vector<double> vr;
  return NPME_ReadSingleBox_V(filename, real, nCharge, vr, V) && !real;
}

// src/SupportFunctions.cpp:88-121
inline void NPME_CalcV1AvgVecMag(double &V_mag, double &dVdr_mag, const size_t nCharge,
                                 const double *V1)
{
  double o[4];
  npme_compare_v(nCharge, 1, V1, V1, o);
  V_mag = o[2];
  dVdr_mag = o[3];
}
inline void NPME_CalcV1AvgVecMag(double &V_mag, double &dVdr_mag, const size_t nCharge,
                                 const _Complex double *V1)
{
  double o[4];
  const double *v = reinterpret_cast<const double *>(V1);
  npme_compare_v(nCharge, 0, v, v, o);
  V_mag = o[2];
  dVdr_mag = o[3];
}

// src/SupportFunctions.cpp:176-305 (PRINT lists every entry of both arrays)
inline void NPME_CompareV1Arrays(double &errorV, double &errordVdr, const size_t nCharge,
                                 const char *desc1, const double *V1, const char *desc2,
                                 const double *V2, bool PRINT, std::ostream &os)
{
  double o[4] = {0, 0, 0, 0};
  npme_compare_v(nCharge, 1, V1, V2, o);
  errorV = o[0];
  errordVdr = o[1];
  if (!PRINT) return;
  const char *nm[4] = {"V0", "VX", "VY", "VZ"};
  char str[256];
  for (size_t i = 0; i < nCharge; i++) {
    double e = 0;
    for (int p = 0; p < 4; p++) {
      std::snprintf(str, sizeof(str), "%s_%s[%4zu] = %15.12f\n%s_%s[%4zu] = %15.12f\n\n", desc1,
                    nm[p], i, V1[4 * i + p], desc2, nm[p], i, V2[4 * i + p]);
      os << str;
      const double d = V1[4 * i + p] - V2[4 * i + p];
      e += d * d;
    }
    std::snprintf(str, sizeof(str), "error = %.2e\n\n", std::sqrt(e / 4));
    os << str;
  }
}
inline void NPME_CompareV1Arrays(double &errorV, double &errordVdr, const size_t nCharge,
                                 const char *desc1, const _Complex double *V1, const char *desc2,
                                 const _Complex double *V2, bool PRINT, std::ostream &os)
{
  const double *a = reinterpret_cast<const double *>(V1), *b = reinterpret_cast<const double *>(V2);
  double o[4] = {0, 0, 0, 0};
  npme_compare_v(nCharge, 0, a, b, o);
  errorV = o[0];
  errordVdr = o[1];
  if (!PRINT) return;
  const char *nm[4] = {"V0", "VX", "VY", "VZ"};
  char str[320];
  for (size_t i = 0; i < nCharge; i++) {
    double e = 0;
    for (int p = 0; p < 4; p++) {
      const size_t k = 2 * (4 * i + p);
      std::snprintf(str, sizeof(str),
                    "%s_%s[%4zu] = %15.12f + %15.12fi\n%s_%s[%4zu] = %15.12f + %15.12fi\n\n", desc1,
                    nm[p], i, a[k], a[k + 1], desc2, nm[p], i, b[k], b[k + 1]);
      os << str;
      e += (a[k] - b[k]) * (a[k] - b[k]) + (a[k + 1] - b[k + 1]) * (a[k + 1] - b[k + 1]);
    }
    std::snprintf(str, sizeof(str), "error = %.2e\n\n", std::sqrt(e / 4));
    os << str;
  }
}

// ------------------------------------------------------------ two-box grids
// NPME_RecSumGrid_FullTrans_Func (src/RecSumGrid.h:99-112): f(r) =
// func(|r + R0|) on the full N1 x N2 x N3 grid.  A built-in full kernel
// object (NPME_Kfunc_Laplace / _Ralpha / _Helmholtz) is sampled on the GPU,
// any other object through its Calc on the host.  nProc is not used.  Where
// the reference exits on a too small separation this returns false.
namespace npme_detail {
inline bool fulltrans_builtin(const NPME_BuiltinKernel &d, _Complex double *f, const long N[3],
                              const double a[3], const double del[3], const double X[3],
                              const double R0[3], bool PRINT, std::ostream &os)
{
  NPME_KeywordInput k;
  if (d.family == NPME_BuiltinKernel::LAPLACE) {
    k.AddKeyword("funcType", "Laplace");
  } else if (d.family == NPME_BuiltinKernel::RALPHA) {
    k.AddKeyword("funcType", "Ralpha");
    k.AddKeyword("alpha", d.alpha);
  } else {
    k.AddKeyword("funcType", "Helmholtz");
    k.AddKeyword("k0_r", d.k0_r);
    k.AddKeyword("k0_i", d.k0_i);
  }
  const int rc = npme_recsumgrid_fulltrans_func(reinterpret_cast<double *>(f), N[0], a[0], del[0], N[1],
                                                a[1], del[1], N[2], a[2], del[2], X[0], X[1], X[2],
                                                k.c_handle(), R0);
  if (rc != NPME_OK) os << "Error in NPME_RecSumGrid_FullTrans_Func.\n" << npme_last_error() << "\n";
  else if (PRINT) os << "\nNPME_RecSumGrid_FullTrans_Func\n";
  return rc == NPME_OK;
}
}  // namespace npme_detail

inline bool NPME_RecSumGrid_FullTrans_Func(_Complex double *f, const long int N1, const double a1,
                                           const double del1, const long int N2, const double a2,
                                           const double del2, const long int N3, const double a3,
                                           const double del3, const double X, const double Y,
                                           const double Z, const NPME_KfuncReal &func,
                                           const double R0[3], int nProc, bool PRINT,
                                           std::ostream &os)
{
  (void)nProc;
  const long N[3] = {N1, N2, N3};
  const double a[3] = {a1, a2, a3}, del[3] = {del1, del2, del3}, Xb[3] = {X, Y, Z};
  const NPME_BuiltinKernel *d = func.npme_builtin();
  if (d && d->part == NPME_BuiltinKernel::FULL)
    return npme_detail::fulltrans_builtin(*d, f, N, a, del, Xb, R0, PRINT, os);
  const npme_kfunc_real c = {const_cast<NPME_KfuncReal *>(&func), npme_detail::tramp_value_r,
                             npme_detail::tramp_grad_r};
  const int rc = npme_recsumgrid_fulltrans_func_user(reinterpret_cast<double *>(f), N1, a1, del1, N2, a2,
                                                     del2, N3, a3, del3, X, Y, Z, &c, nullptr, R0);
  if (rc != NPME_OK) os << "Error in NPME_RecSumGrid_FullTrans_Func.\n" << npme_last_error() << "\n";
  else if (PRINT) os << "\nNPME_RecSumGrid_FullTrans_Func\n";
  return rc == NPME_OK;
}

inline bool NPME_RecSumGrid_FullTrans_Func(_Complex double *f, const long int N1, const double a1,
                                           const double del1, const long int N2, const double a2,
                                           const double del2, const long int N3, const double a3,
                                           const double del3, const double X, const double Y,
                                           const double Z, const NPME_KfuncComplex &func,
                                           const double R0[3], int nProc, bool PRINT,
                                           std::ostream &os)
{
  (void)nProc;
  const long N[3] = {N1, N2, N3};
  const double a[3] = {a1, a2, a3}, del[3] = {del1, del2, del3}, Xb[3] = {X, Y, Z};
  const NPME_BuiltinKernel *d = func.npme_builtin();
  if (d && d->part == NPME_BuiltinKernel::FULL)
    return npme_detail::fulltrans_builtin(*d, f, N, a, del, Xb, R0, PRINT, os);
  const npme_kfunc_complex c = {const_cast<NPME_KfuncComplex *>(&func), npme_detail::tramp_value_c,
                                npme_detail::tramp_grad_c};
  const int rc = npme_recsumgrid_fulltrans_func_user(reinterpret_cast<double *>(f), N1, a1, del1, N2, a2,
                                                     del2, N3, a3, del3, X, Y, Z, nullptr, &c, R0);
  if (rc != NPME_OK) os << "Error in NPME_RecSumGrid_FullTrans_Func.\n" << npme_last_error() << "\n";
  else if (PRINT) os << "\nNPME_RecSumGrid_FullTrans_Func\n";
  return rc == NPME_OK;
}

// NPME_RSTP_CalcTheta_FullF_ReverseSign (src/RecSumV.h:102-113): the compact
// theta T from the compact charge grid Qr and the full DFT F of the kernel
// grid, on the GPU.  The scratch X and the twiddles C1..C3 (fixed by N) are
// not used, nor is nProc.  The reference returns nothing and cannot fail;
// here a failure (no device, bad sizes) prints the error and leaves T as it
// was.
inline void NPME_RSTP_CalcTheta_FullF_ReverseSign(
    _Complex double *T, _Complex double *X, const _Complex double *F, const _Complex double *Qr,
    const long int N1, const long int N2, const long int N3, const _Complex double *C1,
    const _Complex double *C2, const _Complex double *C3, const double *lamda1_2,
    const double *lamda2_2, const double *lamda3_2, bool reverseSignX, bool reverseSignY,
    bool reverseSignZ, int nProc, bool PRINT, std::ostream &os)
{
  (void)X;
  (void)C1;
  (void)C2;
  (void)C3;
  (void)nProc;
  const int rc = npme_rstp_calctheta_fullF_reversesign(
      reinterpret_cast<double *>(T), reinterpret_cast<const double *>(F),
      reinterpret_cast<const double *>(Qr), N1, N2, N3, lamda1_2, lamda2_2, lamda3_2,
      reverseSignX ? 1 : 0, reverseSignY ? 1 : 0, reverseSignZ ? 1 : 0);
  if (rc != NPME_OK) os << "Error in NPME_RSTP_CalcTheta_FullF_ReverseSign.\n" << npme_last_error() << "\n";
  else if (PRINT) os << "\n    NPME_RSTP_CalcT_FullF\n";
}

// A box pair (no reference counterpart; the reference has the two functions
// above and no driver): V at the targets of one box from the sources of
// another, entirely on the translated kernel grid.  The kernel is the
// keywords' funcType (func == nullptr), a built-in full kernel object, or a
// user kernel; a user kernel object is borrowed and must outlive the pair.
class NPME_BoxPair {
 public:
  NPME_BoxPair() {}
  NPME_BoxPair(const NPME_BoxPair &) = delete;
  NPME_BoxPair &operator=(const NPME_BoxPair &) = delete;
  ~NPME_BoxPair() { npme_destroy(_h); }

  bool SetUp(const NPME_KeywordInput &keyword, size_t nTarget, const double *coordTarget,
             size_t nSource, const double *coordSource, const NPME_KfuncReal *func, bool printLog,
             std::ostream &os)
  {
    reset();
    const NPME_BuiltinKernel *d = func ? func->npme_builtin() : nullptr;
    int rc;
    if (!func || (d && d->part == NPME_BuiltinKernel::FULL)) {
      NPME_KeywordInput k = keyword;
      if (d) {
        k.AddKeyword("funcType", d->family == NPME_BuiltinKernel::RALPHA ? "Ralpha" : "Laplace");
        if (d->family == NPME_BuiltinKernel::RALPHA) k.AddKeyword("alpha", d->alpha);
      }
      rc = npme_create_boxpair(&_h, k.c_handle(), nTarget, coordTarget, nSource, coordSource);
    } else {
      _cr = {const_cast<NPME_KfuncReal *>(func), npme_detail::tramp_value_r, npme_detail::tramp_grad_r};
      rc = npme_create_boxpair_genfunc_real(&_h, keyword.c_handle(), nTarget, coordTarget, nSource,
                                            coordSource, &_cr);
    }
    return done(rc, nTarget, nSource, false, printLog, os);
  }
  bool SetUp(const NPME_KeywordInput &keyword, size_t nTarget, const double *coordTarget,
             size_t nSource, const double *coordSource, const NPME_KfuncComplex *func,
             bool printLog, std::ostream &os)
  {
    reset();
    const NPME_BuiltinKernel *d = func ? func->npme_builtin() : nullptr;
    int rc;
    if (d && d->part == NPME_BuiltinKernel::FULL) {
      NPME_KeywordInput k = keyword;
      k.AddKeyword("funcType", "Helmholtz");
      k.AddKeyword("k0_r", d->k0_r);
      k.AddKeyword("k0_i", d->k0_i);
      rc = npme_create_boxpair(&_h, k.c_handle(), nTarget, coordTarget, nSource, coordSource);
    } else if (!func) {
      rc = npme_create_boxpair(&_h, keyword.c_handle(), nTarget, coordTarget, nSource, coordSource);
    } else {
      _cc = {const_cast<NPME_KfuncComplex *>(func), npme_detail::tramp_value_c, npme_detail::tramp_grad_c};
      rc = npme_create_boxpair_genfunc_complex(&_h, keyword.c_handle(), nTarget, coordTarget, nSource,
                                               coordSource, &_cc);
    }
    return done(rc, nTarget, nSource, true, printLog, os);
  }
  // V[nTarget][4]; reverseSign (nullptr: none) mirrors the source box through
  // the target box centre along the flagged axes, with the same grid
  bool CalcV(double *V, const double *chargeSource, const bool *reverseSign, bool printLog,
             std::ostream &os)
  {
    int rv[3];
    return npme_report(npme_boxpair_calcv_real(_h, V, _nT, _nS, chargeSource, flags(reverseSign, rv)),
                       printLog, os);
  }
  bool CalcV(_Complex double *V, const _Complex double *chargeSource, const bool *reverseSign,
             bool printLog, std::ostream &os)
  {
    int rv[3];
    return npme_report(npme_boxpair_calcv_complex(_h, reinterpret_cast<double *>(V), _nT, _nS,
                                                  reinterpret_cast<const double *>(chargeSource),
                                                  flags(reverseSign, rv)),
                       printLog, os);
  }
  // R0 = target centre - source centre (f(r) = func(|r + R0|)) and the
  // distance of the grid box from the singularity
  bool Separation(double R0[3], double &rMin) const { return npme_boxpair_separation(_h, R0, &rMin) == NPME_OK; }

 private:
  void reset()
  {
    npme_destroy(_h);
    _h = nullptr;
    _nT = _nS = 0;
  }
  bool done(int rc, size_t nT, size_t nS, bool cx, bool printLog, std::ostream &os)
  {
    (void)cx;
    if (rc != NPME_OK) {
      _h = nullptr;
      return npme_report(rc, printLog, os);
    }
    _nT = nT;
    _nS = nS;
    return true;
  }
  static const int *flags(const bool *r, int *o)
  {
    if (!r) return nullptr;
    for (int d = 0; d < 3; d++) o[d] = r[d] ? 1 : 0;
    return o;
  }
  npme_handle *_h = nullptr;
  size_t _nT = 0, _nS = 0;
  npme_kfunc_real _cr = {};
  npme_kfunc_complex _cc = {};
};

}  // namespace NPME_Library


#endif
