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
 * split) are evaluated on the GPU.  The NPME_Kfunc* objects of
 * NPME_KernelList are descriptors of those kernels: a user-defined
 * NPME_KfuncReal / NPME_KfuncComplex subclass is not evaluated, and SetUp
 * returns false for it.  Failures return false and print the library's
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

// ------------------------------------------------------------ kernel descriptors
// NPME_KfuncReal / NPME_KfuncComplex (src/KernelFunction.h:25-116) and the
// built-in kernels of NPME_KernelList (src/Interface.h:757-771), with the
// constructors and SetParm of src/KernelFunction{Laplace,Ralpha,Helmholtz}.h.
// A descriptor records its parameters; SetUp passes those that were set
// (alpha, k0, nDeriv, Rdir) to the library in place of the keywords'.
class NPME_KfuncReal {
 public:
  virtual ~NPME_KfuncReal() {}
  virtual bool builtin() const { return false; }
  // parameters carried by a built-in descriptor: alpha, Nder, Rdir
  bool hasAlpha = false, hasDM = false;
  double alpha = 0, Rdir = 0;
  int Nder = 0;
};
class NPME_KfuncComplex {
 public:
  virtual ~NPME_KfuncComplex() {}
  virtual bool builtin() const { return false; }
  bool hasK0 = false, hasDM = false;
  double k0_r = 0, k0_i = 0, Rdir = 0;
  int Nder = 0;
};

// exact kernels: NPME_Kfunc_Laplace (src/KernelFunctionLaplace.h:38),
// NPME_Kfunc_Ralpha(alpha) (src/KernelFunctionRalpha.h:34-37),
// NPME_Kfunc_Helmholtz(k0) (src/KernelFunctionHelmholtz.h:34-37)
class NPME_Kfunc_Laplace : public NPME_KfuncReal {
 public:
  bool builtin() const override { return true; }
};
class NPME_Kfunc_Ralpha : public NPME_KfuncReal {
 public:
  NPME_Kfunc_Ralpha() {}
  explicit NPME_Kfunc_Ralpha(double a) { SetParm(a); }
  bool builtin() const override { return true; }
  bool SetParm(const double a)
  {
    alpha = a;
    hasAlpha = true;
    return true;
  }
};
class NPME_Kfunc_Helmholtz : public NPME_KfuncComplex {
 public:
  NPME_Kfunc_Helmholtz() {}
  explicit NPME_Kfunc_Helmholtz(_Complex double k0) { SetParm(k0); }
  bool builtin() const override { return true; }
  bool SetParm(const _Complex double k0)
  {
    k0_r = __real__ k0;
    k0_i = __imag__ k0;
    hasK0 = true;
    return true;
  }
};

// DM split kernels: (Nder, Rdir) for Laplace (src/KernelFunctionLaplace.h:
// 87-116,155-183), (alpha, Nder, Rdir) for r^alpha
// (src/KernelFunctionRalpha.h:94-126,169-200), (k0, Nder, Rdir) for
// Helmholtz (src/KernelFunctionHelmholtz.h:102-133,182-212)
#define NPME_AMD_LAPLACE_DM(name)                                                     \
  class name : public NPME_KfuncReal {                                                \
   public:                                                                            \
    name() {}                                                                         \
    name(int nd, double rd) { SetParm(nd, rd, false, std::cout); }                    \
    bool builtin() const override { return true; }                                    \
    bool SetParm(const int nd, const double rd, bool, std::ostream &)                 \
    {                                                                                 \
      Nder = nd;                                                                      \
      Rdir = rd;                                                                      \
      hasDM = true;                                                                   \
      return true;                                                                    \
    }                                                                                 \
  };
#define NPME_AMD_RALPHA_DM(name)                                                      \
  class name : public NPME_KfuncReal {                                                \
   public:                                                                            \
    name() {}                                                                         \
    name(double a, int nd, double rd) { SetParm(a, nd, rd, false, std::cout); }       \
    bool builtin() const override { return true; }                                    \
    bool SetParm(const double a, const int nd, const double rd, bool, std::ostream &) \
    {                                                                                 \
      alpha = a;                                                                      \
      Nder = nd;                                                                      \
      Rdir = rd;                                                                      \
      hasAlpha = hasDM = true;                                                        \
      return true;                                                                    \
    }                                                                                 \
  };
#define NPME_AMD_HELMHOLTZ_DM(name)                                                   \
  class name : public NPME_KfuncComplex {                                             \
   public:                                                                            \
    name() {}                                                                         \
    name(_Complex double k0, int nd, double rd) { SetParm(k0, nd, rd, false, std::cout); } \
    bool builtin() const override { return true; }                                    \
    bool SetParm(const _Complex double k0, const int nd, const double rd, bool,       \
                 std::ostream &)                                                      \
    {                                                                                 \
      k0_r = __real__ k0;                                                             \
      k0_i = __imag__ k0;                                                             \
      Nder = nd;                                                                      \
      Rdir = rd;                                                                      \
      hasK0 = hasDM = true;                                                           \
      return true;                                                                    \
    }                                                                                 \
  };
NPME_AMD_LAPLACE_DM(NPME_Kfunc_Laplace_LR_DM)
NPME_AMD_LAPLACE_DM(NPME_Kfunc_Laplace_SR_DM)
NPME_AMD_RALPHA_DM(NPME_Kfunc_Ralpha_LR_DM)
NPME_AMD_RALPHA_DM(NPME_Kfunc_Ralpha_SR_DM)
NPME_AMD_HELMHOLTZ_DM(NPME_Kfunc_Helmholtz_LR_DM)
NPME_AMD_HELMHOLTZ_DM(NPME_Kfunc_Helmholtz_SR_DM)
#undef NPME_AMD_LAPLACE_DM
#undef NPME_AMD_RALPHA_DM
#undef NPME_AMD_HELMHOLTZ_DM

// the parameters a set of descriptors carries, written over the keywords
// (a later descriptor wins; the short-range one is applied last)
inline NPME_KeywordInput npme_apply_kernel_params(const NPME_KeywordInput &keyword,
                                                  const NPME_KfuncReal *const *f, int nf)
{
  NPME_KeywordInput k = keyword;
  for (int i = 0; i < nf; i++) {
    if (!f[i]) continue;
    if (f[i]->hasAlpha) k.AddKeyword("alpha", f[i]->alpha);
    if (f[i]->hasDM) {
      k.AddKeyword("nDeriv", f[i]->Nder);
      k.AddKeyword("Rdir", f[i]->Rdir);
    }
  }
  return k;
}
inline NPME_KeywordInput npme_apply_kernel_params(const NPME_KeywordInput &keyword,
                                                  const NPME_KfuncComplex *const *f, int nf)
{
  NPME_KeywordInput k = keyword;
  for (int i = 0; i < nf; i++) {
    if (!f[i]) continue;
    if (f[i]->hasK0) {
      k.AddKeyword("k0_r", f[i]->k0_r);
      k.AddKeyword("k0_i", f[i]->k0_i);
    }
    if (f[i]->hasDM) {
      k.AddKeyword("nDeriv", f[i]->Nder);
      k.AddKeyword("Rdir", f[i]->Rdir);
    }
  }
  return k;
}

struct NPME_KernelList {
  NPME_Kfunc_Laplace funcLaplace;
  NPME_Kfunc_Laplace_LR_DM funcLaplace_LR_DM;
  NPME_Kfunc_Laplace_SR_DM funcLaplace_SR_DM;
  NPME_Kfunc_Ralpha funcRalpha;
  NPME_Kfunc_Ralpha_LR_DM funcRalpha_LR_DM;
  NPME_Kfunc_Ralpha_SR_DM funcRalpha_SR_DM;
  NPME_Kfunc_Helmholtz funcHelmholtz;
  NPME_Kfunc_Helmholtz_LR_DM funcHelmholtz_LR_DM;
  NPME_Kfunc_Helmholtz_SR_DM funcHelmholtz_SR_DM;
};

// NPME_Interface_SelectKernelPtr (src/Interface.h:775-788): the kernel
// parameters (alpha, k0, Rdir, nDeriv) stay in the keywords
inline bool NPME_Interface_SelectKernelPtr(NPME_KfuncReal *&func, NPME_KfuncReal *&func_LR,
                                           NPME_KfuncReal *&func_SR, NPME_KernelList &kl,
                                           const NPME_KeywordInput &keyword, bool printLog,
                                           std::ostream &os)
{
  std::string ft;
  keyword.GetValue("funcType", ft);
  if (ft == "Laplace") {
    func = &kl.funcLaplace;
    func_LR = &kl.funcLaplace_LR_DM;
    func_SR = &kl.funcLaplace_SR_DM;
    return true;
  }
  if (ft == "Ralpha") {
    func = &kl.funcRalpha;
    func_LR = &kl.funcRalpha_LR_DM;
    func_SR = &kl.funcRalpha_SR_DM;
    return true;
  }
  if (printLog) os << "NPME error: funcType '" << ft << "' is not a real kernel\n";
  return false;
}
inline bool NPME_Interface_SelectKernelPtr(NPME_KfuncComplex *&func, NPME_KfuncComplex *&func_LR,
                                           NPME_KfuncComplex *&func_SR, NPME_KernelList &kl,
                                           const NPME_KeywordInput &keyword, bool printLog,
                                           std::ostream &os)
{
  std::string ft;
  keyword.GetValue("funcType", ft);
  if (ft == "Helmholtz") {
    func = &kl.funcHelmholtz;
    func_LR = &kl.funcHelmholtz_LR_DM;
    func_SR = &kl.funcHelmholtz_SR_DM;
    return true;
  }
  if (printLog) os << "NPME error: funcType '" << ft << "' is not a complex kernel\n";
  return false;
}

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

// NPME_InterfaceReal_GenFunc::SetUp (src/Interface.h:291-296): func, funcLR
// and funcSR must be the built-in descriptors of NPME_KernelList
class NPME_InterfaceReal_GenFunc : public NPME_InterfaceReal {
 public:
  bool SetUp(const NPME_KeywordInput &keyword, const size_t nCharge, const double *coord,
             NPME_KfuncReal *func, NPME_KfuncReal *funcLR, NPME_KfuncReal *funcSR,
             bool printLog, std::ostream &ofs_log)
  {
    if (!func || !funcLR || !funcSR || !func->builtin() || !funcLR->builtin() ||
        !funcSR->builtin()) {
      if (printLog) ofs_log << "NPME error: user-defined kernels are not provided on the GPU\n";
      return false;
    }
    const NPME_KfuncReal *f[3] = {func, funcLR, funcSR};
    return setup(npme_apply_kernel_params(keyword, f, 3), nCharge, coord, printLog, ofs_log);
  }
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

// NPME_Interface_SetUpRealKernel / SetUpComplexKernel
// (src/Interface.h:800-829): allocates with new; the caller deletes
inline bool NPME_Interface_SetUpRealKernel(NPME_InterfaceReal *&npme, NPME_KfuncReal *func,
                                           NPME_KfuncReal *funcLR, NPME_KfuncReal *funcSR,
                                           const size_t nCharge, const double *coord,
                                           const NPME_KeywordInput &keyword,
                                           bool useGenericKernel, bool printLog,
                                           std::ostream &ofs_log)
{
  npme = nullptr;
  std::string ft;
  keyword.GetValue("funcType", ft);
  if (ft == "Laplace" && !useGenericKernel) {
    auto *p = new NPME_InterfaceReal_Laplace_DM;
    if (!p->SetUp(keyword, nCharge, coord, printLog, ofs_log)) {
      delete p;
      return false;
    }
    npme = p;
    return true;
  }
  auto *p = new NPME_InterfaceReal_GenFunc;
  if (!p->SetUp(keyword, nCharge, coord, func, funcLR, funcSR, printLog, ofs_log)) {
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
  (void)useGenericKernel;
  npme = nullptr;
  if (!func || !funcLR || !funcSR || !func->builtin() || !funcLR->builtin() ||
      !funcSR->builtin()) {
    if (printLog) ofs_log << "NPME error: user-defined kernels are not provided on the GPU\n";
    return false;
  }
  auto *p = new NPME_InterfaceComplex_Helmholtz_DM;
  const NPME_KfuncComplex *f[3] = {func, funcLR, funcSR};
  if (!p->SetUp(npme_apply_kernel_params(keyword, f, 3), nCharge, coord, printLog, ofs_log)) {
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

}  // namespace NPME_Library


#endif
