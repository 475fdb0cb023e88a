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

#include <cstring>
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
  void PrintKeywords(std::ostream &os) const
  {
    const int n = npme_keywords_count(_kw);
    for (int i = 0; i < n; i++) {
      const std::string k = key(i);
      std::string v;
      GetValue(k, v);
      os << "  " << k << " " << v << "\n";
    }
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
// built-in kernels of NPME_KernelList (src/Interface.h:757-771)
class NPME_KfuncReal {
 public:
  virtual ~NPME_KfuncReal() {}
  virtual bool builtin() const { return false; }
};
class NPME_KfuncComplex {
 public:
  virtual ~NPME_KfuncComplex() {}
  virtual bool builtin() const { return false; }
};
#define NPME_AMD_BUILTIN_REAL(name)                        \
  class name : public NPME_KfuncReal {                     \
   public:                                                 \
    bool builtin() const override { return true; }         \
  };
#define NPME_AMD_BUILTIN_COMPLEX(name)                     \
  class name : public NPME_KfuncComplex {                  \
   public:                                                 \
    bool builtin() const override { return true; }         \
  };
NPME_AMD_BUILTIN_REAL(NPME_Kfunc_Laplace)
NPME_AMD_BUILTIN_REAL(NPME_Kfunc_Laplace_LR_DM)
NPME_AMD_BUILTIN_REAL(NPME_Kfunc_Laplace_SR_DM)
NPME_AMD_BUILTIN_REAL(NPME_Kfunc_Ralpha)
NPME_AMD_BUILTIN_REAL(NPME_Kfunc_Ralpha_LR_DM)
NPME_AMD_BUILTIN_REAL(NPME_Kfunc_Ralpha_SR_DM)
NPME_AMD_BUILTIN_COMPLEX(NPME_Kfunc_Helmholtz)
NPME_AMD_BUILTIN_COMPLEX(NPME_Kfunc_Helmholtz_LR_DM)
NPME_AMD_BUILTIN_COMPLEX(NPME_Kfunc_Helmholtz_SR_DM)
#undef NPME_AMD_BUILTIN_REAL
#undef NPME_AMD_BUILTIN_COMPLEX

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
    (void)PRINT;
    return npme_report(npme_calcv_real(_h, V1, nCharge, charge, coord), true, ofs_log);
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
    return setup(keyword, nCharge, coord, printLog, ofs_log);
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
    (void)PRINT;
    return npme_report(npme_calcv_complex(_h, reinterpret_cast<double *>(V1), nCharge,
                                          reinterpret_cast<const double *>(charge), coord),
                       true, ofs_log);
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
  if (!p->SetUp(keyword, nCharge, coord, printLog, ofs_log)) {
    delete p;
    return false;
  }
  npme = p;
  return true;
}

}  // namespace NPME_Library

#endif
