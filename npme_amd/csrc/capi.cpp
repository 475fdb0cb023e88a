// C ABI of the MI355X NPME library (include/npme_amd.h).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <thread>

#include "npme_internal.h"

namespace npme {
static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
std::string last_error() { return g_err; }
}  // namespace npme

using namespace npme;

extern "C" {

const char *npme_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------- keywords
npme_keywords *npme_keywords_create(void) { return new (std::nothrow) npme_keywords(); }
void npme_keywords_destroy(npme_keywords *kw) { delete kw; }

int npme_keywords_set(npme_keywords *kw, const char *key, const char *value)
{
  if (!kw || !key || !value) {
    set_error("npme_keywords_set: NULL argument");
    return NPME_ERR_ARG;
  }
  kw->kw.set(key, std::string(value));
  return NPME_OK;
}

int npme_keywords_get(const npme_keywords *kw, const char *key, char *buf, size_t buflen)
{
  if (!kw || !key) return 0;
  std::string v;
  if (!kw->kw.get(key, v)) return 0;
  if (buf && buflen) {
    std::snprintf(buf, buflen, "%s", v.c_str());
  }
  return 1;
}

int npme_keywords_count(const npme_keywords *kw) { return kw ? (int)kw->kw.kv.size() : 0; }

int npme_keywords_key(const npme_keywords *kw, int i, char *buf, size_t buflen)
{
  if (!kw || i < 0 || i >= (int)kw->kw.kv.size()) return NPME_ERR_ARG;
  std::snprintf(buf, buflen, "%s", kw->kw.kv[i].first.c_str());
  return NPME_OK;
}

// "@Keyword" then "key value [comment...]" lines until the next '@' or EOF
// (src/ReadPrint.cpp:38-66)
int npme_keywords_read_file(npme_keywords *kw, const char *path)
{
  return guarded([&] {
    if (!kw || !path) throw Error{NPME_ERR_ARG, "npme_keywords_read_file: NULL argument"};
    std::ifstream in(path);
    if (!in) throw Error{NPME_ERR_ARG, std::string("cannot open ") + path};
    std::string line;
    bool inside = false, found = false;
    while (std::getline(in, line)) {
      std::istringstream ls(line);
      std::string a, b;
      if (!(ls >> a)) continue;
      if (a[0] == '@') {
        inside = (a == "@Keyword");
        found |= inside;
        continue;
      }
      if (inside && (ls >> b)) kw->kw.set(a, b);
    }
    if (!found) throw Error{NPME_ERR_KEYWORD, std::string("no @Keyword section in ") + path};
  });
}

// ---------------------------------------------------------------- defaults + RSEM
// NPME_AddMissingKeywords (src/Interface.cpp:53-517)
int npme_add_missing_keywords(npme_keywords *kwp, size_t nCharge, int isChargeReal,
                              const double *coord, const double *chargeReal,
                              const double *chargeComplex)
{
  return guarded([&] {
    if (!kwp || !coord) throw Error{NPME_ERR_ARG, "npme_add_missing_keywords: NULL argument"};
    add_missing_keywords(kwp->kw, nCharge, isChargeReal != 0, coord, chargeReal, chargeComplex);
  });
}

}  // extern "C"

namespace npme {

void add_missing_keywords(Keywords &kw, size_t nCharge, bool isChargeReal, const double *coord,
                          const double *chargeReal, const double *chargeComplex,
                          const ChargeStats *stats)
{
  {
    std::string s, funcType, calcType, split;
    double Rdir, tol, k0r = 0, k0i = 0;
    long Bn;
    if (!kw.get("nProc", s)) kw.set("nProc", (long)std::max(1u, std::thread::hardware_concurrency()));
    if (!kw.get("FFTmemGB", s)) kw.set("FFTmemGB", 10.0);
    if (!kw.get("vecOption", s)) kw.set("vecOption", std::string("avx"));
    if (!kw.get("BnOrder", Bn)) {
      Bn = 8;
      kw.set("BnOrder", Bn);
    }
    for (const char *k : {"a1", "a2", "a3"})
      if (!kw.get(k, s)) kw.set(k, 4.0);
    if (!kw.get("tol", tol)) {
      tol = 1.0e-6;
      kw.set("tol", tol);
    }
    if (!kw.get("nCellClust1D", s)) kw.set("nCellClust1D", 3L);
    if (!kw.get("calcType", calcType))
      throw Error{NPME_ERR_KEYWORD, "missing required keyword 'calcType' ('pme', 'exact' or 'pme_exact')"};
    if (calcType != "pme" && calcType != "exact" && calcType != "pme_exact")
      throw Error{NPME_ERR_KEYWORD, "'calcType' = " + calcType + " is undefined"};
    if (!kw.get("funcType", funcType))
      throw Error{NPME_ERR_KEYWORD, "missing required keyword 'funcType' ('Laplace', 'Helmholtz' or 'Ralpha')"};
    if (funcType == "Helmholtz") {
      if (!kw.get("k0_r", k0r)) throw Error{NPME_ERR_KEYWORD, "missing required keyword 'k0_r'"};
      if (!kw.get("k0_i", k0i)) throw Error{NPME_ERR_KEYWORD, "missing required keyword 'k0_i'"};
      if (k0i < -1.0e-12) throw Error{NPME_ERR_KEYWORD, "k0_i is negative"};
      if (isChargeReal) throw Error{NPME_ERR_KEYWORD, "input charges are real, but kernel is complex"};
    } else if (funcType == "Ralpha") {
      if (!kw.get("alpha", s)) throw Error{NPME_ERR_KEYWORD, "missing required keyword 'alpha'"};
      if (!isChargeReal) throw Error{NPME_ERR_KEYWORD, "input charges are complex, but kernel is real"};
    } else if (funcType == "Laplace") {
      if (!isChargeReal) throw Error{NPME_ERR_KEYWORD, "input charges are complex, but kernel is real"};
    } else {
      throw Error{NPME_ERR_KEYWORD, "funcType = " + funcType + " is undefined"};
    }
    if (!kw.get("Rdir", Rdir)) throw Error{NPME_ERR_KEYWORD, "missing required keyword 'Rdir'"};
    if (!kw.get("EwaldSplit", split)) {
      split = "DerivMatch";
      kw.set("EwaldSplit", split);
    } else if (split == "LaplaceOrig") {
      if (funcType != "Laplace")
        throw Error{NPME_ERR_KEYWORD, "keywords for EwaldSplit and funcType are incompatible"};
    } else if (split != "DerivMatch") {
      throw Error{NPME_ERR_KEYWORD, "EwaldSplit = " + split + " is undefined"};
    }
    for (const char *k : {"del1", "del2", "del3"})
      if (!kw.get(k, s)) kw.set(k, Rdir);
    if (!kw.get("printV", s)) kw.set("printV", 1L);
    if (!kw.get("printLog", s)) kw.set("printLog", 1L);
    if (!kw.get("nNeigh", s)) kw.set("nNeigh", 2L);
    if (!kw.get("useCompactF", s)) kw.set("useCompactF", 1L);

    bool setFFT = false;
    for (const char *k : {"N1", "N2", "N3", "n1", "n2", "n3"})
      if (!kw.get(k, s)) setFFT = true;
    if (setFFT) {
      long N[3], nb[3];
      int nDeriv = -1, nd = 0;
      const bool haveNder = kw.get("nDeriv", nDeriv);
      const bool cx = funcType == "Helmholtz";
      if (cx && !chargeComplex) throw Error{NPME_ERR_ARG, "complex charges required for the grid choice"};
      if (!cx && !chargeReal) throw Error{NPME_ERR_ARG, "real charges required for the grid choice"};
      ChargeStats own;
      if (!stats) charge_stats(nCharge, coord, cx ? chargeComplex : chargeReal, cx, own);
      const ChargeStats &st = stats ? *stats : own;
      if (cx) {
        std::string err;
        if (!rsem_helmholtz(st, Bn, k0r, k0i, tol, Rdir, haveNder ? nDeriv : -1, N, nb, nd, err))
          throw Error{NPME_ERR_KEYWORD, err};
      } else {
        double alpha = -1.0;
        if (funcType == "Ralpha") kw.get("alpha", alpha);
        if (!rsem_static(st, Bn, alpha, tol, Rdir, haveNder ? nDeriv : -1, N, nb, nd)) {
          std::ostringstream os;
          os << "grid estimate: no error-model parameters for BnOrder = " << Bn
             << " alpha = " << alpha << (haveNder ? " nDeriv = " + std::to_string(nDeriv) : "");
          throw Error{NPME_ERR_KEYWORD, os.str()};
        }
      }
      if (!haveNder) kw.set("nDeriv", (long)nd);
      const char *nk[3] = {"N1", "N2", "N3"}, *bk[3] = {"n1", "n2", "n3"};
      for (int d = 0; d < 3; d++) {
        kw.set(nk[d], N[d]);
        kw.set(bk[d], nb[d]);
      }
    }
    if (split == "DerivMatch") {
      int nDeriv;
      if (!kw.get("nDeriv", nDeriv)) {
        nDeriv = (Bn == 4) ? 3 : (Bn == 6) ? 6 : 8;
        kw.set("nDeriv", (long)nDeriv);
      }
    }
  }
}

}  // namespace npme

extern "C" {

// ---------------------------------------------------------------- handles
int npme_create(npme_handle **out, const npme_keywords *kw, size_t nCharge, const double *coord)
{
  return guarded([&] {
    if (!out || !kw || !coord) throw Error{NPME_ERR_ARG, "npme_create: NULL argument"};
    *out = nullptr;
    std::unique_ptr<npme_handle> h(new npme_handle());
    try {
      handle_setup(h->h, kw->kw, nCharge, coord);
    } catch (...) {
      handle_free_device(h->h);
      handle_destroy_streams(h->h);
      throw;
    }
    *out = h.release();
  });
}

static int create_user(npme_handle **out, const npme_keywords *kw, size_t nCharge,
                       const double *coord, const UserKernels &u)
{
  return guarded([&] {
    if (!out || !kw || !coord) throw Error{NPME_ERR_ARG, "npme_create_genfunc: NULL argument"};
    *out = nullptr;
    std::unique_ptr<npme_handle> h(new npme_handle());
    h->h.userK = true;
    h->h.user = u;
    try {
      handle_setup(h->h, kw->kw, nCharge, coord);
    } catch (...) {
      handle_free_device(h->h);
      handle_destroy_streams(h->h);
      throw;
    }
    *out = h.release();
  });
}

int npme_create_genfunc_real(npme_handle **out, const npme_keywords *kw, size_t nCharge,
                             const double *coord, const npme_kfunc_real *func,
                             const npme_kfunc_real *funcLR, const npme_kfunc_real *funcSR)
{
  const npme_kfunc_real *f[3] = {func, funcLR, funcSR};
  UserKernels u;
  u.complex = false;
  for (int i = 0; i < 3; i++) {
    // src/Interface.cpp:1620-1638: each of func, funcLR, funcSR is required
    if (!f[i] || !f[i]->value || !f[i]->grad) {
      set_error(std::string("npme_create_genfunc_real: ") + (i == 0 ? "func" : i == 1 ? "funcLR" : "funcSR") +
                " == NULL");
      if (out) *out = nullptr;
      return NPME_ERR_ARG;
    }
    u.r[i] = *f[i];
  }
  return create_user(out, kw, nCharge, coord, u);
}

int npme_create_genfunc_complex(npme_handle **out, const npme_keywords *kw, size_t nCharge,
                                const double *coord, const npme_kfunc_complex *func,
                                const npme_kfunc_complex *funcLR,
                                const npme_kfunc_complex *funcSR)
{
  const npme_kfunc_complex *f[3] = {func, funcLR, funcSR};
  UserKernels u;
  u.complex = true;
  for (int i = 0; i < 3; i++) {
    if (!f[i] || !f[i]->value || !f[i]->grad) {
      set_error(std::string("npme_create_genfunc_complex: ") +
                (i == 0 ? "func" : i == 1 ? "funcLR" : "funcSR") + " == NULL");
      if (out) *out = nullptr;
      return NPME_ERR_ARG;
    }
    u.c[i] = *f[i];
  }
  return create_user(out, kw, nCharge, coord, u);
}

int npme_dm_coefficients(const char *funcType, double alpha, double k0_r, double k0_i, int Nder,
                         double Rdir, double *a, double *b)
{
  return guarded([&] {
    if (!funcType || !a || !b) throw Error{NPME_ERR_ARG, "npme_dm_coefficients: NULL argument"};
    const std::string ft(funcType);
    int kind;
    if (ft == "Laplace") kind = KIND_LAPLACE;
    else if (ft == "Ralpha") kind = KIND_RALPHA;
    else if (ft == "Helmholtz") kind = KIND_HELMHOLTZ;
    else throw Error{NPME_ERR_KEYWORD, "funcType = " + ft + " is undefined"};
    DMKernel k;
    std::string err;
    if (!dm_kernel_init(k, kind, alpha, k0_r, k0_i, Nder, Rdir, err)) throw Error{NPME_ERR_KEYWORD, err};
    for (int i = 0; i < 2 * (Nder + 1); i++) {
      a[i] = k.a[i];
      b[i] = k.b[i];
    }
  });
}

int npme_ewald_rdir2beta(double Rdir, double tol, double *beta)
{
  return guarded([&] {
    if (!beta) throw Error{NPME_ERR_ARG, "npme_ewald_rdir2beta: NULL argument"};
    if (!(Rdir > 0) || !(tol > 0) || !(tol < 1)) throw Error{NPME_ERR_ARG, "needs Rdir > 0, 0 < tol < 1"};
    *beta = rdir2beta(Rdir, tol);
  });
}

int npme_get_table_error(const npme_handle *h, double err[2])
{
  if (!h || !err) {
    set_error("npme_get_table_error: NULL argument");
    return NPME_ERR_ARG;
  }
  err[0] = h->h.utErr[0];
  err[1] = h->h.utErr[1];
  return NPME_OK;
}

int npme_reset_coord(npme_handle *h, size_t nCharge, const double *coord)
{
  return guarded([&] {
    if (!h || !coord) throw Error{NPME_ERR_ARG, "npme_reset_coord: NULL argument"};
    if (h->h.trans) throw Error{NPME_ERR_ARG, "a box-pair handle is reset with npme_boxpair_reset_coord"};
    std::lock_guard<std::mutex> lk(h->h.mtx);
    Keywords kw = h->h.kw;
    handle_setup(h->h, kw, nCharge, coord);
  });
}

void npme_destroy(npme_handle *h)
{
  if (!h) return;
  handle_free_device(h->h);
  handle_destroy_streams(h->h);
  delete h;
}

static void check_call(npme_handle *h, size_t nCharge, bool wantComplex)
{
  if (!h) throw Error{NPME_ERR_ARG, "NULL handle"};
  if (h->h.trans) throw Error{NPME_ERR_ARG, "a box-pair handle is evaluated with npme_boxpair_calcv_*"};
  if (!h->h.ready) throw Error{NPME_ERR_ARG, "no valid set-up: the last SetUp or ResetCoord failed"};
  if (nCharge != h->h.n) {
    std::ostringstream os;
    os << "nCharge = " << nCharge << " != " << h->h.n;
    throw Error{NPME_ERR_ARG, os.str()};
  }
  if (wantComplex != h->h.complexK)
    throw Error{NPME_ERR_ARG, wantComplex ? "complex call on a real kernel" : "real call on a complex kernel"};
}

static void calcv_host(npme_handle *h, double *V, size_t n, const double *charge, bool cx)
{
  check_call(h, n, cx);
  if (!V || !charge) throw Error{NPME_ERR_ARG, "NULL V or charge"};
  std::lock_guard<std::mutex> lk(h->h.mtx);
  Handle &H = h->h;
  const size_t qn = (cx ? 2 : 1) * n, vn = (cx ? 8 : 4) * n;
  NPME_HIP_CHECK(hipMemcpyAsync(H.d_qin, charge, qn * sizeof(double), hipMemcpyHostToDevice, H.stream));
  // CalcV reads the charges from both of its streams
  NPME_HIP_CHECK(hipStreamSynchronize(H.stream));
  handle_calcv(H, H.d_qin, H.d_V);
  NPME_HIP_CHECK(hipMemcpyAsync(V, H.d_V, vn * sizeof(double), hipMemcpyDeviceToHost, H.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(H.stream));
}

int npme_calcv_real(npme_handle *h, double *V, size_t nCharge, const double *charge,
                    const double *coord)
{
  (void)coord;
  return guarded([&] { calcv_host(h, V, nCharge, charge, false); });
}

int npme_calcv_complex(npme_handle *h, double *V, size_t nCharge, const double *charge,
                       const double *coord)
{
  (void)coord;
  return guarded([&] { calcv_host(h, V, nCharge, charge, true); });
}

int npme_calcv_device(npme_handle *h, double *dV, const double *dCharge)
{
  return guarded([&] {
    if (!h || !dV || !dCharge) throw Error{NPME_ERR_ARG, "npme_calcv_device: NULL argument"};
    if (h->h.trans) throw Error{NPME_ERR_ARG, "a box-pair handle is evaluated with npme_boxpair_calcv_*"};
    std::lock_guard<std::mutex> lk(h->h.mtx);
    handle_calcv(h->h, dCharge, dV);
  });
}

static void exact_host(npme_handle *h, double *V, size_t n, const double *charge, bool cx)
{
  check_call(h, n, cx);
  if (!V || !charge) throw Error{NPME_ERR_ARG, "NULL V or charge"};
  std::lock_guard<std::mutex> lk(h->h.mtx);
  Handle &H = h->h;
  if (H.userK) {  // the user's full kernel, through its callback on the host
    user_exact_host(H.user, n, H.coordHost.data(), charge, n, nullptr, V);
    return;
  }
  const size_t qn = (cx ? 2 : 1) * n, vn = (cx ? 8 : 4) * n;
  NPME_HIP_CHECK(hipMemcpyAsync(H.d_qin, charge, qn * sizeof(double), hipMemcpyHostToDevice, H.stream));
  handle_exact(H, H.d_qin, n, nullptr, H.d_V);
  NPME_HIP_CHECK(hipMemcpyAsync(V, H.d_V, vn * sizeof(double), hipMemcpyDeviceToHost, H.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(H.stream));
}

int npme_calcv_exact_real(npme_handle *h, double *V, size_t nCharge, const double *charge,
                          const double *coord)
{
  (void)coord;
  return guarded([&] { exact_host(h, V, nCharge, charge, false); });
}

int npme_calcv_exact_complex(npme_handle *h, double *V, size_t nCharge, const double *charge,
                             const double *coord)
{
  (void)coord;
  return guarded([&] { exact_host(h, V, nCharge, charge, true); });
}

int npme_calcv_exact_subset(npme_handle *h, double *V, size_t nTarget, const long *idx,
                            const double *charge)
{
  return guarded([&] {
    if (!h || !V || !idx || !charge) throw Error{NPME_ERR_ARG, "npme_calcv_exact_subset: NULL argument"};
    std::lock_guard<std::mutex> lk(h->h.mtx);
    Handle &H = h->h;
    if (!H.ready) throw Error{NPME_ERR_ARG, "no valid set-up: the last SetUp or ResetCoord failed"};
    for (size_t t = 0; t < nTarget; t++)
      if (idx[t] < 0 || (size_t)idx[t] >= H.n) throw Error{NPME_ERR_ARG, "target index out of range"};
    if (H.userK) {
      user_exact_host(H.user, H.n, H.coordHost.data(), charge, nTarget, idx, V);
      return;
    }
    const bool cx = H.complexK;
    const size_t qn = (cx ? 2 : 1) * H.n, vw = cx ? 8 : 4;
    DevBuf d_idx(std::max<size_t>(nTarget, 1) * sizeof(long));
    DevBuf d_out(std::max<size_t>(nTarget, 1) * vw * sizeof(double));
    NPME_HIP_CHECK(hipMemcpyAsync(H.d_qin, charge, qn * sizeof(double), hipMemcpyHostToDevice, H.stream));
    NPME_HIP_CHECK(hipMemcpyAsync(d_idx.p, idx, nTarget * sizeof(long), hipMemcpyHostToDevice, H.stream));
    handle_exact(H, H.d_qin, nTarget, (const long *)d_idx.p, (double *)d_out.p);
    NPME_HIP_CHECK(hipMemcpyAsync(V, d_out.p, nTarget * vw * sizeof(double), hipMemcpyDeviceToHost,
                                  H.stream));
    NPME_HIP_CHECK(hipStreamSynchronize(H.stream));
  });
}

int npme_get_params(const npme_handle *h, long N[3], long nblk[3], long *Bn, int *Nder,
                    double *Rdir, double *deviceGB)
{
  if (!h) {
    set_error("NULL handle");
    return NPME_ERR_ARG;
  }
  const Handle &H = h->h;
  for (int d = 0; d < 3; d++) {
    if (N) N[d] = H.N[d];
    if (nblk) nblk[d] = H.nblk[d];
  }
  if (Bn) *Bn = H.Bn;
  if (Nder) *Nder = H.ker.Nder;
  if (Rdir) *Rdir = H.ker.Rdir;
  if (deviceGB) *deviceGB = H.deviceBytes / 1.0e9;
  return NPME_OK;
}

int npme_print(const npme_handle *h, char *buf, size_t buflen)
{
  if (!h || !buf) return NPME_ERR_ARG;
  const Handle &H = h->h;
  std::ostringstream os;
  static const char *kn[6] = {"Laplace", "Ralpha", "Helmholtz", "Laplace", "user-defined real",
                               "user-defined complex"};
  os << "NPME (MI355X) kernel " << kn[H.ker.kind]
     << (H.ker.kind == KIND_LAPLACE_ERF ? " LaplaceOrig"
                                        : kind_is_user(H.ker.kind) ? " GenFunc" : " DerivMatch")
     << ", precision " << (H.fp64 ? "fp64" : "mixed") << "\n";
  os << "  nCharge " << H.n << "  Rdir " << H.ker.Rdir << "  Nder " << H.ker.Nder << "  BnOrder "
     << H.Bn << "\n";
  os << "  N1 N2 N3 " << H.N[0] << " " << H.N[1] << " " << H.N[2] << "  n1 n2 n3 " << H.nblk[0]
     << " " << H.nblk[1] << " " << H.nblk[2] << "\n";
  os << "  X Y Z " << H.X[0] << " " << H.X[1] << " " << H.X[2] << "  L " << H.L[0] << " "
     << H.L[1] << " " << H.L[2] << "\n";
  os << "  fself " << H.fself[0] << " " << H.fself[1] << "\n";
  if (H.ker.kind == KIND_LAPLACE_ERF) os << "  beta " << H.ker.beta << "\n";
  if (kind_is_user(H.ker.kind))
    os << "  short-range table: " << kUserOct << " octaves x " << H.utS << " pieces of degree "
       << kUserDeg << ", error " << H.utErr[0] << " (f) " << H.utErr[1] << " (f'/r)\n";
  os << "  cells " << H.nb[0] << "x" << H.nb[1] << "x" << H.nb[2] << " of " << H.csB
     << "  device GB " << H.deviceBytes / 1e9 << "\n";
  os << "  per-charge set-up on the " << (H.deviceSetup ? "device" : "host") << "\n";
  os << "  set-up ms: geometry " << H.setupMs[0] << ", free " << H.setupMs[1] << ", cells "
     << H.setupMs[2] << ", list " << H.setupMs[3] << ", charges " << H.setupMs[4] << ", grid "
     << H.setupMs[5] << "\n";
  os << "  kernel grid G: " << (H.gReused ? "reused from the previous set-up" : "computed") << "\n";
  if (H.trans)
    os << "  direct sum: none (box pair on the translated kernel grid)\n";
  else if (kind_ref_pairs(H.ker.kind))
    os << "  direct sum: the reference's interacting cell pairs, every point pair (FP64)\n";
  else if (H.useNlist)
    os << "  direct sum: neighbour list, " << H.nlistEntries << " entries\n";
  else
    os << "  direct sum: cell rows, " << H.nSegments << " segments\n";
  std::snprintf(buf, buflen, "%s", os.str().c_str());
  return NPME_OK;
}

int npme_count_pairs(npme_handle *h, double *pairs)
{
  return guarded([&] {
    if (!h || !pairs) throw Error{NPME_ERR_ARG, "npme_count_pairs: NULL argument"};
    std::lock_guard<std::mutex> lk(h->h.mtx);
    *pairs = handle_count_pairs(h->h);
  });
}

int npme_count_pairs_noalloc(npme_handle *h, double *pairs)
{
  return guarded([&] {
    if (!h || !pairs) throw Error{NPME_ERR_ARG, "npme_count_pairs_noalloc: NULL argument"};
    std::lock_guard<std::mutex> lk(h->h.mtx);
    *pairs = handle_count_pairs_noalloc(h->h);
  });
}

int npme_get_indices(npme_handle *h, size_t nCharge, unsigned long long *clusterKey,
                     unsigned long long *cellKey, long *mStart, long *order)
{
  return guarded([&] {
    if (!h) throw Error{NPME_ERR_ARG, "NULL handle"};
    std::lock_guard<std::mutex> lk(h->h.mtx);
    const Handle &H = h->h;
    if (!H.ready) throw Error{NPME_ERR_ARG, "no valid set-up: the last SetUp or ResetCoord failed"};
    if (nCharge != H.n) throw Error{NPME_ERR_ARG, "nCharge does not match the set-up"};
    const size_t n = H.n;
    std::vector<int> perm(n);
    std::vector<int4> cell(n), ms(n);
    NPME_HIP_CHECK(hipMemcpyAsync(perm.data(), H.d_perm, n * sizeof(int), hipMemcpyDeviceToHost, H.stream));
    NPME_HIP_CHECK(hipMemcpyAsync(cell.data(), H.d_cell, n * sizeof(int4), hipMemcpyDeviceToHost, H.stream));
    NPME_HIP_CHECK(hipMemcpyAsync(ms.data(), H.d_ms, n * sizeof(int4), hipMemcpyDeviceToHost, H.stream));
    NPME_HIP_CHECK(hipStreamSynchronize(H.stream));
    // the keys of src/PartitionEmbeddedBox.cpp:266-455 from the global cell
    // coordinates g = a nP + b kept by the set-up (a: cluster, b: cell in
    // the cluster)
    const unsigned long nP = (unsigned long)H.nCellClust1D, Nc = nP * nP * nP;
    const unsigned long nA[3] = {H.nb[0] / nP, H.nb[1] / nP, H.nb[2] / nP};
    for (size_t s = 0; s < n; s++) {
      const size_t i = (size_t)perm[s];
      if (i >= n) throw Error{NPME_ERR_DEVICE, "set-up permutation out of range"};
      const unsigned long g[3] = {(unsigned long)cell[s].x, (unsigned long)cell[s].y,
                                  (unsigned long)cell[s].z};
      const unsigned long tA = ((g[0] / nP) * nA[1] + g[1] / nP) * nA[2] + g[2] / nP;
      if (clusterKey) clusterKey[i] = tA;
      if (cellKey) cellKey[i] = tA * Nc + ((g[0] % nP) * nP + g[1] % nP) * nP + g[2] % nP;
      if (mStart) {
        mStart[3 * i + 0] = ms[s].x - H.N[0] / 4;
        mStart[3 * i + 1] = ms[s].y - H.N[1] / 4;
        mStart[3 * i + 2] = ms[s].z - H.N[2] / 4;
      }
      if (order) order[s] = (long)i;
    }
  });
}

void npme_debug_fail_alloc(long k) { dev_fail_after(k); }

int npme_set_calcv_graph(npme_handle *h, int on)
{
  return guarded([&] {
    if (!h) throw Error{NPME_ERR_ARG, "NULL handle"};
    std::lock_guard<std::mutex> lk(h->h.mtx);
    h->h.useGraph = on != 0;
  });
}

int npme_set_profiling(npme_handle *h, int on)
{
  if (!h) return NPME_ERR_ARG;
  h->h.profiling = on != 0;
  return NPME_OK;
}

int npme_get_stage_times(const npme_handle *h, double *ms, int n)
{
  if (!h || !ms) return 0;
  const int k = std::min(n, h->h.nstage);
  for (int i = 0; i < k; i++) ms[i] = h->h.stage_ms[i];
  return k;
}

}  // extern "C"
