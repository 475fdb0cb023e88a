// Two-box (translated-kernel) grids.
//
// * A box pair evaluates V_t = sum_s q_s f(|r_t - r_s|) (and grad_t) for
//   targets t in one box and sources s in another, entirely on the kernel
//   grid: the full kernel f is smooth over the grid box when the boxes are
//   apart, so there is no direct sum and no self term.  Both sets go into one
//   handle, targets then sources, each relative to the centre of its own
//   bounding box, so the grid spans the larger of the two boxes, not the
//   pair; the kernel grid holds f(|r + D|), D = source centre - target
//   centre, over the full N1 x N2 x N3 grid (NPME_RecSumGrid_FullTrans_Func,
//   src/RecSumGrid.cpp:663-794), and the convolution pass reads G at
//   (k1, k2, k3) instead of (|k1|, |k2|, |k3|).  A CalcV may reverse D along
//   any axes with the same G (NPME_RSTP_CalcTheta_FullF_ReverseSign,
//   src/RecSumV.cpp:943-1009): G is read at -k on those axes.
// * npme_recsumgrid_fulltrans_func and npme_rstp_calctheta_fullF_reversesign
//   are device forms of those two reference functions on host arrays in the
//   reference's layouts.
//
// The reference declares both functions (src/RecSumGrid.h:99-112,
// src/RecSumV.h:102-113) and calls neither from CalcV; it has no box-pair
// driver.  The pairing of the two sets in one handle is this library's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <sstream>

#include "npme_internal.h"

namespace npme {
namespace {

// keyword defaults of a box pair: the grid must be given (the error model of
// the DM split does not apply to a translated kernel); the rest as
// NPME_AddMissingKeywords sets them (src/Interface.cpp:53-517).  Rdir only
// sizes the cells of the charge sort here.
void boxpair_keywords(Keywords &k, bool user)
{
  std::string s;
  for (const char *key : {"N1", "N2", "N3"})
    if (!k.get(key, s))
      throw Error{NPME_ERR_KEYWORD, std::string("a box pair needs the grid keyword '") + key + "'"};
  if (!k.get("calcType", s)) k.set("calcType", std::string("pme"));
  double Rdir = 8.0;
  if (!k.get("Rdir", Rdir)) k.set("Rdir", Rdir);
  long Bn = 8;
  if (!k.get("BnOrder", Bn)) k.set("BnOrder", Bn);
  const char *nk[3] = {"N1", "N2", "N3"}, *bk[3] = {"n1", "n2", "n3"};
  const char *dk[3] = {"del1", "del2", "del3"}, *ak[3] = {"a1", "a2", "a3"};
  for (int d = 0; d < 3; d++) {
    long N = 0;
    k.get(nk[d], N);
    if (!k.get(bk[d], s)) {
      // the largest block edge <= 20 with N a multiple of 2 n and n <= N/4
      long b = std::min<long>(20, N / 4);
      while (b > 1 && N % (2 * b) != 0) b--;
      k.set(bk[d], std::max<long>(b, 1));
    }
    if (!k.get(dk[d], s)) k.set(dk[d], Rdir);
    if (!k.get(ak[d], s)) k.set(ak[d], 4.0);
  }
  if (!k.get("nNeigh", s)) k.set("nNeigh", 2L);
  if (!k.get("nCellClust1D", s)) k.set("nCellClust1D", 3L);
  if (!user) {
    std::string funcType;
    if (!k.get("funcType", funcType))
      throw Error{NPME_ERR_KEYWORD, "missing required keyword 'funcType' ('Laplace', 'Helmholtz' or 'Ralpha')"};
    if (!k.get("EwaldSplit", s)) k.set("EwaldSplit", std::string("DerivMatch"));
    if (!k.get("nDeriv", s)) k.set("nDeriv", 6L);
    if (funcType == "Helmholtz") {
      if (!k.get("k0_r", s)) throw Error{NPME_ERR_KEYWORD, "missing required keyword 'k0_r'"};
      if (!k.get("k0_i", s)) throw Error{NPME_ERR_KEYWORD, "missing required keyword 'k0_i'"};
    } else if (funcType == "Ralpha") {
      if (!k.get("alpha", s)) throw Error{NPME_ERR_KEYWORD, "missing required keyword 'alpha'"};
    }
  }
}

// bounding-box centre (NPME_FindBoxDimCenter, src/SupportFunctions.cpp)
void centre(size_t n, const double *c, double ctr[3])
{
  double mn[3], mx[3], X0[3];
  for (int d = 0; d < 3; d++) mn[d] = mx[d] = c[d];
  for (size_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      const double x = c[3 * i + d];
      if (!std::isfinite(x)) throw Error{NPME_ERR_ARG, "coordinates must be finite"};
      mn[d] = std::min(mn[d], x);
      mx[d] = std::max(mx[d], x);
    }
  box_from_bounds(mn, mx, X0, ctr);
}

void boxpair_setup(Handle &h, const Keywords &kw, size_t nT, const double *cT, size_t nS,
                   const double *cS)
{
  if (nT < 1 || nS < 1) throw Error{NPME_ERR_ARG, "a box pair needs at least one target and one source"};
  double ct[3], cs[3];
  centre(nT, cT, ct);
  centre(nS, cS, cs);
  std::vector<double> u(3 * (nT + nS));
  for (size_t i = 0; i < nT; i++)
    for (int d = 0; d < 3; d++) u[3 * i + d] = cT[3 * i + d] - ct[d];
  for (size_t i = 0; i < nS; i++)
    for (int d = 0; d < 3; d++) u[3 * (nT + i) + d] = cS[3 * i + d] - cs[d];
  h.trans = true;
  h.nTarget = nT;
  for (int d = 0; d < 3; d++) h.transD[d] = cs[d] - ct[d];
  h.transRev = h.transFlags = 0;
  // the positions of this set-up go up with it (staged positions of another
  // set are not used)
  h.coordStaged = false;
  handle_setup(h, kw, nT + nS, u.data());
}

DMKernel kernel_from_keywords(const Keywords &k)
{
  std::string funcType;
  if (!k.get("funcType", funcType)) throw Error{NPME_ERR_KEYWORD, "missing keyword 'funcType'"};
  DMKernel ker;
  if (funcType == "Laplace") {
    ker.kind = KIND_LAPLACE;
  } else if (funcType == "Ralpha") {
    ker.kind = KIND_RALPHA;
    if (!k.get("alpha", ker.alpha)) throw Error{NPME_ERR_KEYWORD, "funcType = Ralpha needs 'alpha'"};
  } else if (funcType == "Helmholtz") {
    ker.kind = KIND_HELMHOLTZ;
    if (!k.get("k0_r", ker.k0r) || !k.get("k0_i", ker.k0i))
      throw Error{NPME_ERR_KEYWORD, "funcType = Helmholtz needs k0_r, k0_i"};
    if (ker.k0i < -1.0e-12) throw Error{NPME_ERR_KEYWORD, "k0_i is negative"};
  } else {
    throw Error{NPME_ERR_KEYWORD, "funcType = " + funcType + " is undefined"};
  }
  return ker;
}

// the arguments of NPME_RecSumGrid_FullTrans_Func: grid sizes even
// (NPME_Set_xArray, src/RecSumGrid.cpp:1026-1032), box and extension
// positive, and the separation check (src/RecSumGrid.cpp:694-705)
void fulltrans_args(const long N[3], const double del[3], const double X[3], const double R0[3])
{
  for (int d = 0; d < 3; d++) {
    if (N[d] < 2 || N[d] % 2 != 0 || N[d] > (1L << 12))
      throw Error{NPME_ERR_GRID, "grid sizes must be even and at most 4096"};
    if (!(X[d] > 0) || !(del[d] > 0)) throw Error{NPME_ERR_ARG, "X, Y, Z and del must be positive"};
  }
  const double l[3] = {2 * (X[0] + del[0]), 2 * (X[1] + del[1]), 2 * (X[2] + del[2])};
  const double rMin = min_distance_r0(R0, l);
  if (rMin < 1.0e-6) {
    std::ostringstream os;
    os << "rMin = " << rMin << " < 1.0E-6: box-box separation R0 = " << R0[0] << " " << R0[1] << " "
       << R0[2] << " should be increased";
    throw Error{NPME_ERR_ARG, os.str()};
  }
}

struct TempStream {
  hipStream_t s = nullptr;
  TempStream() { NPME_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~TempStream()
  {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

}  // namespace
}  // namespace npme

using namespace npme;

extern "C" {

// ------------------------------------------------------------- box pairs
static int create_boxpair(npme_handle **out, const npme_keywords *kw, size_t nT, const double *cT,
                          size_t nS, const double *cS, const UserKernels *u)
{
  return guarded([&] {
    if (!out || !kw || !cT || !cS) throw Error{NPME_ERR_ARG, "npme_create_boxpair: NULL argument"};
    *out = nullptr;
    Keywords k = kw->kw;
    boxpair_keywords(k, u != nullptr);
    std::unique_ptr<npme_handle> h(new npme_handle());
    if (u) {
      h->h.userK = true;
      h->h.user = *u;
    }
    try {
      boxpair_setup(h->h, k, nT, cT, nS, cS);
    } catch (...) {
      handle_free_device(h->h);
      handle_destroy_streams(h->h);
      throw;
    }
    *out = h.release();
  });
}

int npme_create_boxpair(npme_handle **out, const npme_keywords *kw, size_t nTarget,
                        const double *coordTarget, size_t nSource, const double *coordSource)
{
  return create_boxpair(out, kw, nTarget, coordTarget, nSource, coordSource, nullptr);
}

int npme_create_boxpair_genfunc_real(npme_handle **out, const npme_keywords *kw, size_t nTarget,
                                     const double *coordTarget, size_t nSource,
                                     const double *coordSource, const npme_kfunc_real *func)
{
  if (!func || !func->value) {
    set_error("npme_create_boxpair_genfunc_real: func == NULL");
    if (out) *out = nullptr;
    return NPME_ERR_ARG;
  }
  UserKernels u;
  u.complex = false;
  u.r[0] = *func;
  return create_boxpair(out, kw, nTarget, coordTarget, nSource, coordSource, &u);
}

int npme_create_boxpair_genfunc_complex(npme_handle **out, const npme_keywords *kw, size_t nTarget,
                                        const double *coordTarget, size_t nSource,
                                        const double *coordSource, const npme_kfunc_complex *func)
{
  if (!func || !func->value) {
    set_error("npme_create_boxpair_genfunc_complex: func == NULL");
    if (out) *out = nullptr;
    return NPME_ERR_ARG;
  }
  UserKernels u;
  u.complex = true;
  u.c[0] = *func;
  return create_boxpair(out, kw, nTarget, coordTarget, nSource, coordSource, &u);
}

int npme_boxpair_reset_coord(npme_handle *h, size_t nTarget, const double *coordTarget,
                             size_t nSource, const double *coordSource)
{
  return guarded([&] {
    if (!h || !coordTarget || !coordSource) throw Error{NPME_ERR_ARG, "npme_boxpair_reset_coord: NULL argument"};
    if (!h->h.trans) throw Error{NPME_ERR_ARG, "not a box-pair handle"};
    std::lock_guard<std::mutex> lk(h->h.mtx);
    Keywords kw = h->h.kw;
    boxpair_setup(h->h, kw, nTarget, coordTarget, nSource, coordSource);
  });
}

int npme_boxpair_separation(const npme_handle *h, double R0[3], double *rMin)
{
  return guarded([&] {
    if (!h || !R0 || !rMin) throw Error{NPME_ERR_ARG, "npme_boxpair_separation: NULL argument"};
    const Handle &H = h->h;
    if (!H.trans) throw Error{NPME_ERR_ARG, "not a box-pair handle"};
    if (!H.ready) throw Error{NPME_ERR_ARG, "no valid set-up: the last SetUp or ResetCoord failed"};
    // the reference's R0 of the pair: f(r) = func(|r + R0|) with r the
    // target-minus-source separation in the local frames
    for (int d = 0; d < 3; d++) R0[d] = -H.transD[d];
    *rMin = min_distance_r0(R0, H.L);
  });
}

static void boxpair_calcv(npme_handle *h, double *V, size_t nT, size_t nS, const double *qS,
                          const int *rev, bool cx)
{
  if (!h) throw Error{NPME_ERR_ARG, "NULL handle"};
  Handle &H = h->h;
  if (!H.trans) throw Error{NPME_ERR_ARG, "not a box-pair handle"};
  if (!H.ready) throw Error{NPME_ERR_ARG, "no valid set-up: the last SetUp or ResetCoord failed"};
  if (!V || !qS) throw Error{NPME_ERR_ARG, "NULL V or charge"};
  if (nT != H.nTarget || nT + nS != H.n) {
    std::ostringstream os;
    os << "nTarget, nSource = " << nT << ", " << nS << " != " << H.nTarget << ", " << H.n - H.nTarget;
    throw Error{NPME_ERR_ARG, os.str()};
  }
  if (cx != H.complexK)
    throw Error{NPME_ERR_ARG, cx ? "complex call on a real kernel" : "real call on a complex kernel"};
  const int r = rev ? ((rev[0] ? 1 : 0) | (rev[1] ? 2 : 0) | (rev[2] ? 4 : 0)) : 0;
  std::lock_guard<std::mutex> lk(H.mtx);
  handle_set_trans_rev(H, r);
  const size_t w = cx ? 2 : 1;
  // the charges of the handle: zero at the targets, then the sources
  NPME_HIP_CHECK(hipMemsetAsync(H.d_qin, 0, w * nT * sizeof(double), H.stream));
  NPME_HIP_CHECK(hipMemcpyAsync(H.d_qin + w * nT, qS, w * nS * sizeof(double), hipMemcpyHostToDevice,
                                H.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(H.stream));
  handle_calcv(H, H.d_qin, H.d_V);
  NPME_HIP_CHECK(hipMemcpyAsync(V, H.d_V, 4 * w * nT * sizeof(double), hipMemcpyDeviceToHost, H.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(H.stream));
}

int npme_boxpair_calcv_real(npme_handle *h, double *V, size_t nTarget, size_t nSource,
                            const double *chargeSource, const int reverseSign[3])
{
  return guarded([&] { boxpair_calcv(h, V, nTarget, nSource, chargeSource, reverseSign, false); });
}

int npme_boxpair_calcv_complex(npme_handle *h, double *V, size_t nTarget, size_t nSource,
                               const double *chargeSource, const int reverseSign[3])
{
  return guarded([&] { boxpair_calcv(h, V, nTarget, nSource, chargeSource, reverseSign, true); });
}

// ------------------------------------------------ reference library forms
static void fulltrans_common(double *f, const long N[3], const double ta[3], const double del[3],
                             const double X[3], const double R0[3], const DMKernel *ker,
                             const UserKernels *u)
{
  if (!f) throw Error{NPME_ERR_ARG, "NULL output"};
  fulltrans_args(N, del, X, R0);
  const std::vector<double> tab = fulltrans_tables(N, ta, del, X, R0);
  const size_t tot = (size_t)N[0] * N[1] * N[2];
  if (u) {
    // the user's func through its callback on the host
    std::vector<double> F;
    user_fgrid_host_n(*u, 0, N, tab, true, F);
    std::memcpy(f, F.data(), 2 * tot * sizeof(double));
    return;
  }
  TempStream ts;
  DevBuf dTab(tab.size() * sizeof(double)), dF(tot * sizeof(double2));
  NPME_HIP_CHECK(hipMemcpyAsync(dTab.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, ts.s));
  launch_fgrid_trans(*ker, N, (const double *)dTab.p, (double2 *)dF.p, ts.s);
  NPME_HIP_CHECK(hipMemcpyAsync(f, dF.p, tot * sizeof(double2), hipMemcpyDeviceToHost, ts.s));
  NPME_HIP_CHECK(hipStreamSynchronize(ts.s));
}

int npme_recsumgrid_fulltrans_func(double *f, long N1, double a1, double del1, long N2, double a2,
                                   double del2, long N3, double a3, double del3, double X, double Y,
                                   double Z, const npme_keywords *kernel, const double R0[3])
{
  return guarded([&] {
    if (!kernel || !R0) throw Error{NPME_ERR_ARG, "npme_recsumgrid_fulltrans_func: NULL argument"};
    const long N[3] = {N1, N2, N3};
    const double ta[3] = {a1, a2, a3}, del[3] = {del1, del2, del3}, Xb[3] = {X, Y, Z};
    const DMKernel ker = kernel_from_keywords(kernel->kw);
    fulltrans_common(f, N, ta, del, Xb, R0, &ker, nullptr);
  });
}

int npme_recsumgrid_fulltrans_func_user(double *f, long N1, double a1, double del1, long N2,
                                        double a2, double del2, long N3, double a3, double del3,
                                        double X, double Y, double Z, const npme_kfunc_real *func,
                                        const npme_kfunc_complex *cfunc, const double R0[3])
{
  return guarded([&] {
    if (!R0 || (!func == !cfunc))
      throw Error{NPME_ERR_ARG, "npme_recsumgrid_fulltrans_func_user: give R0 and exactly one of func, cfunc"};
    if ((func && !func->value) || (cfunc && !cfunc->value))
      throw Error{NPME_ERR_ARG, "npme_recsumgrid_fulltrans_func_user: NULL callback"};
    UserKernels u;
    u.complex = cfunc != nullptr;
    if (func) u.r[0] = *func;
    else u.c[0] = *cfunc;
    const long N[3] = {N1, N2, N3};
    const double ta[3] = {a1, a2, a3}, del[3] = {del1, del2, del3}, Xb[3] = {X, Y, Z};
    fulltrans_common(f, N, ta, del, Xb, R0, nullptr, &u);
  });
}

int npme_rstp_calctheta_fullF_reversesign(double *T, const double *F, const double *Qr, long N1,
                                          long N2, long N3, const double *lamda1_2,
                                          const double *lamda2_2, const double *lamda3_2,
                                          int reverseSignX, int reverseSignY, int reverseSignZ)
{
  return guarded([&] {
    if (!T || !F || !Qr || !lamda1_2 || !lamda2_2 || !lamda3_2)
      throw Error{NPME_ERR_ARG, "npme_rstp_calctheta_fullF_reversesign: NULL argument"};
    const long N[3] = {N1, N2, N3};
    FFTPlan1D plan[3];
    for (int d = 0; d < 3; d++) {
      if (N[d] < 8 || N[d] % 4 != 0 || N[d] > 4095)
        throw Error{NPME_ERR_GRID, "grid sizes must be multiples of 4 in [8, 4095]"};
      plan[d] = make_plan((int)N[d]);
      if (plan[d].nst < 0) throw Error{NPME_ERR_GRID, "FFT size has a prime factor > 7"};
    }
    const long M1 = N1 / 2, M2 = N2 / 2, M3 = N3 / 2, Mt = M1 * M2 * M3;
    const long Kp = (N3 + 7) / 8 * 8;
    const size_t tot = (size_t)N1 * N2 * N3;
    // the compact grid in the device order u = m + N/4 from the reference's
    // p = m mod N/2, conjugated: theta(p) = sum_n conj(Q(n)) f(n - p)
    // (src/RecSumV.cpp:895-935 multiplies conj(DFT(Q)) by F) is a
    // convolution of conj(Q) with the grid read at -k
    std::vector<double> q(2 * Mt);
    for (long u1 = 0; u1 < M1; u1++)
      for (long u2 = 0; u2 < M2; u2++)
        for (long u3 = 0; u3 < M3; u3++) {
          const long p = (((u1 + M1 / 2) % M1) * M2 + (u2 + M2 / 2) % M2) * M3 + (u3 + M3 / 2) % M3;
          const long o = (u1 * M2 + u2) * M3 + u3;
          q[2 * o] = Qr[2 * p];
          q[2 * o + 1] = -Qr[2 * p + 1];
        }
    std::vector<double> lam;
    lam.insert(lam.end(), lamda1_2, lamda1_2 + N1);
    lam.insert(lam.end(), lamda2_2, lamda2_2 + N2);
    lam.insert(lam.end(), lamda3_2, lamda3_2 + N3);
    std::vector<double> tw[3];
    for (int d = 0; d < 3; d++) tw[d] = twiddles((int)N[d]);
    TempStream ts;
    hipStream_t s = ts.s;
    DevBuf dF(tot * sizeof(double2)), dLam(lam.size() * sizeof(double));
    DevBuf dGrid(Mt * sizeof(double2)), dB1(M1 * M2 * Kp * sizeof(double2)),
        dB2(M1 * N2 * Kp * sizeof(double2));
    const size_t totG = (size_t)N1 * N2 * Kp;
    DevBuf dG(totG * sizeof(double2));
    DevBuf dTw[3];
    for (int d = 0; d < 3; d++) {
      dTw[d].reset(tw[d].size() * sizeof(double));
      NPME_HIP_CHECK(hipMemcpyAsync(dTw[d].p, tw[d].data(), tw[d].size() * sizeof(double),
                                    hipMemcpyHostToDevice, s));
    }
    NPME_HIP_CHECK(hipMemcpyAsync(dF.p, F, tot * sizeof(double2), hipMemcpyHostToDevice, s));
    NPME_HIP_CHECK(hipMemcpyAsync(dLam.p, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice, s));
    NPME_HIP_CHECK(hipMemcpyAsync(dGrid.p, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice, s));
    launch_make_G_full(true, N, N3, (const double2 *)dF.p, (const double *)dLam.p, dG.p, s);
    // reversal of the reference's axis d reads F at -k_d; the convolution
    // reads it at -k on every axis besides
    const int rev = (reverseSignX ? 1 : 0) | (reverseSignY ? 2 : 0) | (reverseSignZ ? 4 : 0);
    const int gfl = (~rev) & 7;
    PassGeom a{(int)N3, (int)M3, (int)N3, M1 * M2, M1 * M2, 0, M3, 1, 0, Kp, 1};
    fft_pass_complex(true, a, plan[2], dTw[2].p, dGrid.p, dB1.p, false, s);
    PassGeom b{(int)N2, (int)M2, (int)N2, M1 * Kp, Kp, M2 * Kp, 1, Kp, N2 * Kp, 1, Kp};
    b.validInner = N3;
    fft_pass_complex(true, b, plan[1], dTw[1].p, dB1.p, dB2.p, false, s);
    fft_pass_convolve(true, true, (int)N1, (int)M1, N2 * Kp, N2 * Kp, (int)N2, (int)(N2 / 2 + 1), Kp, N3,
                      plan[0], dTw[0].p, dG.p, dB2.p, s, true, true, gfl);
    PassGeom dd{(int)N2, (int)N2, (int)M2, M1 * Kp, Kp, N2 * Kp, 1, Kp, M2 * Kp, 1, Kp};
    dd.validInner = N3;
    fft_pass_complex(true, dd, plan[1], dTw[1].p, dB2.p, dB1.p, true, s);
    PassGeom e{(int)N3, (int)N3, (int)M3, M1 * M2, M1 * M2, 0, Kp, 1, 0, M3, 1};
    fft_pass_complex(true, e, plan[2], dTw[2].p, dB1.p, dGrid.p, true, s);
    NPME_HIP_CHECK(hipMemcpyAsync(q.data(), dGrid.p, q.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    NPME_HIP_CHECK(hipStreamSynchronize(s));
    for (long u1 = 0; u1 < M1; u1++)
      for (long u2 = 0; u2 < M2; u2++)
        for (long u3 = 0; u3 < M3; u3++) {
          const long p = (((u1 + M1 / 2) % M1) * M2 + (u2 + M2 / 2) % M2) * M3 + (u3 + M3 / 2) % M3;
          const long o = (u1 * M2 + u2) * M3 + u3;
          T[2 * p] = q[2 * o];
          T[2 * p + 1] = q[2 * o + 1];
        }
  });
}

}  // extern "C"
