// Particle-side kernels of the NPME CalcV pipeline on gfx950:
//   gather   charges into cell-sorted order, for both sums in one pass
//   direct   short-range DM sum: the neighbour-list walk (a team of 4 lanes
//            per target, FP32 pair math carried into FP64 in the mixed
//            form), the cell-row walk for sets whose list does not fit, and
//            the reference's interacting cell pairs for the erfc and user
//            kernels; the list build (count and fill passes)
//   spread   B-spline charge spreading, one grid tile per wavefront (16 x 8
//            x 8 on the matrix cores for real charges, 8^3 on the VALUs for
//            complex ones), the tile accumulated in registers (no atomics;
//            every grid point sums its charges in list order, so results
//            are deterministic)
//   interp   B-spline interpolation of theta + combine with the direct sum
//            and the self term, scattered back to input order in FP64
//   exact    O(N^2) FP64 sum of the full kernel (CalcV_exact)
//   fgrid    FP64 real-space compact grid of f_LR with the overlapped
//            top-hat images (set-up)
#include "npme_internal.h"

namespace npme {

template <typename R> struct V4t;
template <> struct V4t<float> { using T = float4; };
template <> struct V4t<double> { using T = double4; };

struct DirectDev {
  int kind, Nder, nNeigh;
  double Rdir, alpha, k0r, k0i, cs;
  double beta;     // Ewald split (KIND_LAPLACE_ERF)
  double a[62], b[62];    // DM coefficients (re, im), zero beyond Nder
  float af[62], bf[62];   // the same in FP32
  int nb0, nb1, nb2;
  double mn0, mn1, mn2, csA, cellTol;  // cell boxes (neighbour-list build)
  int fxShift;     // S (mixed form)
  int fxUnit;      // 2^S: one cluster edge in fixed-point units (mixed form)
  float fxInv;     // cluster edge / 2^S: length of one fixed-point unit
  float fxClose2;  // below this r^2 a pair adds the sub-unit residuals
  int eshift;        // neighbour-list entries: source byte offsets, log2 of the source size
  double fxInvD;     // fxInv in FP64 (scale of the list walk's gradient sums)
  float fxClose2u;   // list walk: squared distance in fixed-point units below which
                     // a pair adds the residuals (2^48: pairs closer than 2^24 units)
  const void *utab;  // user kernels: short-range table in the working precision
  int utS;           // pieces per octave of the table
  double invRdir;
};

// coefficient array of the working precision
template <typename Real> __device__ __forceinline__ const Real *coef_a(const DirectDev &p);
template <> __device__ __forceinline__ const float *coef_a<float>(const DirectDev &p) { return p.af; }
template <> __device__ __forceinline__ const double *coef_a<double>(const DirectDev &p) { return p.a; }
template <typename Real> __device__ __forceinline__ const Real *coef_b(const DirectDev &p);
template <> __device__ __forceinline__ const float *coef_b<float>(const DirectDev &p) { return p.bf; }
template <> __device__ __forceinline__ const double *coef_b<double>(const DirectDev &p) { return p.b; }

// Horner in r2 of degree D with compile-time indices (the coefficients are
// zero above Nder), so the kernel-argument loads are hoisted out of the
// pair loop
template <typename Real, int D>
__device__ __forceinline__ Real horner_fix(const Real *c, Real r2)
{
  Real s = c[2 * D];
#pragma unroll
  for (int i = D - 1; i >= 0; i--) s = s * r2 + c[2 * i];
  return s;
}
// the same with the leading coefficient c[2D] passed in a register (a VGPR
// copy made once per thread keeps the first FMA to one scalar operand)
template <typename Real, int D>
__device__ __forceinline__ Real horner_top(Real top, const Real *c, Real r2)
{
  Real s = top;
#pragma unroll
  for (int i = D - 1; i >= 0; i--) s = s * r2 + c[2 * i];
  return s;
}
template <typename Real, int D>
__device__ __forceinline__ void horner_fix_cx(const Real *c, Real r2, Real &re, Real &im)
{
  Real sr = c[2 * D], si = c[2 * D + 1];
#pragma unroll
  for (int i = D - 1; i >= 0; i--) {
    sr = sr * r2 + c[2 * i];
    si = si * r2 + c[2 * i + 1];
  }
  re = sr;
  im = si;
}

// ------------------------------------------------------------------ gather
// charges in sorted order for the reciprocal sum (re and im in separate
// arrays), so spreading does not read the direct sum's packed sources
template <typename Real, bool CX>
__global__ void k_gather_q(size_t n, const int *__restrict__ perm, const double *__restrict__ q,
                           Real *__restrict__ qre, Real *__restrict__ qim)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int o = perm[s];
  if (CX) {
    qre[s] = (Real)q[2 * (size_t)o];
    qim[s] = (Real)q[2 * (size_t)o + 1];
  } else {
    qre[s] = (Real)q[o];
  }
}

// both at once: one read of each charge (the random gather through perm)
// writes the direct sum's packed source charge and the spreading copy
template <typename Real, bool CX>
__global__ void k_gather_both(size_t n, const int *__restrict__ perm, const double *__restrict__ q,
                              typename V4t<Real>::T *__restrict__ src, Real *__restrict__ qi,
                              Real *__restrict__ qre, Real *__restrict__ qim)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int o = perm[s];
  if (CX) {
    const Real re = (Real)q[2 * (size_t)o], im = (Real)q[2 * (size_t)o + 1];
    src[s].w = re;
    qi[s] = im;
    qre[s] = re;
    qim[s] = im;
  } else {
    const Real re = (Real)q[o];
    src[s].w = re;
    qre[s] = re;
  }
}

void launch_gather_both(bool fp64, bool complexK, size_t n, const int *perm, const double *q_in,
                        void *src, void *qi, void *qre, void *qim, hipStream_t s)
{
  const dim3 grid((unsigned)((n + 255) / 256));
#define NPME_GB(R, CXV)                                                                        \
  hipLaunchKernelGGL((k_gather_both<R, CXV>), grid, dim3(256), 0, s, n, perm, q_in,             \
                     (typename V4t<R>::T *)src, (R *)qi, (R *)qre, (R *)qim)
  if (fp64) {
    if (complexK) NPME_GB(double, true);
    else NPME_GB(double, false);
  } else {
    if (complexK) NPME_GB(float, true);
    else NPME_GB(float, false);
  }
#undef NPME_GB
  NPME_HIP_CHECK(hipGetLastError());
}

void launch_gather_q(bool fp64, bool complexK, size_t n, const int *perm, const double *q_in,
                     void *qre, void *qim, hipStream_t s)
{
  const dim3 grid((unsigned)((n + 255) / 256));
#define NPME_GQ(R, CXV)                                                                        \
  hipLaunchKernelGGL((k_gather_q<R, CXV>), grid, dim3(256), 0, s, n, perm, q_in, (R *)qre, (R *)qim)
  if (fp64) {
    if (complexK) NPME_GQ(double, true);
    else NPME_GQ(double, false);
  } else {
    if (complexK) NPME_GQ(float, true);
    else NPME_GQ(float, false);
  }
#undef NPME_GQ
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ direct
template <typename Real> __device__ __forceinline__ Real rsqrt_r(Real x);
__device__ __forceinline__ void sincos_r(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ void sincos_r(double x, double *s, double *c) { sincos(x, s, c); }
// v_rsq_f32 (1 ulp): r2 of a pair is never denormal, so the range fix-up
// of rsqrtf is not needed
template <> __device__ __forceinline__ float rsqrt_r<float>(float x) { return __builtin_amdgcn_rsqf(x); }
// FP64: the hardware estimate (v_rsq_f64, about 23 bits) refined by two
// Newton steps y <- y + y (1 - x y^2) / 2, whose quadratic convergence
// reaches full FP64 precision (about 1 ulp; the correctly rounded
// 1 / sqrt(x) costs a square root and a division with their fix-ups)
template <> __device__ __forceinline__ double rsqrt_r<double>(double x)
{
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

// Horner in r2 of a real coefficient set (stride-2 storage: re parts)
template <typename Real>
__device__ __forceinline__ Real horner_re(const double *c, int deg, Real r2)
{
  Real s = (Real)c[2 * deg];
  for (int i = deg - 1; i >= 0; i--) s = s * r2 + (Real)c[2 * i];
  return s;
}
template <typename Real>
__device__ __forceinline__ void horner_cx(const double *c, int deg, Real r2, Real &re, Real &im)
{
  Real sr = (Real)c[2 * deg], si = (Real)c[2 * deg + 1];
  for (int i = deg - 1; i >= 0; i--) {
    sr = sr * r2 + (Real)c[2 * i];
    si = si * r2 + (Real)c[2 * i + 1];
  }
  re = sr;
  im = si;
}

// Short-range DM direct sum.  One wavefront takes a block of 8 consecutive
// cell-sorted targets; lane (i, j) = (lane / 8, lane % 8) pairs target i with
// every 8th source of the block's source stream.  The stream is a list of
// contiguous segments of the sorted charges, built at set-up: for every row
// of cells (x, y fixed) within Rdir of the box spanned by the block's
// targets, the run of cells along z whose slab is within Rdir, cut where the
// run crosses a cluster of nCellClust1D^3 cells (inside one cluster such a
// run is contiguous in the sorted order).  All loops are uniform over the
// wavefront; only the r < Rdir mask differs per lane.  Positions are
// relative to the low corner of their cluster.  In the mixed form they are
// int32 fixed point (one cluster edge = 2^S units) in the x, y, z words of
// the float4 whose w word is the FP32 charge: the pair difference is exact in
// integers and rounded to FP32 once.  The rounding of the positions to whole
// units (an absolute error of one unit) only matters for pairs closer than
// 2^24 units, where it would exceed FP32 relative accuracy; those rare pairs
// add the sub-unit residuals of both positions (a second array, read only
// then), and their integer difference converts to FP32 exactly.  Per segment the lanes accumulate in the
// working precision; segment partials are added in FP64 and the 8 lanes of
// a target reduced by shuffles at the end.  Workgroups are remapped so that
// each XCD works through a contiguous range of blocks (its L2 then holds the
// shared neighbourhoods).
__device__ __forceinline__ int xcd_remap(int wg, int nwg)
{
  const int x = wg & 7, local = wg >> 3, q = nwg >> 3, r = nwg & 7;
  return x * q + min(x, r) + local;
}

template <typename Real, int KIND, bool FIXP, int MAXD>
__global__ __launch_bounds__(256) void k_direct(DirectDev p, size_t n,
                                                const typename V4t<Real>::T *__restrict__ src,
                                                const float4 *__restrict__ res,
                                                const Real *__restrict__ qi,
                                                const int4 *__restrict__ cell,
                                                const int *__restrict__ segStart,
                                                const int4 *__restrict__ seg, int nBlocks,
                                                int nP, double csA, double *__restrict__ Vdir)
{
  using V4 = typename V4t<Real>::T;
  constexpr bool CX = kind_is_complex(KIND);
  const int lane = threadIdx.x & 63;
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const int blk = wg * 4 + (threadIdx.x >> 6);
  if (blk >= nBlocks) return;  // uniform per wavefront
  const int ti = lane >> 3, tj = lane & 7;
  const size_t t = (size_t)blk * 8 + ti;
  const bool tvalid = t < n;
  const size_t tt = tvalid ? t : (size_t)blk * 8;
  const V4 me = src[tt];
  int mx = 0, my = 0, mz = 0;
  float4 mres = make_float4(0, 0, 0, 0);
  if constexpr (FIXP) {
    mx = __float_as_int(me.x);
    my = __float_as_int(me.y);
    mz = __float_as_int(me.z);
    mres = res[tt];
  }
  const int4 ct = cell[tt];
  const int atx = ct.x / nP, aty = ct.y / nP, atz = ct.z / nP;
  const Real R2 = (Real)(p.Rdir * p.Rdir);
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int sb = segStart[blk], se = segStart[blk + 1];
  for (int g0 = sb; g0 < se; g0 += 64) {
    // one descriptor per lane, then broadcast
    const int4 mine = (g0 + lane < se) ? seg[g0 + lane] : make_int4(0, 0, 0, 0);
    const int cnt = min(64, se - g0);
    for (int k = 0; k < cnt; k++) {
      const int segB = __shfl(mine.x, k, 64), segE = __shfl(mine.y, k, 64);
      const int ax = __shfl(mine.z, k, 64), ayz = __shfl(mine.w, k, 64);
      const Real offx = (Real)((ax - atx) * csA);
      const Real offy = (Real)(((ayz >> 16) - aty) * csA);
      const Real offz = (Real)(((ayz & 0xffff) - atz) * csA);
      const int iox = mx - (ax - atx) * p.fxUnit, ioy = my - ((ayz >> 16) - aty) * p.fxUnit,
                ioz = mz - ((ayz & 0xffff) - atz) * p.fxUnit;
      Real part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      int js = segB + tj;
      V4 o;
      if (js < segE) o = src[js];
      for (int s0 = segB; s0 < segE; s0 += 8) {
        const int jn = js + 8;
        V4 on;
        if (jn < segE) on = src[jn];  // prefetch the next source of this lane
        if (js < segE && tvalid && (size_t)js != t) {
          Real rx, ry, rz, r2;
          if constexpr (FIXP) {
            const int dx = iox - __float_as_int(o.x), dy = ioy - __float_as_int(o.y),
                      dz = ioz - __float_as_int(o.z);
            rx = (Real)dx * (Real)p.fxInv;
            ry = (Real)dy * (Real)p.fxInv;
            rz = (Real)dz * (Real)p.fxInv;
            r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < (Real)p.fxClose2) {
              const float4 ro = res[js];
              rx = ((Real)dx + (mres.x - ro.x)) * (Real)p.fxInv;
              ry = ((Real)dy + (mres.y - ro.y)) * (Real)p.fxInv;
              rz = ((Real)dz + (mres.z - ro.z)) * (Real)p.fxInv;
              r2 = rx * rx + ry * ry + rz * rz;
            }
          } else {
            rx = me.x - (o.x + offx);
            ry = me.y - (o.y + offy);
            rz = me.z - (o.z + offz);
            r2 = rx * rx + ry * ry + rz * rz;
          }
          if (r2 < R2)
            pair_sr<Real, KIND, MAXD>(p, rx, ry, rz, r2, o.w, CX ? qi[js] : (Real)0, part,
                                      coef_a<Real>(p)[2 * MAXD], coef_b<Real>(p)[2 * MAXD]);
        }
        js = jn;
        o = on;
      }
#pragma unroll
      for (int c = 0; c < (CX ? 8 : 4); c++) acc[c] += (double)part[c];
    }
  }
  // reduce over the 8 source lanes of each target
#pragma unroll
  for (int c = 0; c < (CX ? 8 : 4); c++) {
    double v = acc[c];
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    acc[c] = v;
  }
  if (tvalid && tj == 0) {
    if (CX) {
#pragma unroll
      for (int c = 0; c < 8; c++) Vdir[8 * t + c] = acc[c];
    } else {
#pragma unroll
      for (int c = 0; c < 4; c++) Vdir[4 * t + c] = acc[c];
    }
  }
}

static DirectDev direct_dev(const DirectParams &p)
{
  DirectDev d;
  d.kind = p.kind;
  d.Nder = p.Nder;
  d.nNeigh = p.nNeigh;
  d.Rdir = p.Rdir;
  d.alpha = p.alpha;
  d.k0r = p.k0r;
  d.k0i = p.k0i;
  d.cs = p.cs;
  d.beta = p.beta;
  for (int i = 0; i < 62; i++) {
    const bool used = i < 2 * (p.Nder + 1);
    d.a[i] = used ? p.a[i] : 0.0;
    d.b[i] = used ? p.b[i] : 0.0;
    d.af[i] = (float)d.a[i];
    d.bf[i] = (float)d.b[i];
  }
  d.nb0 = p.nb[0];
  d.nb1 = p.nb[1];
  d.nb2 = p.nb[2];
  d.fxShift = p.fxShift;
  d.fxUnit = 1 << p.fxShift;
  d.fxInv = (float)p.fxInv;
  // pairs closer than 2^24 units: one unit exceeds FP32 relative accuracy
  const double rc = std::ldexp(p.fxInv, 24);
  d.fxClose2 = (float)(rc * rc);
  d.utab = p.utab;
  d.utS = p.utS;
  d.invRdir = 1.0 / p.Rdir;
  d.eshift = p.nlEshift;
  d.fxInvD = p.fxInv;
  d.fxClose2u = 281474976710656.0f;  // 2^48
  return d;
}

void launch_direct(bool fp64, const DirectParams &p, size_t n, const void *src, const void *res,
                   const void *qi, const int4 *cell, const int *segStart, const void *segs, int nP, double csA,
                   double *Vdir, hipStream_t s)
{
  const DirectDev d = direct_dev(p);
  const int nBlocks = (int)((n + 7) / 8);
  const dim3 grid((unsigned)((nBlocks + 3) / 4));  // 4 wavefronts per workgroup
  const int4 *sg = (const int4 *)segs;
#define NPME_DIRECT_D(R, K, HL, D)                                                              \
  hipLaunchKernelGGL((k_direct<R, K, HL, D>), grid, dim3(256), 0, s, d, n,                      \
                     (const typename V4t<R>::T *)src, (const float4 *)res, (const R *)qi, cell, \
                     segStart, sg, nBlocks, nP, csA, Vdir)
#define NPME_DIRECT(R, K, HL)                                                                   \
  do {                                                                                          \
    if (p.Nder <= 8) NPME_DIRECT_D(R, K, HL, 8);                                                \
    else if (p.Nder <= 16) NPME_DIRECT_D(R, K, HL, 16);                                         \
    else NPME_DIRECT_D(R, K, HL, 30);                                                           \
  } while (0)
  // the erfc split has no polynomial: one instance
  if (fp64) {
    if (p.kind == KIND_LAPLACE) NPME_DIRECT(double, KIND_LAPLACE, false);
    else if (p.kind == KIND_LAPLACE_ERF) NPME_DIRECT_D(double, KIND_LAPLACE_ERF, false, 8);
    else if (p.kind == KIND_RALPHA) NPME_DIRECT(double, KIND_RALPHA, false);
    else if (p.kind == KIND_USER) NPME_DIRECT_D(double, KIND_USER, false, 4);
    else if (p.kind == KIND_USER_CX) NPME_DIRECT_D(double, KIND_USER_CX, false, 4);
    else NPME_DIRECT(double, KIND_HELMHOLTZ, false);
  } else {
    if (p.kind == KIND_LAPLACE) NPME_DIRECT(float, KIND_LAPLACE, true);
    else if (p.kind == KIND_LAPLACE_ERF) NPME_DIRECT_D(float, KIND_LAPLACE_ERF, true, 8);
    else if (p.kind == KIND_RALPHA) NPME_DIRECT(float, KIND_RALPHA, true);
    else if (p.kind == KIND_USER) NPME_DIRECT_D(float, KIND_USER, true, 4);
    else if (p.kind == KIND_USER_CX) NPME_DIRECT_D(float, KIND_USER_CX, true, 4);
    else NPME_DIRECT(float, KIND_HELMHOLTZ, true);
  }
#undef NPME_DIRECT_D
#undef NPME_DIRECT
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ neighbour list
// The geometry is fixed between CalcV calls, so the pairs inside the cutoff
// are found once at set-up and stored as a sliced-ELLPACK list.  The sorted
// targets are cut into slices of kNlSlice consecutive charges (one
// wavefront), and a team of kNlTeam lanes walks each target's list: lane
// t kNlTeam + j of a slice takes entries j, j + kNlTeam, j + 2 kNlTeam, ...
// of the slice's target t.  A slice is padded to its longest list rounded
// up to 4 kNlTeam, and each lane's entries go in groups of 4: entry k of
// target t sits at
//   sliceStart + 256 (k / (4 kNlTeam)) + 4 (t kNlTeam + k % kNlTeam)
//              + (k / kNlTeam) % 4,
// so group g of a wavefront is 1 KB contiguous, read by one 16-byte load
// per lane.  The lanes of a team read consecutive entries of one list at
// each step, which the build keeps in stencil-cell order and ascending
// source order within a cell, so one gather of a wavefront touches a few
// cache lines per team rather than one per lane.  An entry is the byte
// offset of its source in the sorted source array.  A padding entry is the
// sentinel source n (stored after the n charges, charge 0, placed so that
// its distance to every target is finite and non-zero): it adds exactly
// zero and the loop needs no test for it.  The pairs are found in FP64 with
// a cutoff inflated by 1e-9 relative and every entry is evaluated (the
// short-range function vanishes at Rdir with its first Nder derivatives).
unsigned nlist_pad(size_t sentinel, int eshift) { return (unsigned)sentinel << eshift; }

// word of entry 0 of sorted target s, and of its entry k
__device__ __forceinline__ long nl_target_base(const long *__restrict__ sliceStart, size_t s)
{
  return sliceStart[s / kNlSlice] + 4L * kNlTeam * (long)(s % kNlSlice);
}
__device__ __forceinline__ long nl_slot(long base, int k)
{
  return base + 256L * (k >> (2 + kNlTeamLog)) + 4 * (k & (kNlTeam - 1)) + ((k >> kNlTeamLog) & 3);
}

// walk the cells of the stencil of sorted target s in FP64 input
// coordinates; COUNT mode counts, FILL mode writes entries
// where the cell (gx, gy, gz) lies for a target at (x, y, z): 0 no charge of
// it is within the cutoff, 2 every charge is, 1 each must be tested.  The
// cell's box is widened by cellTol (charges sit in their cell's box up to
// the rounding of the cell search), and the two decisions keep a relative
// margin of 1e-9 to the cutoff test, so they agree with it charge by charge.
__device__ __forceinline__ int nl_cell_class(const DirectDev &p, int gx, int gy, int gz, int nP,
                                             double x, double y, double z, double R2)
{
  double near2 = 0, far2 = 0;
  const int g[3] = {gx, gy, gz};
  const double mn[3] = {p.mn0, p.mn1, p.mn2}, pt[3] = {x, y, z};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double lo = mn[d] + (double)(g[d] / nP) * p.csA + (double)(g[d] % nP) * p.cs - p.cellTol;
    const double hi = lo + p.cs + 2.0 * p.cellTol;
    const double a = fmax(0.0, fmax(lo - pt[d], pt[d] - hi));
    const double b = fmax(pt[d] - lo, hi - pt[d]);
    near2 += a * a;
    far2 += b * b;
  }
  if (near2 > R2 * (1.0 + 1.0e-9)) return 0;
  if (far2 < p.Rdir * p.Rdir) return 2;
  return 1;
}

template <bool FILL>
__global__ __launch_bounds__(256) void k_nlist(DirectDev p, size_t n, const int4 *__restrict__ cell,
                                               const int2 *__restrict__ cellSE,
                                               const double4 *__restrict__ xs, int nP,
                                               int *__restrict__ count,
                                               const long *__restrict__ sliceStart,
                                               unsigned *__restrict__ list)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int4 c = cell[s];
  const double4 me = xs[s];
  const double xi = me.x, yi = me.y, zi = me.z;
  const double R = p.Rdir * (1.0 + 1.0e-9);
  const double R2 = R * R;
  const int nN = p.nNeigh;
  const int ax = c.x / nP, ay = c.y / nP, az = c.z / nP;
  long base = 0;  // the list is filled with padding before this pass
  if (FILL) base = nl_target_base(sliceStart, s);
  int k = 0;
  for (int dx = -nN; dx <= nN; dx++)
    for (int dy = -nN; dy <= nN; dy++)
      for (int dz = -nN; dz <= nN; dz++) {
        const int gx = c.x + dx, gy = c.y + dy, gz = c.z + dz;
        if (gx < 0 || gy < 0 || gz < 0 || gx >= p.nb0 || gy >= p.nb1 || gz >= p.nb2) continue;
        const int cls = nl_cell_class(p, gx, gy, gz, nP, xi, yi, zi, R2);
        if (cls == 0) continue;
        const int2 se = cellSE[((long)gx * p.nb1 + gy) * p.nb2 + gz];
        if (!FILL && cls == 2) {
          k += se.y - ((size_t)se.x <= s && s < (size_t)(se.x + se.y) ? 1 : 0);
          continue;
        }
        const unsigned code = (unsigned)((gx / nP - ax + 1) | ((gy / nP - ay + 1) << 2) |
                                         ((gz / nP - az + 1) << 4));
        for (int j = se.x; j < se.x + se.y; j++) {
          if ((size_t)j == s) continue;
          bool in = cls == 2;
          if (!in) {
            const double4 o = xs[j];
            const double rx = xi - o.x, ry = yi - o.y, rz = zi - o.z;
            in = rx * rx + ry * ry + rz * rz < R2;
          }
          if (in) {
            if (FILL) list[nl_slot(base, k)] = (unsigned)j << p.eshift;
            k++;
          }
        }
      }
  if (!FILL) count[s] = k;
}

// the same walk with a wavefront per target, for small charge sets (a
// thread per target leaves most of the GPU idle below ~2e5 charges, and its
// loop over ~1e3 candidates is latency-bound).  The wave first reads the
// (start, count) of its stencil cells in parallel and scans the counts, so
// the candidates of all cells form one sequence t = 0..T-1 (cell by cell in
// the loop order of k_nlist, j ascending within a cell); it then tests 64
// candidates at a time, each lane finding its cell by a binary search over
// the scanned counts, and places the hits with a ballot.  The entries and
// their order are those of k_nlist.  FILL writes single entries into a
// list already filled with padding.  Stencils up to nNeigh = 3.

// charge sets up to this size take the wavefront-per-target walk, larger
// ones the wavefront-per-cell walk (measured crossover ~3e4 charges)
constexpr size_t NLIST_WAVE_MAX = 32768;
constexpr int NLW_NN = 3;
constexpr unsigned NLW_ALLIN = 0x80u;  // stencil cell wholly within the cutoff
constexpr int NLW_MAXC = (2 * NLW_NN + 1) * (2 * NLW_NN + 1) * (2 * NLW_NN + 1);

template <bool FILL>
__global__ __launch_bounds__(256) void k_nlist_wave(DirectDev p, size_t n,
                                                    const int4 *__restrict__ cell,
                                                    const int2 *__restrict__ cellSE,
                                                    const double4 *__restrict__ xs, int nP,
                                                    int *__restrict__ count,
                                                    const long *__restrict__ sliceStart,
                                                    unsigned *__restrict__ list)
{
  __shared__ int shStart[4][NLW_MAXC];
  __shared__ int shPre[4][NLW_MAXC];  // candidates before the cell
  __shared__ unsigned char shCode[4][NLW_MAXC];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const size_t s = (size_t)blockIdx.x * 4 + w;
  const bool live = s < n;
  const size_t sc = live ? s : n - 1;
  const int4 c = cell[sc];
  const double4 me = xs[sc];
  const double R = p.Rdir * (1.0 + 1.0e-9);
  const double R2 = R * R;
  const int nN = p.nNeigh, W = 2 * nN + 1, M = W * W * W;
  const int ax = c.x / nP, ay = c.y / nP, az = c.z / nP;
  int T = 0;
  for (int c0 = 0; c0 < M; c0 += 64) {
    const int ci = c0 + lane;
    int st = 0, cnt = 0;
    unsigned code = 0;
    if (ci < M) {
      const int gx = c.x + ci / (W * W) - nN, gy = c.y + (ci / W) % W - nN, gz = c.z + ci % W - nN;
      const int cls = (gx >= 0 && gy >= 0 && gz >= 0 && gx < p.nb0 && gy < p.nb1 && gz < p.nb2)
                          ? nl_cell_class(p, gx, gy, gz, nP, me.x, me.y, me.z, R2) : 0;
      if (cls) {
        const int2 se = cellSE[((long)gx * p.nb1 + gy) * p.nb2 + gz];
        st = se.x;
        cnt = se.y;
        code = (unsigned)((gx / nP - ax + 1) | ((gy / nP - ay + 1) << 2) | ((gz / nP - az + 1) << 4)) |
               (cls == 2 ? NLW_ALLIN : 0u);
      }
    }
    int inc = cnt;
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(inc, o);
      if (lane >= o) inc += v;
    }
    if (ci < M) {
      shStart[w][ci] = st;
      shPre[w][ci] = T + inc - cnt;
      shCode[w][ci] = (unsigned char)code;
    }
    T += __shfl(inc, 63);
  }
  __syncthreads();
  if (!live) return;  // after the only barrier
  long base = 0;
  if (FILL) base = nl_target_base(sliceStart, s);
  const unsigned long long below = (1ull << lane) - 1;
  int k = 0;
  for (int t0 = 0; t0 < T; t0 += 64) {
    const int t = t0 + lane;
    bool hit = false;
    unsigned ent = 0;
    if (t < T) {
      // last cell whose first candidate is at or before t (empty cells
      // share their start with the next cell, so it is the cell holding t)
      int lo = 0, hi = M - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (shPre[w][mid] <= t) lo = mid;
        else hi = mid - 1;
      }
      const int j = shStart[w][lo] + (t - shPre[w][lo]);
      const unsigned code = shCode[w][lo];
      if ((size_t)j != s) {
        if (code & NLW_ALLIN) {
          hit = true;
        } else {
          const double4 o = xs[j];
          const double rx = me.x - o.x, ry = me.y - o.y, rz = me.z - o.z;
          hit = rx * rx + ry * ry + rz * rz < R2;
        }
        ent = (unsigned)j << p.eshift;
      }
    }
    const unsigned long long m = __ballot(hit);
    if (FILL && hit) {
      const int pos = k + __popcll(m & below);
      list[nl_slot(base, pos)] = ent;
    }
    k += __popcll(m);
  }
  if (!FILL && lane == 0) count[s] = k;
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int l)
{
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}

// stencil of the home cell (hx, hy, hz) for a wavefront: per stencil cell
// (in the loop order of k_nlist) its first charge, the candidates before
// it and its cluster-offset code, in LDS; returns the candidate count T
__device__ __forceinline__ int nl_stencil_scan(const DirectDev &p, const int2 *__restrict__ cellSE,
                                               int nP, int hx, int hy, int hz, int lane,
                                               int *shStart, int *shPre, unsigned char *shCode)
{
  const int nN = p.nNeigh, W = 2 * nN + 1, M = W * W * W;
  const int ax = hx / nP, ay = hy / nP, az = hz / nP;
  int T = 0;
  for (int c0 = 0; c0 < M; c0 += 64) {
    const int ci = c0 + lane;
    int st = 0, cnt = 0;
    unsigned code = 0;
    if (ci < M) {
      const int gx = hx + ci / (W * W) - nN, gy = hy + (ci / W) % W - nN, gz = hz + ci % W - nN;
      if (gx >= 0 && gy >= 0 && gz >= 0 && gx < p.nb0 && gy < p.nb1 && gz < p.nb2) {
        const int2 se = cellSE[((long)gx * p.nb1 + gy) * p.nb2 + gz];
        st = se.x;
        cnt = se.y;
        code = (unsigned)((gx / nP - ax + 1) | ((gy / nP - ay + 1) << 2) | ((gz / nP - az + 1) << 4));
      }
    }
    int inc = cnt;
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(inc, o);
      if (lane >= o) inc += v;
    }
    if (ci < M) {
      shStart[ci] = st;
      shPre[ci] = T + inc - cnt;
      shCode[ci] = (unsigned char)code;
    }
    T += __shfl(inc, 63);
  }
  return T;
}

// candidate t of the scanned stencil: source index and list entry
__device__ __forceinline__ int nl_candidate(int t, int M, const int *shStart, const int *shPre,
                                            const unsigned char *shCode, int eshift, unsigned &ent)
{
  // last cell whose first candidate is at or before t (empty cells share
  // their start with the next cell, so it is the cell holding t)
  int lo = 0, hi = M - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (shPre[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const int j = shStart[lo] + (t - shPre[lo]);
  (void)shCode;
  ent = (unsigned)j << eshift;
  return j;
}

// the same walk with a wavefront per cell, for large charge sets: the
// targets of one cell share their stencil, so the wave loads the stencil's
// candidates once, 64 at a time (the flattened sequence of k_nlist_wave),
// and tests each chunk against the cell's targets in turn, one ballot per
// target.  Lane i holds target i's position, list offset and entry count
// (targets in batches of 64).  This is the count pass; it keeps its
// ballots, the hit mask of target tb + i for chunk k at
// masks[maskStart[c] + k * (targets of c) + tb + i], and
// k_nlist_cell_expand fills the list from them, so the entries and their
// order are those of k_nlist.
__global__ __launch_bounds__(256) void k_nlist_cell_count(DirectDev p, long ncell,
                                                          const int2 *__restrict__ cellSE,
                                                          const double4 *__restrict__ xs, int nP,
                                                          int *__restrict__ count,
                                                          const long *__restrict__ maskStart,
                                                          unsigned long long *__restrict__ masks)
{
  __shared__ int shStart[4][NLW_MAXC];
  __shared__ int shPre[4][NLW_MAXC];
  __shared__ unsigned char shCode[4][NLW_MAXC];
  // wave-uniform values live in scalar registers: the cell, its targets
  // and their positions (scalar loads of xs at a uniform address)
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const long c = (long)blockIdx.x * 4 + w;
  const int2 home = c < ncell ? cellSE[c] : make_int2(0, 0);
  const int hz = (int)(c % p.nb2), hy = (int)((c / p.nb2) % p.nb1), hx = (int)(c / ((long)p.nb1 * p.nb2));
  const double R = p.Rdir * (1.0 + 1.0e-9);
  const double R2 = R * R;
  const int nN = p.nNeigh, W = 2 * nN + 1, M = W * W * W;
  int T = 0;
  if (home.y > 0) T = nl_stencil_scan(p, cellSE, nP, hx, hy, hz, lane, shStart[w], shPre[w], shCode[w]);
  __syncthreads();
  if (home.y == 0) return;  // after the only barrier
  const long mbase = maskStart[c];
  for (int tb = 0; tb < home.y; tb += 64) {
    const int mt = min(64, home.y - tb);
    const int s0 = home.x + tb;  // target i is s0 + i
    const double4 *tpos = xs + s0;
    int kreg = 0;  // lane i: entries of target s0 + i so far
    for (int t0 = 0; t0 < T; t0 += 64) {
      const int t = t0 + lane;
      // a lane past the candidates tests a point far outside every cutoff
      int j = -1;
      double ox = 1.0e150, oy = 1.0e150, oz = 1.0e150;
      if (t < T) {
        unsigned ent;
        j = nl_candidate(t, M, shStart[w], shPre[w], shCode[w], p.eshift, ent);
        const double4 o = xs[j];
        ox = o.x;
        oy = o.y;
        oz = o.z;
      }
      unsigned long long mym = 0;
      // two targets per step: their distance tests are independent
      int i = 0;
      for (; i + 1 < mt; i += 2) {
        const double4 ta = tpos[i], tb2 = tpos[i + 1];
        const double ax = ta.x - ox, ay = ta.y - oy, az = ta.z - oz;
        const double bx = tb2.x - ox, by = tb2.y - oy, bz = tb2.z - oz;
        // & rather than &&: no branch around the distance test
        const bool ha = (j != s0 + i) & (ax * ax + ay * ay + az * az < R2);
        const bool hb = (j != s0 + i + 1) & (bx * bx + by * by + bz * bz < R2);
        const unsigned long long ma = __ballot(ha), mb = __ballot(hb);
        mym = lane == i ? ma : (lane == i + 1 ? mb : mym);
      }
      if (i < mt) {
        const double4 tp = tpos[i];
        const double rx = tp.x - ox, ry = tp.y - oy, rz = tp.z - oz;
        const bool hit = (j != s0 + i) & (rx * rx + ry * ry + rz * rz < R2);
        const unsigned long long m = __ballot(hit);
        mym = lane == i ? m : mym;
      }
      if (lane < mt) masks[mbase + (long)(t0 >> 6) * home.y + tb + lane] = mym;
      kreg += __popcll(mym);
    }
    if (lane < mt) count[s0 + lane] = kreg;
  }
}

// fill pass from the count pass's hit masks: the same candidate sequence,
// without positions or distance tests.  Target i's mask and list offset are
// wave-uniform (scalar loads); a hit lane's place is the target's entry
// count plus the hits in lanes below it (mbcnt).
__global__ __launch_bounds__(256) void k_nlist_cell_expand(DirectDev p, long ncell,
                                                           const int2 *__restrict__ cellSE, int nP,
                                                           const long *__restrict__ maskStart,
                                                           const unsigned long long *__restrict__ masks,
                                                           const long *__restrict__ sliceStart,
                                                           unsigned *__restrict__ list)
{
  __shared__ int shStart[4][NLW_MAXC];
  __shared__ int shPre[4][NLW_MAXC];
  __shared__ unsigned char shCode[4][NLW_MAXC];
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const long c = (long)blockIdx.x * 4 + w;
  const int2 home = c < ncell ? cellSE[c] : make_int2(0, 0);
  const int hz = (int)(c % p.nb2), hy = (int)((c / p.nb2) % p.nb1), hx = (int)(c / ((long)p.nb1 * p.nb2));
  const int nN = p.nNeigh, W = 2 * nN + 1, M = W * W * W;
  int T = 0;
  if (home.y > 0) T = nl_stencil_scan(p, cellSE, nP, hx, hy, hz, lane, shStart[w], shPre[w], shCode[w]);
  __syncthreads();
  if (home.y == 0) return;  // after the only barrier
  const long mbase = maskStart[c];
  for (int tb = 0; tb < home.y; tb += 64) {
    const int mt = min(64, home.y - tb);
    const int s0 = home.x + tb;
    // lane i: list offset of target s0 + i (its entry 0), taken per
    // target with readlane
    const int si = s0 + min(lane, mt - 1);
    const long obase = nl_target_base(sliceStart, (size_t)si);
    int kreg = 0;
    for (int t0 = 0; t0 < T; t0 += 64) {
      const int t = t0 + lane;
      unsigned ent = 0;
      if (t < T) (void)nl_candidate(t, M, shStart[w], shPre[w], shCode[w], p.eshift, ent);
      const unsigned long long *mrow = masks + mbase + (long)(t0 >> 6) * home.y + tb;
      const unsigned long long mym = lane < mt ? mrow[lane] : 0ull;
      for (int i = 0; i < mt; i++) {
        const unsigned long long m = mrow[i];
        if (m == 0) continue;  // uniform
        const int ki = __builtin_amdgcn_readlane(kreg, i);
        const long ob = (long)readlane_u64((unsigned long long)obase, i);
        if ((m >> lane) & 1ull) {
          const int pos = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)m, (unsigned)ki));
          list[nl_slot(ob, pos)] = ent;
        }
      }
      kreg += __popcll(mym);
    }
  }
}

// per cell: the hit masks the count pass keeps (targets x candidate
// chunks); words[ncell] = 0 closes the scan
__global__ __launch_bounds__(256) void k_nlist_cell_words(DirectDev p, long ncell,
                                                          const int2 *__restrict__ cellSE,
                                                          long *__restrict__ words)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncell) return;
  if (c == ncell) {
    words[c] = 0;
    return;
  }
  const int2 home = cellSE[c];
  long T = 0;
  if (home.y > 0) {
    const int hz = (int)(c % p.nb2), hy = (int)((c / p.nb2) % p.nb1), hx = (int)(c / ((long)p.nb1 * p.nb2));
    const int nN = p.nNeigh;
    for (int dx = -nN; dx <= nN; dx++)
      for (int dy = -nN; dy <= nN; dy++)
        for (int dz = -nN; dz <= nN; dz++) {
          const int gx = hx + dx, gy = hy + dy, gz = hz + dz;
          if (gx < 0 || gy < 0 || gz < 0 || gx >= p.nb0 || gy >= p.nb1 || gz >= p.nb2) continue;
          T += cellSE[((long)gx * p.nb1 + gy) * p.nb2 + gz].y;
        }
  }
  words[c] = (long)home.y * ((T + 63) / 64);
}

void launch_nlist(bool fill, const DirectParams &p, size_t n, const int4 *cell, const int *cellStart,
                  const double4 *xs, int nP, int *count, const long *sliceStart, unsigned *list,
                  NlistWork &work, hipStream_t s)
{
  DirectDev d;
  d.Rdir = p.Rdir;
  d.nNeigh = p.nNeigh;
  d.nb0 = p.nb[0];
  d.nb1 = p.nb[1];
  d.nb2 = p.nb[2];
  d.cs = p.cs;
  d.csA = p.csA;
  d.mn0 = p.mn[0];
  d.mn1 = p.mn[1];
  d.mn2 = p.mn[2];
  d.cellTol = p.cellTol;
  d.eshift = p.nlEshift;
  // 1: a thread per target (stencils wider than NLW_NN), 2: a wavefront per
  // target (small sets), 3: a wavefront per cell
  const int pick = p.nNeigh > NLW_NN ? 1 : n <= NLIST_WAVE_MAX ? 2 : 3;
  if (pick == 3) {
    const long ncell = (long)p.nb[0] * p.nb[1] * p.nb[2];
    const dim3 g((unsigned)((ncell + 3) / 4));  // 4 cells (wavefronts) per workgroup
    const int2 *cse = (const int2 *)cellStart;
    if (!fill) {
      // hit masks of the count pass: about an eighth of the list's size
      work.maskStart.reset((ncell + 1) * sizeof(long));
      DevBuf words((ncell + 1) * sizeof(long));
      hipLaunchKernelGGL(k_nlist_cell_words, dim3((unsigned)((ncell + 256) / 256)), dim3(256), 0, s, d,
                         ncell, cse, (long *)words.p);
      NPME_HIP_CHECK(hipGetLastError());
      const long total = scan_words_device((const long *)words.p, (long *)work.maskStart.p,
                                           (size_t)ncell + 1, s);
      work.masks.reset(std::max<size_t>((size_t)total, 1) * 8);
      work.masked = true;
      hipLaunchKernelGGL(k_nlist_cell_count, g, dim3(256), 0, s, d, ncell, cse, xs, nP, count,
                         (const long *)work.maskStart.p, (unsigned long long *)work.masks.p);
    } else {
      if (!work.masked) throw Error{NPME_ERR_ARG, "neighbour list: fill pass without its count pass"};
      hipLaunchKernelGGL(k_nlist_cell_expand, g, dim3(256), 0, s, d, ncell, cse, nP,
                         (const long *)work.maskStart.p, (const unsigned long long *)work.masks.p,
                         sliceStart, list);
    }
    NPME_HIP_CHECK(hipGetLastError());
    return;
  }
  if (pick == 2) {
    const dim3 g((unsigned)((n + 3) / 4));  // 4 targets (wavefronts) per workgroup
    if (fill)
      hipLaunchKernelGGL(k_nlist_wave<true>, g, dim3(256), 0, s, d, n, cell, (const int2 *)cellStart,
                         xs, nP, count, sliceStart, list);
    else
      hipLaunchKernelGGL(k_nlist_wave<false>, g, dim3(256), 0, s, d, n, cell, (const int2 *)cellStart,
                         xs, nP, count, sliceStart, list);
    NPME_HIP_CHECK(hipGetLastError());
    return;
  }
  const dim3 g((unsigned)((n + 255) / 256));
  if (fill)
    hipLaunchKernelGGL(k_nlist<true>, g, dim3(256), 0, s, d, n, cell, (const int2 *)cellStart, xs,
                       nP, count, sliceStart, list);
  else
    hipLaunchKernelGGL(k_nlist<false>, g, dim3(256), 0, s, d, n, cell, (const int2 *)cellStart, xs,
                       nP, count, sliceStart, list);
  NPME_HIP_CHECK(hipGetLastError());
}

// short-range part of a user kernel from its table: f[0], f[1] = Re, Im of
// f(r) and f[2], f[3] = Re, Im of f'(r)/r (real kernels: f[1] = f[3] = 0).
// r = Rdir 2^e m with m in [0.5, 1) gives octave k = -e and x = 2m - 1 in
// [0, 1); r = Rdir takes the table's end, r > Rdir (pairs listed within a
// relative 1e-9 of the cutoff) adds zero, as the reference's short-range
// kernels do outside Rdir, and r below the last octave takes its start.
// The piece and octave indices are clamped, so every table read stays
// inside it for any r2.
template <typename Real, bool CXU>
__device__ __forceinline__ void user_sr(const DirectDev &p, Real r2, Real *f)
{
  constexpr int P = kUserDeg + 1, NQ = CXU ? 4 : 2;
  const Real u = sqrt(r2) * (Real)p.invRdir;
  int e;
  const Real m = frexp(u, &e);
  int k = -e;
  Real x = (Real)2 * m - (Real)1;
  const Real keep = (e > 1 || (e == 1 && m > (Real)0.5)) ? (Real)0 : (Real)1;
  if (k < 0) {
    k = 0;
    x = (Real)1;
  }
  if (k >= kUserOct) {
    k = kUserOct - 1;
    x = (Real)0;
  }
  const int S = p.utS;
  const Real xs = x * (Real)S;
  int j = (int)xs;
  j = min(max(j, 0), S - 1);
  const Real t = (Real)2 * (xs - (Real)j) - (Real)1;
  const Real *c = static_cast<const Real *>(p.utab) + ((size_t)k * S + j) * (NQ * P);
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    Real s = c[q * P + P - 1];
#pragma unroll
    for (int d = P - 2; d >= 0; d--) s = s * t + c[q * P + d];
    f[CXU ? q : 2 * q] = s * keep;
  }
  if (!CXU) f[1] = f[3] = (Real)0;
}

// short-range DM pair term of one source (charge qr + i qim) at offset
// (rx, ry, rz), r2 = |r|^2 < Rdir^2, added to part: V and grad (real), or
// (re, im) of V, d/dx, d/dy, d/dz (complex)
// topA, topB: the leading coefficients a[MAXD], b[MAXD] (real kernels)
template <typename Real, int KIND, int MAXD>
__device__ __forceinline__ void pair_sr(const DirectDev &p, Real rx, Real ry, Real rz, Real r2,
                                        Real qr, Real qim, Real *part, Real topA, Real topB)
{
  if constexpr (KIND == KIND_LAPLACE) {
    const Real ri = rsqrt_r<Real>(r2);
    const Real f0 = ri - horner_top<Real, MAXD>(topA, coef_a<Real>(p), r2);
    const Real f1 = -ri * ri * ri - horner_top<Real, MAXD>(topB, coef_b<Real>(p), r2);
    const Real fq = f1 * qr;
    part[0] += f0 * qr;
    part[1] += rx * fq;
    part[2] += ry * fq;
    part[3] += rz * fq;
  } else if constexpr (KIND == KIND_LAPLACE_ERF) {
    // erfc(beta r)/r and (1/r) d/dr of it
    // (src/KernelFunctionLaplace.cpp:1344-1373)
    (void)topA;
    (void)topB;
    const Real ri = rsqrt_r<Real>(r2);
    const Real b = (Real)p.beta;
    const Real f0 = erfc(b * (r2 * ri)) * ri;
    const Real g = (Real)(p.beta * 1.12837916709551257390) * exp(-(b * b) * r2);  // 2 beta/sqrt(pi)
    const Real f1 = -(f0 + g) * ri * ri;
    const Real fq = f1 * qr;
    part[0] += f0 * qr;
    part[1] += rx * fq;
    part[2] += ry * fq;
    part[3] += rz * fq;
  } else if constexpr (KIND == KIND_USER) {
    (void)topA;
    (void)topB;
    Real f[4];
    user_sr<Real, false>(p, r2, f);
    const Real fq = f[2] * qr;
    part[0] += f[0] * qr;
    part[1] += rx * fq;
    part[2] += ry * fq;
    part[3] += rz * fq;
  } else if constexpr (KIND == KIND_USER_CX) {
    (void)topA;
    (void)topB;
    Real f[4];
    user_sr<Real, true>(p, r2, f);
    part[0] += f[0] * qr - f[1] * qim;
    part[1] += f[0] * qim + f[1] * qr;
    const Real gr = f[2] * qr - f[3] * qim, gi = f[2] * qim + f[3] * qr;
    part[2] += rx * gr;
    part[3] += rx * gi;
    part[4] += ry * gr;
    part[5] += ry * gi;
    part[6] += rz * gr;
    part[7] += rz * gi;
  } else if constexpr (KIND == KIND_RALPHA) {
    const Real fe = pow(r2, (Real)(0.5 * p.alpha));
    const Real f0 = fe - horner_top<Real, MAXD>(topA, coef_a<Real>(p), r2);
    const Real f1 = (Real)p.alpha * fe / r2 - horner_top<Real, MAXD>(topB, coef_b<Real>(p), r2);
    const Real fq = f1 * qr;
    part[0] += f0 * qr;
    part[1] += rx * fq;
    part[2] += ry * fq;
    part[3] += rz * fq;
  } else {
    (void)topA;
    (void)topB;
    // exp(i k0 r)/r with the phase k0r r reduced modulo 2 pi in FP64
    const Real r = sqrt(r2);
    const Real ri = (Real)1 / r;
    const double ph = p.k0r * (double)r;
    const double red = ph - 6.283185307179586477 * rint(ph * 0.15915494309189533577);
    Real sn, cn;
    sincos_r((Real)red, &sn, &cn);
    const Real damp = exp((Real)(-p.k0i) * r) * ri;
    const Real fr = damp * cn, fi = damp * sn;
    const Real k0r = (Real)p.k0r, k0i = (Real)p.k0i;
    const Real ikr = -k0i - ri, iki = k0r;  // i k0 - 1/r
    const Real f1r = (ikr * fr - iki * fi) * ri, f1i = (ikr * fi + iki * fr) * ri;
    Real pr, pim, br, bi;
    horner_fix_cx<Real, MAXD>(coef_a<Real>(p), r2, pr, pim);
    horner_fix_cx<Real, MAXD>(coef_b<Real>(p), r2, br, bi);
    const Real s0r = fr - pr, s0i = fi - pim, s1r = f1r - br, s1i = f1i - bi;
    part[0] += s0r * qr - s0i * qim;
    part[1] += s0r * qim + s0i * qr;
    const Real gr = s1r * qr - s1i * qim, gi = s1r * qim + s1i * qr;
    part[2] += rx * gr;
    part[3] += rx * gi;
    part[4] += ry * gr;
    part[5] += ry * gi;
    part[6] += rz * gr;
    part[7] += rz * gi;
  }
}

// direct sum over the neighbour list: a team of kNlTeam lanes per target,
// the kNlSlice targets of a slice walk their lists in lockstep (lists of one
// slice have similar lengths, so little of the wavefront idles); the team's
// sums are added across its lanes at the end
template <typename Real, int KIND, bool FIXP, int MAXD>
__device__ __forceinline__ void direct_slice(const DirectDev &p, size_t n, size_t slice, int lane,
                                             const typename V4t<Real>::T *__restrict__ src,
                                             const float4 *__restrict__ res,
                                             const Real *__restrict__ qi,
                                             const long *__restrict__ sliceStart,
                                             const int *__restrict__ sliceLen,
                                             const unsigned *__restrict__ list, unsigned pad,
                                             double csA, double *__restrict__ Vdir, Real topA,
                                             Real topB)
{
  using V4 = typename V4t<Real>::T;
  constexpr bool CX = kind_is_complex(KIND);
  const size_t s = slice * kNlSlice + (lane >> kNlTeamLog);
  const bool tvalid = s < n;
  const size_t tt = tvalid ? s : slice * kNlSlice;
  const V4 me = src[tt];
  // positions are in one frame for all charges (FP64: relative to the low
  // corner of the cells; mixed: int32 fixed point of that, wrapped modulo
  // 2^32), so a pair's difference needs no cluster offset.  Listed pairs lie
  // within Rdir < 2^31 units, where the wrapped int32 difference is exact.
  int mx = 0, my = 0, mz = 0;
  float4 mres = make_float4(0, 0, 0, 0);
  if constexpr (FIXP) {
    mx = __float_as_int(me.x);
    my = __float_as_int(me.y);
    mz = __float_as_int(me.z);
    mres = res[tt];
  }
  (void)csA;
  // the slice is the wavefront's: its length is uniform, so the loop and
  // the bounds of the entry loads branch on scalars
  const int len = __builtin_amdgcn_readfirstlane(sliceLen[slice]);
  const uint4 *e4 = reinterpret_cast<const uint4 *>(list + sliceStart[slice]) + lane;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  Real part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // U entries per step.  Two buffer sets alternate (the loop body is the
  // step unrolled twice), so while step i is evaluated the source gathers
  // of step i+1 and the entries of step i+2 are in flight without register
  // copies of loads in flight.  Lanes past n hold padding entries only.
  // An entry is the byte offset of its source in src (and, for the mixed
  // form, in res, whose elements have the same 16-byte size).
  constexpr int U = 4;
  static_assert(U == 4, "one group of 4 list entries per step");
  const char *srcb = reinterpret_cast<const char *>(src);
  const char *resb = reinterpret_cast<const char *>(res);
  constexpr int ESH = sizeof(V4) == 32 ? 5 : 4;
  auto load_ent = [&](int k0, unsigned *ent) {  // k0 a multiple of 4, len too
    uint4 g = make_uint4(pad, pad, pad, pad);
    if (k0 < len) g = e4[(long)(k0 >> 2) * 64];
    ent[0] = g.x;
    ent[1] = g.y;
    ent[2] = g.z;
    ent[3] = g.w;
  };
  auto gather = [&](const unsigned *ent, V4 *o, Real *oqi) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      o[u] = *reinterpret_cast<const V4 *>(srcb + ent[u]);
      if constexpr (CX) oqi[u] = qi[ent[u] >> ESH];
    }
  };
  // the mixed form keeps the pair offsets in fixed-point units: the squared
  // distance is scaled once (with a floor of 1e-24 that keeps the padding
  // sentinel's terms finite and changes no real pair), and the gradient
  // sums are scaled by the unit length once per target
  const Real u2 = (Real)(p.fxInvD * p.fxInvD);
  auto eval = [&](const unsigned *ent, const V4 *o, const Real *oqi) {
    Real rx[U], ry[U], rz[U], r2[U];
    if constexpr (FIXP) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        rx[u] = (Real)(mx - __float_as_int(o[u].x));
        ry[u] = (Real)(my - __float_as_int(o[u].y));
        rz[u] = (Real)(mz - __float_as_int(o[u].z));
        r2[u] = rx[u] * rx[u] + ry[u] * ry[u] + rz[u] * rz[u];
      }
      // rare (pairs closer than 2^24 units): add the sub-unit residuals of
      // both positions; one wave-uniform test per step on the closest pair
      const Real rmin = fmin(fmin(r2[0], r2[1]), fmin(r2[2], r2[3]));
      if (__builtin_amdgcn_ballot_w64(rmin < (Real)p.fxClose2u)) {
#pragma unroll
        for (int u = 0; u < U; u++) {
          if (r2[u] < (Real)p.fxClose2u) {
            const float4 ro = *reinterpret_cast<const float4 *>(resb + ent[u]);
            const int dx = mx - __float_as_int(o[u].x), dy = my - __float_as_int(o[u].y),
                      dz = mz - __float_as_int(o[u].z);
            rx[u] = (Real)dx + (mres.x - ro.x);
            ry[u] = (Real)dy + (mres.y - ro.y);
            rz[u] = (Real)dz + (mres.z - ro.z);
            r2[u] = rx[u] * rx[u] + ry[u] * ry[u] + rz[u] * rz[u];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) r2[u] = fma(r2[u], u2, (Real)1.0e-24);
    } else {
#pragma unroll
      for (int u = 0; u < U; u++) {
        rx[u] = me.x - o[u].x;
        ry[u] = me.y - o[u].y;
        rz[u] = me.z - o[u].z;
        r2[u] = rx[u] * rx[u] + ry[u] * ry[u] + rz[u] * rz[u];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      pair_sr<Real, KIND, MAXD>(p, rx[u], ry[u], rz[u], r2[u], o[u].w, CX ? oqi[u] : (Real)0, part,
                                topA, topB);
  };
  auto carry = [&](int k0) {
    if (((k0 + U) & 7) == 0) {  // FP64 carry every 8 entries
#pragma unroll
      for (int c = 0; c < (CX ? 8 : 4); c++) {
        acc[c] += (double)part[c];
        part[c] = 0;
      }
    }
  };
  unsigned E[U], DA[U], DB[U];
  V4 OA[U], OB[U];
  Real QA[U], QB[U];
  if (len > 0) {
    load_ent(0, DA);
    gather(DA, OA, QA);
    load_ent(U, E);
  }
  for (int k0 = 0; k0 < len; k0 += 2 * U) {
    // step k0 from (DA, OA); prefetch step k0+U into (DB, OB), entries k0+2U
    const bool moreA = k0 + U < len;  // uniform per wavefront
    if (moreA) {
#pragma unroll
      for (int u = 0; u < U; u++) DB[u] = E[u];
      gather(DB, OB, QB);
      load_ent(k0 + 2 * U, E);
    }
    eval(DA, OA, QA);
    carry(k0);
    if (!moreA) break;
    // step k0+U from (DB, OB); prefetch step k0+2U into (DA, OA), entries k0+3U
    if (k0 + 2 * U < len) {
#pragma unroll
      for (int u = 0; u < U; u++) DA[u] = E[u];
      gather(DA, OA, QA);
      load_ent(k0 + 3 * U, E);
    }
    eval(DB, OB, QB);
    carry(k0 + U);
  }
#pragma unroll
  for (int c = 0; c < (CX ? 8 : 4); c++) {
    acc[c] += (double)part[c];
#pragma unroll
    for (int o = 1; o < kNlTeam; o <<= 1) acc[c] += __shfl_xor(acc[c], o);
  }
  if (!tvalid || (lane & (kNlTeam - 1)) != 0) return;
  if constexpr (FIXP) {  // gradient sums from fixed-point units to lengths
#pragma unroll
    for (int c = (CX ? 2 : 1); c < (CX ? 8 : 4); c++) acc[c] *= p.fxInvD;
  }
  if (CX) {
#pragma unroll
    for (int c = 0; c < 8; c++) Vdir[8 * s + c] = acc[c];
  } else {
#pragma unroll
    for (int c = 0; c < 4; c++) Vdir[4 * s + c] = acc[c];
  }
}


// Each wavefront walks a contiguous range of slices: one slice with the
// grid of a wavefront per slice used here.  (Grids of a few workgroups per
// CU, which leave wave slots to the reciprocal sum's kernels on the other
// stream, measured slower at N = 1M.)
template <typename Real, int KIND, bool FIXP, int MAXD>
__global__ __launch_bounds__(256) void k_direct_nl(DirectDev p, size_t n,
                                                   const typename V4t<Real>::T *__restrict__ src,
                                                   const float4 *__restrict__ res,
                                                   const Real *__restrict__ qi,
                                                   const long *__restrict__ sliceStart,
                                                   const int *__restrict__ sliceLen,
                                                   const unsigned *__restrict__ list, unsigned pad,
                                                   double csA, double *__restrict__ Vdir)
{
  const int lane = threadIdx.x & 63;
  const size_t nSlices = (n + kNlSlice - 1) / kNlSlice;
  const size_t w = (size_t)xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  const size_t nw = (size_t)gridDim.x * 4;
  const size_t per = nSlices / nw, extra = nSlices - per * nw;
  const size_t beg = w * per + (w < extra ? w : extra);
  const size_t end = beg + per + (w < extra ? 1 : 0);
  Real topA = coef_a<Real>(p)[2 * MAXD], topB = coef_b<Real>(p)[2 * MAXD];
  asm volatile("" : "+v"(topA), "+v"(topB));  // held in VGPRs for the pair loop
  for (size_t sl = beg; sl < end; sl++)  // uniform per wavefront
    direct_slice<Real, KIND, FIXP, MAXD>(p, n, sl, lane, src, res, qi, sliceStart, sliceLen,
                                         list, pad, csA, Vdir, topA, topB);
}

void launch_direct_nl(bool fp64, const DirectParams &p, size_t n, const void *src, const void *res,
                      const void *qi, const long *sliceStart, const int *sliceLen, const unsigned *list,
                      double csA, double *Vdir, hipStream_t s)
{
  const unsigned pad = nlist_pad(n, p.nlEshift);
  const DirectDev d = direct_dev(p);
  // a workgroup (4 wavefronts) per 4 slices
  const long nwg = (long)((n + 4 * kNlSlice - 1) / (4 * kNlSlice));
  const dim3 grid((unsigned)std::max<long>(nwg, 1));
#define NPME_DNL_D(R, K, HL, D)                                                                 \
  hipLaunchKernelGGL((k_direct_nl<R, K, HL, D>), grid, dim3(256), 0, s, d, n,                   \
                     (const typename V4t<R>::T *)src, (const float4 *)res, (const R *)qi,      \
                     sliceStart, sliceLen, list, pad, csA, Vdir)
// Horner at the exact degree for Nder 4..8 (the coefficients above Nder
// are zero, so a larger compile-time degree is also exact, only slower)
#define NPME_DNL(R, K, HL)                                                                      \
  do {                                                                                          \
    if (p.Nder <= 4) NPME_DNL_D(R, K, HL, 4);                                                   \
    else if (p.Nder == 5) NPME_DNL_D(R, K, HL, 5);                                              \
    else if (p.Nder == 6) NPME_DNL_D(R, K, HL, 6);                                              \
    else if (p.Nder == 7) NPME_DNL_D(R, K, HL, 7);                                              \
    else if (p.Nder == 8) NPME_DNL_D(R, K, HL, 8);                                              \
    else if (p.Nder <= 16) NPME_DNL_D(R, K, HL, 16);                                            \
    else NPME_DNL_D(R, K, HL, 30);                                                              \
  } while (0)
  if (fp64) {
    if (p.kind == KIND_LAPLACE) NPME_DNL(double, KIND_LAPLACE, false);
    else if (p.kind == KIND_LAPLACE_ERF) NPME_DNL_D(double, KIND_LAPLACE_ERF, false, 4);
    else if (p.kind == KIND_RALPHA) NPME_DNL(double, KIND_RALPHA, false);
    else if (p.kind == KIND_USER) NPME_DNL_D(double, KIND_USER, false, 4);
    else if (p.kind == KIND_USER_CX) NPME_DNL_D(double, KIND_USER_CX, false, 4);
    else NPME_DNL(double, KIND_HELMHOLTZ, false);
  } else {
    if (p.kind == KIND_LAPLACE) NPME_DNL(float, KIND_LAPLACE, true);
    else if (p.kind == KIND_LAPLACE_ERF) NPME_DNL_D(float, KIND_LAPLACE_ERF, true, 4);
    else if (p.kind == KIND_RALPHA) NPME_DNL(float, KIND_RALPHA, true);
    else if (p.kind == KIND_USER) NPME_DNL_D(float, KIND_USER, true, 4);
    else if (p.kind == KIND_USER_CX) NPME_DNL_D(float, KIND_USER_CX, true, 4);
    else NPME_DNL(float, KIND_HELMHOLTZ, true);
  }
#undef NPME_DNL
#undef NPME_DNL_D
  NPME_HIP_CHECK(hipGetLastError());
}

// number of ordered pairs i != j with r < Rdir (FP64 input coordinates),
// visiting the same cells as the direct sum
__global__ __launch_bounds__(256) void k_count_pairs(DirectDev p, size_t n,
                                                     const int4 *__restrict__ cell,
                                                     const int2 *__restrict__ cellSE,
                                                     const double4 *__restrict__ xs,
                                                     unsigned long long *__restrict__ total)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt = 0;
  if (s < n) {
    const int4 c = cell[s];
    const double xi = xs[s].x, yi = xs[s].y, zi = xs[s].z;
    const double R2 = p.Rdir * p.Rdir;
    const int nN = p.nNeigh;
    for (int dx = -nN; dx <= nN; dx++)
      for (int dy = -nN; dy <= nN; dy++)
        for (int dz = -nN; dz <= nN; dz++) {
          const int gx = c.x + dx, gy = c.y + dy, gz = c.z + dz;
          if (gx < 0 || gy < 0 || gz < 0 || gx >= p.nb0 || gy >= p.nb1 || gz >= p.nb2) continue;
          const int2 se = cellSE[((long)gx * p.nb1 + gy) * p.nb2 + gz];
          for (int j = se.x; j < se.x + se.y; j++) {
            if ((size_t)j == s) continue;
            const double4 o = xs[j];
            const double rx = xi - o.x, ry = yi - o.y, rz = zi - o.z;
            if (rx * rx + ry * ry + rz * rz < R2) cnt++;
          }
        }
  }
  // integer sum: the total does not depend on the order of the additions
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(total, cnt);
}

void launch_count_pairs(const DirectParams &p, size_t n, const int4 *cell, const int *cellStart,
                        const double4 *xs, unsigned long long *total, hipStream_t s)
{
  DirectDev d;
  d.Rdir = p.Rdir;
  d.nNeigh = p.nNeigh;
  d.nb0 = p.nb[0];
  d.nb1 = p.nb[1];
  d.nb2 = p.nb[2];
  hipLaunchKernelGGL(k_count_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, n, cell,
                     (const int2 *)cellStart, xs, total);
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------- LaplaceOrig, user kernels: the reference's pairs
// The reference's erfc direct sum and its GenFunc direct sum (funcSR.Calc)
// have no cutoff of their own: they add every point pair of the cell pairs
// of its interaction list (src/PotentialLaplace.cpp:267-318, :613-668,
// src/PotentialGenFunc.cpp:140-320, :523-645, with the list of
// NPME_FindInteractionList_C, src/PartitionBox.cpp:789-922).  A cell pair
// (c, c') is listed when c' = c, or when c' is in the +-nNeigh stencil of c,
// the two cell boxes are closer than 1.0001 Rdir (from the cells' integer
// offsets: sum_d max(0, |dn_d| - 1)^2 <= nNeigh^2) and some point pair of
// the two cells is closer than Rdir (squared distance below 1.0001 Rdir^2,
// Rdir = nNeigh * cell edge).  k_cellpair_contact sets one bit per (cell,
// stencil offset) for the listed pairs; k_direct_refpairs sums, per target,
// every source of the flagged cells in FP64.
__global__ __launch_bounds__(256) void k_cellpair_contact(DirectDev p, long ncell, int W, double R2,
                                                          const int2 *__restrict__ cellSE,
                                                          const double4 *__restrict__ xs,
                                                          unsigned *__restrict__ flags)
{
  const int nN = p.nNeigh, S = 2 * nN + 1, nOff = S * S * S;
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= ncell * nOff) return;
  const long c = id / nOff;
  const int o = (int)(id - c * nOff);
  const int2 se = cellSE[c];
  if (se.y == 0) return;
  const int cz = (int)(c % p.nb2), cy = (int)((c / p.nb2) % p.nb1), cx = (int)(c / ((long)p.nb1 * p.nb2));
  const int dx = o / (S * S) - nN, dy = (o / S) % S - nN, dz = o % S - nN;
  const int gx = cx + dx, gy = cy + dy, gz = cz + dz;
  if (gx < 0 || gy < 0 || gz < 0 || gx >= p.nb0 || gy >= p.nb1 || gz >= p.nb2) return;
  const int ex = max(abs(dx) - 1, 0), ey = max(abs(dy) - 1, 0), ez = max(abs(dz) - 1, 0);
  if (ex * ex + ey * ey + ez * ez > nN * nN) return;
  const long c2 = ((long)gx * p.nb1 + gy) * p.nb2 + gz;
  const int2 se2 = cellSE[c2];
  if (se2.y == 0) return;
  bool on = c2 == c;
  for (int i = se.x; i < se.x + se.y && !on; i++) {
    const double4 a = xs[i];
    for (int j = se2.x; j < se2.x + se2.y; j++) {
      const double4 b = xs[j];
      const double rx = a.x - b.x, ry = a.y - b.y, rz = a.z - b.z;
      if (rx * rx + ry * ry + rz * rz < R2) {
        on = true;
        break;
      }
    }
  }
  if (on) atomicOr(&flags[c * W + (o >> 5)], 1u << (o & 31));
}

template <int KIND>
__global__ __launch_bounds__(256) void k_direct_refpairs(DirectDev p, size_t n, int W,
                                                         const int4 *__restrict__ cell,
                                                         const int2 *__restrict__ cellSE,
                                                         const double4 *__restrict__ xs,
                                                         const int *__restrict__ perm,
                                                         const double *__restrict__ q,
                                                         const unsigned *__restrict__ flags,
                                                         double *__restrict__ Vdir)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int nN = p.nNeigh, S = 2 * nN + 1, nOff = S * S * S;
  const int4 c = cell[s];
  const long cl = ((long)c.x * p.nb1 + c.y) * p.nb2 + c.z;
  const double4 me = xs[s];
  constexpr bool CX = kind_is_complex(KIND);
  double part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int o = 0; o < nOff; o++) {
    if (!((flags[cl * W + (o >> 5)] >> (o & 31)) & 1u)) continue;
    const int gx = c.x + o / (S * S) - nN, gy = c.y + (o / S) % S - nN, gz = c.z + o % S - nN;
    const int2 se = cellSE[((long)gx * p.nb1 + gy) * p.nb2 + gz];
    for (int j = se.x; j < se.x + se.y; j++) {
      if ((size_t)j == s) continue;
      const double4 b = xs[j];
      const double rx = me.x - b.x, ry = me.y - b.y, rz = me.z - b.z;
      const size_t oj = (size_t)perm[j];
      const double qr = CX ? q[2 * oj] : q[oj], qim = CX ? q[2 * oj + 1] : 0.0;
      pair_sr<double, KIND, 8>(p, rx, ry, rz, rx * rx + ry * ry + rz * rz, qr, qim, part, 0.0, 0.0);
    }
  }
#pragma unroll
  for (int k = 0; k < (CX ? 8 : 4); k++) Vdir[(CX ? 8 : 4) * s + k] = part[k];
}

static DirectDev refpair_dev(const DirectParams &pp)
{
  DirectDev d{};
  d.kind = pp.kind;
  d.nNeigh = pp.nNeigh;
  d.Rdir = pp.Rdir;
  d.beta = pp.beta;
  d.nb0 = pp.nb[0];
  d.nb1 = pp.nb[1];
  d.nb2 = pp.nb[2];
  // user kernels: the FP64 short-range table over (0, utR]
  d.utab = pp.utab;
  d.utS = pp.utS;
  d.invRdir = pp.utR > 0 ? 1.0 / pp.utR : 0.0;
  return d;
}

int refpair_words(int nNeigh)
{
  const int S = 2 * nNeigh + 1;
  return (S * S * S + 31) / 32;
}

void launch_cellpair_contact(const DirectParams &pp, const int *cellStart, const double4 *xs,
                             unsigned *flags, hipStream_t s)
{
  const DirectDev d = refpair_dev(pp);
  const long ncell = (long)pp.nb[0] * pp.nb[1] * pp.nb[2];
  const int S = 2 * pp.nNeigh + 1, W = refpair_words(pp.nNeigh);
  // the reference's test: Rdir recomputed as nNeigh cell edges, squared
  // distance below 1.0001 Rdir^2 (src/PartitionBox.cpp:767-787, :841-842)
  const double cs = pp.Rdir / pp.nNeigh, Rr = pp.nNeigh * cs, R2 = Rr * Rr * 1.0001;
  NPME_HIP_CHECK(hipMemsetAsync(flags, 0, (size_t)ncell * W * sizeof(unsigned), s));
  const long tot = ncell * S * S * S;
  hipLaunchKernelGGL(k_cellpair_contact, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, d, ncell, W,
                     R2, (const int2 *)cellStart, xs, flags);
  NPME_HIP_CHECK(hipGetLastError());
}

void launch_direct_refpairs(const DirectParams &pp, size_t n, const int4 *cell, const int *cellStart,
                            const double4 *xs, const int *perm, const double *q,
                            const unsigned *flags, double *Vdir, hipStream_t s)
{
  const DirectDev d = refpair_dev(pp);
  const dim3 g((unsigned)((n + 255) / 256));
  const int W = refpair_words(pp.nNeigh);
  if (pp.kind == KIND_USER)
    hipLaunchKernelGGL(k_direct_refpairs<KIND_USER>, g, dim3(256), 0, s, d, n, W, cell,
                       (const int2 *)cellStart, xs, perm, q, flags, Vdir);
  else if (pp.kind == KIND_USER_CX)
    hipLaunchKernelGGL(k_direct_refpairs<KIND_USER_CX>, g, dim3(256), 0, s, d, n, W, cell,
                       (const int2 *)cellStart, xs, perm, q, flags, Vdir);
  else if (pp.kind == KIND_LAPLACE_ERF)
    hipLaunchKernelGGL(k_direct_refpairs<KIND_LAPLACE_ERF>, g, dim3(256), 0, s, d, n, W, cell,
                       (const int2 *)cellStart, xs, perm, q, flags, Vdir);
  else
    throw Error{NPME_ERR_UNSUPPORTED, "no reference-pair direct sum for this kernel"};
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ B-splines
// M_n(t + j), j = 0..n-1 (Cox-de Boor over all shifts), n <= 16.
// The recursion always runs in FP64 and the weights are rounded once to the
// output type.  Run in FP32, its rounding accumulates with the order (the
// weights' sum drifts from 1 by several ulp per level), and that alone gave
// the FP32 form relative errors in V of 5e-7 at order 4 up to 2.7e-6 at
// order 14; rounded once, the weights contribute ~5e-8 at every order.
// FP64 VALU operations issue at the FP32 rate on this chip (no packed FP32
// here), so the cost is registers, not issue slots.
__device__ __forceinline__ void bsp_rec(int n, double t, double *w)
{
  if (n == 1) {
    w[0] = 1.0;
    return;
  }
  w[0] = t;
  w[1] = 1.0 - t;
  for (int k = 3; k <= n; k++) {
    const double inv = 1.0 / (double)(k - 1);
    double prev = 0.0;  // M_{k-1}(t + j - 1) for j = 0
    for (int j = 0; j < k; j++) {
      const double left = (j < k - 1) ? w[j] : 0.0;
      const double val = ((t + j) * left + ((double)k - t - j) * prev) * inv;
      prev = left;
      w[j] = val;
    }
  }
}

template <typename Real> __device__ __forceinline__ void bsp_shifts(int n, double t, Real *m)
{
  double w[16];
  bsp_rec(n, t, w);
  for (int j = 0; j < n; j++) m[j] = (Real)w[j];
}

// one level of the unnormalised recursion, order k - 1 -> k, in place:
// U_k(t + j) = (t + j) U_{k-1}(t + j) + (k - t - j) U_{k-1}(t + j - 1), with
// U_k = (k - 1)! M_k; the end terms j = 0 and j = k - 1 have one factor
template <int K> __device__ __forceinline__ void bsp_level(double t, double *w)
{
  const double top = (1.0 - t) * w[K - 2];
#pragma unroll
  for (int j = K - 2; j >= 1; j--) w[j] = (t + j) * w[j] + ((double)(K - j) - t) * w[j - 1];
  w[0] *= t;
  w[K - 1] = top;
}
template <int N> struct Fact {
  static constexpr double v = N * Fact<N - 1>::v;
};
template <> struct Fact<0> {
  static constexpr double v = 1.0;
};

// the recursion with the order known at compile time (fully unrolled,
// registers); w[0..N-1] = M_N(t + j), scaled by 1 / (N - 1)! once
template <int N> __device__ __forceinline__ void bsp_rec_c(double t, double *w)
{
  if constexpr (N == 1) {
    w[0] = 1.0;
  } else {
    w[0] = t;
    w[1] = 1.0 - t;
    if constexpr (N >= 3) bsp_level<3>(t, w);
    if constexpr (N >= 4) bsp_level<4>(t, w);
    if constexpr (N >= 5) bsp_level<5>(t, w);
    if constexpr (N >= 6) bsp_level<6>(t, w);
    if constexpr (N >= 7) bsp_level<7>(t, w);
    if constexpr (N >= 8) bsp_level<8>(t, w);
    if constexpr (N >= 9) bsp_level<9>(t, w);
    if constexpr (N >= 10) bsp_level<10>(t, w);
    if constexpr (N >= 11) bsp_level<11>(t, w);
    if constexpr (N >= 12) bsp_level<12>(t, w);
    if constexpr (N >= 13) bsp_level<13>(t, w);
    if constexpr (N >= 14) bsp_level<14>(t, w);
    if constexpr (N >= 15) bsp_level<15>(t, w);
    if constexpr (N >= 16) bsp_level<16>(t, w);
    static_assert(N <= 16, "spline orders up to 16");
    if constexpr (N >= 3) {
      constexpr double sc = 1.0 / Fact<N - 1>::v;
#pragma unroll
      for (int j = 0; j < N; j++) w[j] *= sc;
    }
  }
}

template <typename Real, int N> __device__ __forceinline__ void bsp_shifts_c(double t, Real *m)
{
  double w[N];
  bsp_rec_c<N>(t, w);
#pragma unroll
  for (int j = 0; j < N; j++) m[j] = (Real)w[j];
}

// weights M_N(t + j) and gradient weights dM_N(t + j)/dt =
// M_{N-1}(t + j) - M_{N-1}(t + j - 1) from one FP64 recursion (the order
// N - 1 stage is the last but one level of the order N recursion), each
// rounded once to the output type
template <typename Real, int N>
__device__ __forceinline__ void bsp_weights_grad_c(double t, Real *B, Real *dB)
{
  double w[N];
  bsp_rec_c<N - 1>(t, w);
#pragma unroll
  for (int j = 0; j < N; j++)
    dB[j] = (Real)(((j < N - 1) ? w[j] : 0.0) - ((j >= 1) ? w[j - 1] : 0.0));
  {
    // one more level of the recursion: order N - 1 -> N (the end terms
    // have one factor)
    constexpr double inv = 1.0 / (double)(N - 1);
    const double top = (1.0 - t) * w[N - 2];
#pragma unroll
    for (int j = N - 2; j >= 1; j--) w[j] = ((t + j) * w[j] + ((double)(N - j) - t) * w[j - 1]) * inv;
    w[0] *= t * inv;
    w[N - 1] = top * inv;
  }
#pragma unroll
  for (int j = 0; j < N; j++) B[j] = (Real)w[j];
}

// the same for an order known only at run time (n <= 16)
template <typename Real>
__device__ __forceinline__ void bsp_weights_grad(int n, double t, Real *B, Real *dB)
{
  double w[16];
  bsp_rec(n - 1, t, w);
  for (int j = 0; j < n; j++)
    dB[j] = (Real)(((j < n - 1) ? w[j] : 0.0) - ((j >= 1) ? w[j - 1] : 0.0));
  const double inv = 1.0 / (double)(n - 1);
  double prev = 0.0;
  for (int j = 0; j < n; j++) {
    const double left = (j < n - 1) ? w[j] : 0.0;
    w[j] = ((t + j) * left + ((double)n - t - j) * prev) * inv;
    prev = left;
  }
  for (int j = 0; j < n; j++) B[j] = (Real)w[j];
}

// ------------------------------------------------------------------ spread
// tile = 8 x 8 x 8 grid points, 64 lanes; lane (i, j) owns the 8-point
// column (u1 + i, u2 + j, u3 .. u3+7).  Charges whose spline window meets the
// tile are listed per tile (set-up); each chunk of 64 charges stages its
// per-axis weights restricted to the tile in LDS, then every lane adds
// q w1[i] w2[j] w3[k] into its column.
constexpr int TS = 8;

template <typename Real, bool CX, int BN>
__global__ __launch_bounds__(64) void k_spread(int BnRun, long M1, long M2, long M3, int nt1, int nt2,
                                               int nt3, const int *__restrict__ tileStart,
                                               const int *__restrict__ tileList,
                                               const Real *__restrict__ qre,
                                               const Real *__restrict__ qim,
                                               const int4 *__restrict__ ms,
                                               const typename V4t<Real>::T *__restrict__ frac,
                                               Real *__restrict__ grid)
{
  // per staged charge: w1[8], w2[8], w3[8] restricted to the tile, q, q_im
  __shared__ __align__(16) Real w[64][3 * TS];
  __shared__ Real qs[64][2];
  const int Bn = BN > 0 ? BN : BnRun;
  constexpr int BA = BN > 0 ? BN : 16;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int t3 = t % nt3, t2 = (t / nt3) % nt2, t1 = t / (nt3 * nt2);
  const int u1 = t1 * TS, u2 = t2 * TS, u3 = t3 * TS;
  const int li = threadIdx.x / TS, lj = threadIdx.x % TS;
  double acc[2 * TS];
#pragma unroll
  for (int k = 0; k < 2 * TS; k++) acc[k] = 0;
  const int beg = tileStart[t], end = tileStart[t + 1];
  for (int base = beg; base < end; base += 64) {
    const int cnt = min(64, end - base);
    if ((int)threadIdx.x < cnt) {
      const int s = tileList[base + threadIdx.x];
      const int4 m0 = ms[s];
      const typename V4t<Real>::T f = frac[s];
      Real sh[BA];
      const int mstart[3] = {m0.x, m0.y, m0.z};
      const Real fr[3] = {f.x, f.y, f.z};
      const int ub[3] = {u1, u2, u3};
#pragma unroll
      for (int d = 0; d < 3; d++) {
        if constexpr (BN > 0) bsp_shifts_c<Real, BN>(fr[d], sh);
        else bsp_shifts<Real>(Bn, fr[d], sh);
#pragma unroll
        for (int k = 0; k < TS; k++) {
          const int idx = ub[d] + k - mstart[d];
          Real v = (Real)0;
#pragma unroll
          for (int j = 0; j < BA; j++)
            if (j == idx && j < Bn) v = sh[j];
          w[threadIdx.x][d * TS + k] = v;
        }
      }
      qs[threadIdx.x][0] = qre[s];
      qs[threadIdx.x][1] = CX ? qim[s] : (Real)0;
    }
    __syncthreads();
    // FP32 (or FP64) sums over the chunk, carried into FP64 per chunk
    Real part[2 * TS];
#pragma unroll
    for (int k = 0; k < 2 * TS; k++) part[k] = 0;
    for (int c = 0; c < cnt; c++) {
      const Real w12 = w[c][li] * w[c][TS + lj];
      Real w3[TS];
#pragma unroll
      for (int k = 0; k < TS; k++) w3[k] = w[c][2 * TS + k];
      const Real cr = qs[c][0] * w12;
      if (CX) {
        const Real ci = qs[c][1] * w12;
#pragma unroll
        for (int k = 0; k < TS; k++) {
          part[2 * k] += cr * w3[k];
          part[2 * k + 1] += ci * w3[k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < TS; k++) part[k] += cr * w3[k];
      }
    }
#pragma unroll
    for (int k = 0; k < (CX ? 2 * TS : TS); k++) acc[k] += (double)part[k];
    __syncthreads();
  }
  const long p1 = u1 + li, p2 = u2 + lj;
  if (p1 < M1 && p2 < M2) {
    const long row = (p1 * M2 + p2) * M3;
#pragma unroll
    for (int k = 0; k < TS; k++) {
      const long p3 = u3 + k;
      if (p3 < M3) {
        if (CX) {
          grid[2 * (row + p3)] = (Real)acc[2 * k];
          grid[2 * (row + p3) + 1] = (Real)acc[2 * k + 1];
        } else {
          grid[row + p3] = (Real)acc[k];
        }
      }
    }
  }
}

// Matrix-core spreading (FP32, real charges).  A tile of 16 x 8 x 8 grid
// points is one wavefront; over the charges c listed for the tile
//   Q[i][(j, k)] = sum_c (q_c w1_c[i]) (w2_c[j] w3_c[k]),
// a product of a 16 x C and a C x 64 matrix, which v_mfma_f32_16x16x4f32
// computes four charges at a time in four 16-column blocks.  Operand
// layout of that instruction: lane l gives A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[4 (l / 16) + r][l % 16] in register r.
// Block b holds the columns (j, k) = (2b + n / 8, n % 8), n = l % 16.
// Each chunk of 64 charges stages its tile-restricted weights in LDS (one
// row per charge, zero outside the spline window), then runs the products;
// the FP32 block sums are carried into FP64 per chunk.
typedef float npme_f4 __attribute__((ext_vector_type(4)));
template <int BN>
__global__ __launch_bounds__(64) void k_spread_mfma(int BnRun, long M1, long M2, long M3, int nt1,
                                                    int nt2, int nt3,
                                                    const int *__restrict__ tileStart,
                                                    const int *__restrict__ tileList,
                                                    const float *__restrict__ qre,
                                                    const int4 *__restrict__ ms,
                                                    const float4 *__restrict__ frac,
                                                    float *__restrict__ grid)
{
  constexpr int TX = 16, TY = 8, TZ = 8;
  // rows of 16 floats: the 16 lanes of an operand group read one row, the
  // four groups of a step read four consecutive rows (no bank conflicts).
  // The three row blocks are one 8 KB block, cleared per chunk by the
  // wavefront in contiguous 1 KB runs
  __shared__ __align__(16) float stage[64 * (TX + TY + TZ)];
  float(*sA)[TX] = reinterpret_cast<float(*)[TX]>(stage);
  float(*sW2)[TY] = reinterpret_cast<float(*)[TY]>(stage + 64 * TX);
  float(*sW3)[TZ] = reinterpret_cast<float(*)[TZ]>(stage + 64 * (TX + TY));
  const int Bn = BN > 0 ? BN : BnRun;
  constexpr int BA = BN > 0 ? BN : 16;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int t3 = t % nt3, t2 = (t / nt3) % nt2, t1 = t / (nt3 * nt2);
  const int u1 = t1 * TX, u2 = t2 * TY, u3 = t3 * TZ;
  const int l = threadIdx.x, n16 = l & 15, kq = l >> 4;
  double acc[4][4];
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[b][r] = 0;
  const int beg = tileStart[t], end = tileStart[t + 1];
  // charge l of a chunk: index, window start, fractions, q; the next
  // chunk's are loaded while the current chunk's products run
  int4 cm = make_int4(0, 0, 0, 0);
  float4 cf = make_float4(0.f, 0.f, 0.f, 0.f);
  float cq = 0.f;
  auto fetch = [&](int base, int4 &m0, float4 &f, float &q) {
    if (base + l < end) {
      const int c = tileList[base + l];
      m0 = ms[c];
      f = frac[c];
      q = qre[c];
    }
  };
  fetch(beg, cm, cf, cq);
  for (int base = beg; base < end; base += 64) {
    const int cnt = min(64, end - base);
    // row l: zeros, then the spline window of charge l at its tile slots.
    // The w2 row is stored as (w2[0], w2[2], w2[4], w2[6], w2[1], w2[3],
    // w2[5], w2[7]): a lane's four B columns j = 2b + h (h = n16 / 8) are
    // then one 16-byte read
    {
      // the workgroup is this one wavefront, whose LDS accesses execute in
      // program order, so the window writes below follow the clearing
      const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int k = 0; k < (TX + TY + TZ) / 4; k++) reinterpret_cast<float4 *>(stage)[64 * k + l] = z4;
    }
    if (l < cnt) {
      float sh[BA];
      if constexpr (BN > 0) bsp_shifts_c<float, BN>(cf.x, sh);
      else bsp_shifts<float>(Bn, cf.x, sh);
      int off = cm.x - u1;
#pragma unroll
      for (int j = 0; j < BA; j++)
        if (j < Bn && off + j >= 0 && off + j < TX) sA[l][off + j] = cq * sh[j];
      if constexpr (BN > 0) bsp_shifts_c<float, BN>(cf.y, sh);
      else bsp_shifts<float>(Bn, cf.y, sh);
      off = cm.y - u2;
#pragma unroll
      for (int j = 0; j < BA; j++) {
        const int jj = off + j;
        if (j < Bn && jj >= 0 && jj < TY) sW2[l][4 * (jj & 1) + (jj >> 1)] = sh[j];
      }
      if constexpr (BN > 0) bsp_shifts_c<float, BN>(cf.z, sh);
      else bsp_shifts<float>(Bn, cf.z, sh);
      off = cm.z - u3;
#pragma unroll
      for (int j = 0; j < BA; j++)
        if (j < Bn && off + j >= 0 && off + j < TZ) sW3[l][off + j] = sh[j];
    }
    __syncthreads();
    fetch(base + 64, cm, cf, cq);
    npme_f4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = d0, d2 = d0, d3 = d0;
    const int nstep = (cnt + 3) >> 2;  // rows past cnt are zero
#pragma unroll 2
    for (int st = 0; st < nstep; st++) {
      const int c = 4 * st + kq;
      const float av = sA[c][n16];
      const float w3 = sW3[c][n16 & 7];
      const float4 w2 = *reinterpret_cast<const float4 *>(&sW2[c][4 * (n16 >> 3)]);
      d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, w2.x * w3, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, w2.y * w3, d1, 0, 0, 0);
      d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, w2.z * w3, d2, 0, 0, 0);
      d3 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, w2.w * w3, d3, 0, 0, 0);
    }
    const npme_f4 dd[4] = {d0, d1, d2, d3};
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[b][r] += (double)dd[b][r];
    __syncthreads();  // the rows are rewritten by the next chunk
  }
  const long p3 = u3 + (n16 & 7);
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const long p2 = u2 + 2 * b + (n16 >> 3);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const long p1 = u1 + 4 * kq + r;
      if (p1 < M1 && p2 < M2 && p3 < M3) grid[(p1 * M2 + p2) * M3 + p3] = (float)acc[b][r];
    }
  }
}

// FP64 form of k_spread_mfma, on v_mfma_f64_16x16x4f64 (same operand
// layout: lane l gives A[l % 16][l / 16] and B[l / 16][l % 16]; it holds
// D[l / 16 + 4 r][l % 16], checked on the GPU against a host product).  The
// products accumulate in FP64 in the matrix cores, so there is no carry.
typedef double npme_d4 __attribute__((ext_vector_type(4)));
template <int BN>
__global__ __launch_bounds__(64) void k_spread_mfma64(int BnRun, long M1, long M2, long M3, int nt1,
                                                      int nt2, int nt3,
                                                      const int *__restrict__ tileStart,
                                                      const int *__restrict__ tileList,
                                                      const double *__restrict__ qre,
                                                      const int4 *__restrict__ ms,
                                                      const double4 *__restrict__ frac,
                                                      double *__restrict__ grid)
{
  constexpr int TX = 16, TY = 8, TZ = 8;
  // one 16 KB block of rows, cleared per chunk by the wavefront in
  // contiguous runs (as in k_spread_mfma)
  __shared__ __align__(16) double stage[64 * (TX + TY + TZ)];
  double(*sA)[TX] = reinterpret_cast<double(*)[TX]>(stage);
  double(*sW2)[TY] = reinterpret_cast<double(*)[TY]>(stage + 64 * TX);
  double(*sW3)[TZ] = reinterpret_cast<double(*)[TZ]>(stage + 64 * (TX + TY));
  const int Bn = BN > 0 ? BN : BnRun;
  constexpr int BA = BN > 0 ? BN : 16;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int t3 = t % nt3, t2 = (t / nt3) % nt2, t1 = t / (nt3 * nt2);
  const int u1 = t1 * TX, u2 = t2 * TY, u3 = t3 * TZ;
  const int l = threadIdx.x, n16 = l & 15, kq = l >> 4;
  const int beg = tileStart[t], end = tileStart[t + 1];
  int4 cm = make_int4(0, 0, 0, 0);
  double4 cf = make_double4(0, 0, 0, 0);
  double cq = 0;
  auto fetch = [&](int base, int4 &m0, double4 &f, double &q) {
    if (base + l < end) {
      const int c = tileList[base + l];
      m0 = ms[c];
      f = frac[c];
      q = qre[c];
    }
  };
  fetch(beg, cm, cf, cq);
  npme_d4 d0 = {0, 0, 0, 0}, d1 = d0, d2 = d0, d3 = d0;
  for (int base = beg; base < end; base += 64) {
    const int cnt = min(64, end - base);
    // row l as in k_spread_mfma: w2 stored (w2[0], w2[2], w2[4], w2[6],
    // w2[1], w2[3], w2[5], w2[7])
    {
      const double2 z2 = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < (TX + TY + TZ) / 2; k++) reinterpret_cast<double2 *>(stage)[64 * k + l] = z2;
    }
    if (l < cnt) {
      double sh[BA];
      if constexpr (BN > 0) bsp_shifts_c<double, BN>(cf.x, sh);
      else bsp_shifts<double>(Bn, cf.x, sh);
      int off = cm.x - u1;
#pragma unroll
      for (int j = 0; j < BA; j++)
        if (j < Bn && off + j >= 0 && off + j < TX) sA[l][off + j] = cq * sh[j];
      if constexpr (BN > 0) bsp_shifts_c<double, BN>(cf.y, sh);
      else bsp_shifts<double>(Bn, cf.y, sh);
      off = cm.y - u2;
#pragma unroll
      for (int j = 0; j < BA; j++) {
        const int jj = off + j;
        if (j < Bn && jj >= 0 && jj < TY) sW2[l][4 * (jj & 1) + (jj >> 1)] = sh[j];
      }
      if constexpr (BN > 0) bsp_shifts_c<double, BN>(cf.z, sh);
      else bsp_shifts<double>(Bn, cf.z, sh);
      off = cm.z - u3;
#pragma unroll
      for (int j = 0; j < BA; j++)
        if (j < Bn && off + j >= 0 && off + j < TZ) sW3[l][off + j] = sh[j];
    }
    __syncthreads();
    fetch(base + 64, cm, cf, cq);
    const int nstep = (cnt + 3) >> 2;  // rows past cnt are zero
    for (int st = 0; st < nstep; st++) {
      const int c = 4 * st + kq;
      const double av = sA[c][n16];
      const double w3 = sW3[c][n16 & 7];
      const double *w2 = &sW2[c][4 * (n16 >> 3)];
      d0 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, w2[0] * w3, d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, w2[1] * w3, d1, 0, 0, 0);
      d2 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, w2[2] * w3, d2, 0, 0, 0);
      d3 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, w2[3] * w3, d3, 0, 0, 0);
    }
    __syncthreads();  // the rows are rewritten by the next chunk
  }
  const npme_d4 dd[4] = {d0, d1, d2, d3};
  const long p3 = u3 + (n16 & 7);
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const long p2 = u2 + 2 * b + (n16 >> 3);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const long p1 = u1 + kq + 4 * r;
      if (p1 < M1 && p2 < M2 && p3 < M3) grid[(p1 * M2 + p2) * M3 + p3] = dd[b][r];
    }
  }
}

void launch_spread(bool fp64, bool complexK, long Bn, const long M[3], const int T[3],
                   const int nt[3], const int *tileStart, const int *tileList, const void *qre,
                   const void *qim, const int4 *ms, const void *frac, void *grid, hipStream_t s)
{
  const dim3 g((unsigned)(nt[0] * nt[1] * nt[2]));
  if (!fp64 && !complexK && T[0] == 16 && T[1] == 8 && T[2] == 8) {
#define NPME_SPREAD_M(BNV)                                                                      \
  hipLaunchKernelGGL((k_spread_mfma<BNV>), g, dim3(64), 0, s, (int)Bn, M[0], M[1], M[2], nt[0],  \
                     nt[1], nt[2], tileStart, tileList, (const float *)qre, ms,                 \
                     (const float4 *)frac, (float *)grid)
    switch (Bn) {
      case 4: NPME_SPREAD_M(4); break;
      case 6: NPME_SPREAD_M(6); break;
      case 8: NPME_SPREAD_M(8); break;
      case 10: NPME_SPREAD_M(10); break;
      default: NPME_SPREAD_M(0); break;
    }
#undef NPME_SPREAD_M
    NPME_HIP_CHECK(hipGetLastError());
    return;
  }
  if (fp64 && !complexK && T[0] == 16 && T[1] == 8 && T[2] == 8) {
#define NPME_SPREAD_M64(BNV)                                                                    \
  hipLaunchKernelGGL((k_spread_mfma64<BNV>), g, dim3(64), 0, s, (int)Bn, M[0], M[1], M[2], nt[0],\
                     nt[1], nt[2], tileStart, tileList, (const double *)qre, ms,                \
                     (const double4 *)frac, (double *)grid)
    switch (Bn) {
      case 4: NPME_SPREAD_M64(4); break;
      case 6: NPME_SPREAD_M64(6); break;
      case 8: NPME_SPREAD_M64(8); break;
      case 10: NPME_SPREAD_M64(10); break;
      default: NPME_SPREAD_M64(0); break;
    }
#undef NPME_SPREAD_M64
    NPME_HIP_CHECK(hipGetLastError());
    return;
  }
  if (T[0] != 8 || T[1] != 8 || T[2] != 8) throw Error{NPME_ERR_UNSUPPORTED, "spreading tile shape"};
#define NPME_SPREAD_B(R, CXV, BNV)                                                              \
  hipLaunchKernelGGL((k_spread<R, CXV, BNV>), g, dim3(64), 0, s, (int)Bn, M[0], M[1], M[2],     \
                     nt[0], nt[1], nt[2], tileStart, tileList, (const R *)qre, (const R *)qim,  \
                     ms, (const typename V4t<R>::T *)frac, (R *)grid)
#define NPME_SPREAD(R, CXV)                                                                     \
  do {                                                                                          \
    switch (Bn) {                                                                               \
      case 4: NPME_SPREAD_B(R, CXV, 4); break;                                                  \
      case 6: NPME_SPREAD_B(R, CXV, 6); break;                                                  \
      case 8: NPME_SPREAD_B(R, CXV, 8); break;                                                  \
      case 10: NPME_SPREAD_B(R, CXV, 10); break;                                                \
      default: NPME_SPREAD_B(R, CXV, 0); break;                                                 \
    }                                                                                           \
  } while (0)
  if (fp64) {
    if (complexK) NPME_SPREAD(double, true);
    else NPME_SPREAD(double, false);
  } else {
    if (complexK) NPME_SPREAD(float, true);
    else NPME_SPREAD(float, false);
  }
#undef NPME_SPREAD
#undef NPME_SPREAD_B
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ interp
// B-spline interpolation of theta, one workgroup per T^3 tile of the compact
// grid: the tile plus its (Bn-1)-point spline halo is staged in LDS once,
// then every charge whose spline window starts in the tile (listed per tile
// at set-up) is interpolated from LDS, combined with its direct sum and
// self term, and written to its input position in FP64.
template <typename Real, bool CX, int BN>
__global__ __launch_bounds__(256) void k_interp_tile(int BnRun, int T, long M1, long M2, long M3,
                                                     int nt1, int nt2, int nt3, double C1,
                                                     double C2, double C3,
                                                     const int *__restrict__ tileStart,
                                                     const int *__restrict__ tileList,
                                                     const int4 *__restrict__ ms,
                                                     const typename V4t<Real>::T *__restrict__ frac,
                                                     const Real *__restrict__ grid,
                                                     const double *__restrict__ Vdir,
                                                     const int *__restrict__ perm,
                                                     const double *__restrict__ q, double fsr,
                                                     double fsi, double *__restrict__ Vout)
{
  extern __shared__ char smem_raw[];
  Real *th = reinterpret_cast<Real *>(smem_raw);
  // BN > 0: spline order fixed at compile time (arrays stay in registers);
  // BN == 0: generic order
  const int Bn = BN > 0 ? BN : BnRun;
  constexpr int BA = BN > 0 ? BN : 16;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int beg = tileStart[t], end = tileStart[t + 1];
  if (beg == end) return;  // uniform per workgroup
  const int t3 = t % nt3, t2 = (t / nt3) % nt2, t1 = t / (nt3 * nt2);
  const int u1 = t1 * T, u2 = t2 * T, u3 = t3 * T;
  const int E = T + Bn - 1;
  const int e1 = (int)min((long)E, M1 - u1), e2 = (int)min((long)E, M2 - u2),
            e3 = (int)min((long)E, M3 - u3);
  constexpr int W = CX ? 2 : 1;
  {
    // the e1 x e2 x e3 block of theta, copied from global memory straight
    // into LDS (global_load_lds: lane l of a wavefront writes dword l of the
    // 64-dword run the wavefront names, no registers in between), so every
    // load of the block is in flight at once.  Element (i, j, k) sits at
    // element ((i e2 + j) e3 + k) of the block, D dwords each; dword
    // d = b + threadIdx.x of the copy is part d % D of element d / D, and as
    // d advances by the block size each thread's part stays the same while
    // its element advances by nb / D, with carries instead of divisions
    constexpr int D = (int)sizeof(Real) * W / 4;
    const int nb = blockDim.x, tot = e1 * e2 * e3, totw = tot * D, step = nb / D;
    const int dk = step % e3, dj = (step / e3) % e2, di = step / (e3 * e2);
    const int part = threadIdx.x % D;
    int e = threadIdx.x / D, k = e % e3, j = (e / e3) % e2, i = e / (e3 * e2);
    const unsigned *g0 = reinterpret_cast<const unsigned *>(grid + W * (((long)u1 * M2 + u2) * M3 + u3));
    unsigned *thw = reinterpret_cast<unsigned *>(th) + (threadIdx.x & ~63);
    for (int b = 0; b < totw; b += nb) {  // uniform per workgroup
      if (b + (int)threadIdx.x < totw) {
        const unsigned *src = g0 + (long)D * (((long)i * M2 + j) * M3 + k) + part;
        __builtin_amdgcn_global_load_lds(
            (const void __attribute__((address_space(1))) *)src,
            (void __attribute__((address_space(3))) *)(thw + b), 4, 0, 0);
      }
      k += dk;
      j += dj;
      if (k >= e3) {
        k -= e3;
        j++;
      }
      i += di;
      if (j >= e2) {
        j -= e2;
        i++;
      }
    }
    __builtin_amdgcn_s_waitcnt(0);  // the copies have landed in LDS
  }
  __syncthreads();
  for (int c = beg + threadIdx.x; c < end; c += blockDim.x) {
    const int s = tileList[c];
    const int4 m0 = ms[s];
    const typename V4t<Real>::T f = frac[s];
    // the combine's inputs are loaded before the interpolation, so their
    // latency overlaps it
    const int o = perm[s];
    double vd[W * 4], qv[W];
#pragma unroll
    for (int k = 0; k < W * 4; k++) vd[k] = Vdir[W * 4 * (size_t)s + k];
#pragma unroll
    for (int k = 0; k < W; k++) qv[k] = q[W * (size_t)o + k];
    Real B[3][BA], dB[3][BA];
    const Real fr[3] = {f.x, f.y, f.z};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if constexpr (BN > 0) bsp_weights_grad_c<Real, BN>(fr[d], B[d], dB[d]);
      else bsp_weights_grad<Real>(Bn, fr[d], B[d], dB[d]);
    }
    const int l1 = m0.x - u1, l2 = m0.y - u2, l3 = m0.z - u3;
    // per x-plane j1: sums over the (j2, j3) window in the working
    // precision (36 terms), carried into FP64 once per plane
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j1 = 0; j1 < BA; j1++) {
      if (j1 >= Bn) break;
      Real p0 = 0, py = 0, pz = 0, p0i = 0, pyi = 0, pzi = 0;
#pragma unroll
      for (int j2 = 0; j2 < BA; j2++) {
        if (j2 >= Bn) break;
        const Real *row = th + W * (((l1 + j1) * e2 + (l2 + j2)) * e3 + l3);
        Real tb = 0, tdb = 0, tbi = 0, tdbi = 0;
#pragma unroll
        for (int j3 = 0; j3 < BA; j3++) {
          if (j3 >= Bn) break;
          const Real tv = row[W * j3];
          tb += tv * B[2][j3];
          tdb += tv * dB[2][j3];
          if (CX) {
            const Real ti = row[W * j3 + 1];
            tbi += ti * B[2][j3];
            tdbi += ti * dB[2][j3];
          }
        }
        p0 += B[1][j2] * tb;
        py += dB[1][j2] * tb;
        pz += B[1][j2] * tdb;
        if (CX) {
          p0i += B[1][j2] * tbi;
          pyi += dB[1][j2] * tbi;
          pzi += B[1][j2] * tdbi;
        }
      }
      const double b0 = (double)B[0][j1], d0 = (double)dB[0][j1];
      v[0] += b0 * (double)p0;
      v[2] += d0 * (double)p0;
      v[4] += b0 * (double)py;
      v[6] += b0 * (double)pz;
      if (CX) {
        v[1] += b0 * (double)p0i;
        v[3] += d0 * (double)p0i;
        v[5] += b0 * (double)pyi;
        v[7] += b0 * (double)pzi;
      }
    }
    if (CX) {
      const double qr = qv[0], qim = qv[W - 1];
      const double Cc[4] = {1.0, C1, C2, C3};
      double out[8];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        out[2 * k] = v[2 * k] * Cc[k] + vd[2 * k];
        out[2 * k + 1] = v[2 * k + 1] * Cc[k] + vd[(2 * k + 1) % (W * 4)];
      }
      out[0] -= fsr * qr - fsi * qim;
      out[1] -= fsr * qim + fsi * qr;
#pragma unroll
      for (int k = 0; k < 8; k++) Vout[8 * (size_t)o + k] = out[k];
    } else {
      Vout[4 * (size_t)o + 0] = v[0] + vd[0] - fsr * qv[0];
      Vout[4 * (size_t)o + 1] = v[2] * C1 + vd[1 % (W * 4)];
      Vout[4 * (size_t)o + 2] = v[4] * C2 + vd[2 % (W * 4)];
      Vout[4 * (size_t)o + 3] = v[6] * C3 + vd[3 % (W * 4)];
    }
  }
}

void launch_interp(bool fp64, bool complexK, long Bn, const long M[3], const double C[3], int T,
                   const int nt[3], const int *tileStart, const int *tileList, const int4 *ms,
                   const void *frac, const void *grid, const double *Vdir, const int *perm,
                   const double *q_in, const double fself[2], double *Vout, hipStream_t s)
{
  const dim3 g((unsigned)(nt[0] * nt[1] * nt[2]));
  const int E = T + (int)Bn - 1;
  const size_t sh = (size_t)E * E * E * (fp64 ? 8 : 4) * (complexK ? 2 : 1);
#define NPME_INTERP_B(R, CXV, BNV)                                                              \
  hipLaunchKernelGGL((k_interp_tile<R, CXV, BNV>), g, dim3(256), sh, s, (int)Bn, T, M[0], M[1],   \
                     M[2], nt[0], nt[1], nt[2], C[0], C[1], C[2], tileStart, tileList, ms,       \
                     (const typename V4t<R>::T *)frac, (const R *)grid, Vdir, perm, q_in,        \
                     fself[0], fself[1], Vout)
#define NPME_INTERP(R, CXV)                                                                      \
  do {                                                                                           \
    switch (Bn) {                                                                                \
      case 4: NPME_INTERP_B(R, CXV, 4); break;                                                   \
      case 6: NPME_INTERP_B(R, CXV, 6); break;                                                   \
      case 8: NPME_INTERP_B(R, CXV, 8); break;                                                   \
      case 10: NPME_INTERP_B(R, CXV, 10); break;                                                 \
      default: NPME_INTERP_B(R, CXV, 0); break;                                                  \
    }                                                                                            \
  } while (0)
  if (fp64) {
    if (complexK) NPME_INTERP(double, true);
    else NPME_INTERP(double, false);
  } else {
    if (complexK) NPME_INTERP(float, true);
    else NPME_INTERP(float, false);
  }
#undef NPME_INTERP
#undef NPME_INTERP_B
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ exact
// V_t = sum_{j != idx[t]} f(r) q_j in FP64, sources streamed through LDS.
// Coordinates and charges in input order.
template <int KIND>
__global__ __launch_bounds__(256) void k_exact(double alpha, double k0r, double k0i, size_t n,
                                               const double *__restrict__ coord,
                                               const double *__restrict__ q, size_t nt,
                                               const long *__restrict__ idx,
                                               double *__restrict__ Vout)
{
  constexpr bool CX = kind_is_complex(KIND);
  __shared__ double sx[256], sy[256], sz[256], sq[256], sqi[256];
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = t < nt;
  const long i = active ? (idx ? idx[t] : (long)t) : 0;
  const double xi = active ? coord[3 * i] : 0, yi = active ? coord[3 * i + 1] : 0,
               zi = active ? coord[3 * i + 2] : 0;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t j0 = 0; j0 < n; j0 += 256) {
    const size_t j = j0 + threadIdx.x;
    if (j < n) {
      sx[threadIdx.x] = coord[3 * j];
      sy[threadIdx.x] = coord[3 * j + 1];
      sz[threadIdx.x] = coord[3 * j + 2];
      sq[threadIdx.x] = CX ? q[2 * j] : q[j];
      sqi[threadIdx.x] = CX ? q[2 * j + 1] : 0.0;
    }
    __syncthreads();
    const int cnt = (int)min((size_t)256, n - j0);
    if (active) {
      for (int k = 0; k < cnt; k++) {
        if ((long)(j0 + k) == i) continue;
        const double rx = xi - sx[k], ry = yi - sy[k], rz = zi - sz[k];
        const double r2 = rx * rx + ry * ry + rz * rz;
        if (KIND == KIND_LAPLACE) {
          const double ri = 1.0 / sqrt(r2);
          const double f1 = -ri * ri * ri * sq[k];
          v[0] += ri * sq[k];
          v[1] += rx * f1;
          v[2] += ry * f1;
          v[3] += rz * f1;
        } else if (KIND == KIND_RALPHA) {
          const double fe = pow(r2, 0.5 * alpha);
          const double f1 = alpha * fe / r2 * sq[k];
          v[0] += fe * sq[k];
          v[1] += rx * f1;
          v[2] += ry * f1;
          v[3] += rz * f1;
        } else {
          const double r = sqrt(r2), ri = 1.0 / r;
          double sn, cn;
          sincos(k0r * r, &sn, &cn);
          const double dmp = exp(-k0i * r) * ri;
          const double fr = dmp * cn, fi = dmp * sn;
          const double ikr = -k0i - ri, iki = k0r;
          const double f1r = (ikr * fr - iki * fi) * ri, f1i = (ikr * fi + iki * fr) * ri;
          const double qr = sq[k], qim = sqi[k];
          const double vr = fr * qr - fi * qim, vi = fr * qim + fi * qr;
          const double gr = f1r * qr - f1i * qim, gi = f1r * qim + f1i * qr;
          v[0] += vr;
          v[1] += vi;
          v[2] += rx * gr;
          v[3] += rx * gi;
          v[4] += ry * gr;
          v[5] += ry * gi;
          v[6] += rz * gr;
          v[7] += rz * gi;
        }
      }
    }
    __syncthreads();
  }
  if (active) {
    if (CX)
      for (int k = 0; k < 8; k++) Vout[8 * t + k] = v[k];
    else
      for (int k = 0; k < 4; k++) Vout[4 * t + k] = v[k];
  }
}

void launch_exact(int kind, double alpha, double k0r, double k0i, size_t n,
                  const double *coordS, const int *perm, const double *q_in, size_t nt,
                  const long *idx, double *Vout, hipStream_t s)
{
  (void)perm;
  const dim3 g((unsigned)((nt + 255) / 256));
  if (kind == KIND_LAPLACE || kind == KIND_LAPLACE_ERF)  // the full kernel 1/r
    hipLaunchKernelGGL(k_exact<KIND_LAPLACE>, g, dim3(256), 0, s, alpha, k0r, k0i, n, coordS,
                       q_in, nt, idx, Vout);
  else if (kind == KIND_RALPHA)
    hipLaunchKernelGGL(k_exact<KIND_RALPHA>, g, dim3(256), 0, s, alpha, k0r, k0i, n, coordS,
                       q_in, nt, idx, Vout);
  else
    hipLaunchKernelGGL(k_exact<KIND_HELMHOLTZ>, g, dim3(256), 0, s, alpha, k0r, k0i, n, coordS,
                       q_in, nt, idx, Vout);
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ F grid
struct LRDev {
  int kind, Nder;
  double Rdir, alpha, k0r, k0i, beta;
  double a[62], b[62];
};

template <int KIND>
__device__ __forceinline__ void lr_eval(const LRDev &k, double r2, double &fr, double &fi)
{
  const double r = sqrt(fabs(r2));
  fi = 0;
  if (KIND == KIND_LAPLACE_ERF) {
    // erf(beta r)/r, 2 beta/sqrt(pi) (1 - x^2/3) below x = beta r = 1e-8
    // (src/KernelFunctionLaplace.cpp:1087-1116)
    const double x = k.beta * r;
    fr = x < 1e-8 ? k.beta * 1.12837916709551257390 * (1.0 - x * x / 3.0) : erf(x) / r;
    return;
  }
  const bool inside = (KIND == KIND_RALPHA) ? !(r2 > k.Rdir * k.Rdir) : !(r > k.Rdir);
  if (inside) {
    double sr = k.a[2 * k.Nder];
    if (KIND == KIND_HELMHOLTZ) {
      double si = k.a[2 * k.Nder + 1];
      for (int i = k.Nder - 1; i >= 0; i--) {
        sr = sr * r2 + k.a[2 * i];
        si = si * r2 + k.a[2 * i + 1];
      }
      fi = si;
    } else {
      for (int i = k.Nder - 1; i >= 0; i--) sr = sr * r2 + k.a[2 * i];
    }
    fr = sr;
  } else if (KIND == KIND_LAPLACE) {
    fr = 1.0 / r;
  } else if (KIND == KIND_RALPHA) {
    fr = pow(r2, 0.5 * k.alpha);
  } else {
    double sn, cn;
    sincos(k.k0r * r, &sn, &cn);
    const double d = exp(-k.k0i * r) / r;
    fr = d * cn;
    fi = d * sn;
  }
}

// tab layout per axis d: [x1 | x2 | T1 | T2] each Mf[d] long, axes consecutive
template <int KIND>
__global__ __launch_bounds__(256) void k_fgrid(LRDev k, long Mf0, long Mf1, long Mf2,
                                               const double *__restrict__ tab,
                                               void *__restrict__ Fout)
{
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = Mf0 * Mf1 * Mf2;
  if (id >= tot) return;
  const long i3 = id % Mf2, i2 = (id / Mf2) % Mf1, i1 = id / (Mf2 * Mf1);
  const double *t0 = tab, *t1 = tab + 4 * Mf0, *t2 = t1 + 4 * Mf1;
  const double X[2] = {t0[i1], t0[Mf0 + i1]}, TX[2] = {t0[2 * Mf0 + i1], t0[3 * Mf0 + i1]};
  const double Y[2] = {t1[i2], t1[Mf1 + i2]}, TY[2] = {t1[2 * Mf1 + i2], t1[3 * Mf1 + i2]};
  const double Z[2] = {t2[i3], t2[Mf2 + i3]}, TZ[2] = {t2[2 * Mf2 + i3], t2[3 * Mf2 + i3]};
  double sr = 0, si = 0;
#pragma unroll
  for (int ax = 0; ax < 2; ax++)
#pragma unroll
    for (int ay = 0; ay < 2; ay++)
#pragma unroll
      for (int az = 0; az < 2; az++) {
        // outside the extension band of an axis its second weight is 0
        // (with a finite dummy argument), so the term adds exactly zero
        const double w = TX[ax] * TY[ay] * TZ[az];
        if (w == 0.0) continue;
        double fr, fi;
        lr_eval<KIND>(k, X[ax] * X[ax] + Y[ay] * Y[ay] + Z[az] * Z[az], fr, fi);
        sr += w * fr;
        si += w * fi;
      }
  if (KIND == KIND_HELMHOLTZ) {
    double2 o;
    o.x = sr;
    o.y = si;
    static_cast<double2 *>(Fout)[id] = o;
  } else {
    static_cast<double *>(Fout)[id] = sr;
  }
}

void launch_fgrid_real_space(const DMKernel &km, const long Mf[3], const double *d_tab, void *F,
                             hipStream_t s)
{
  LRDev k;
  k.kind = km.kind;
  k.Nder = km.Nder;
  k.Rdir = km.Rdir;
  k.alpha = km.alpha;
  k.k0r = km.k0r;
  k.k0i = km.k0i;
  k.beta = km.beta;
  for (int i = 0; i < 62; i++) {
    k.a[i] = km.a[i];
    k.b[i] = km.b[i];
  }
  const long tot = Mf[0] * Mf[1] * Mf[2];
  const dim3 g((unsigned)((tot + 255) / 256));
  if (k.kind == KIND_LAPLACE)
    hipLaunchKernelGGL(k_fgrid<KIND_LAPLACE>, g, dim3(256), 0, s, k, Mf[0], Mf[1], Mf[2], d_tab, F);
  else if (k.kind == KIND_LAPLACE_ERF)
    hipLaunchKernelGGL(k_fgrid<KIND_LAPLACE_ERF>, g, dim3(256), 0, s, k, Mf[0], Mf[1], Mf[2], d_tab, F);
  else if (k.kind == KIND_RALPHA)
    hipLaunchKernelGGL(k_fgrid<KIND_RALPHA>, g, dim3(256), 0, s, k, Mf[0], Mf[1], Mf[2], d_tab, F);
  else
    hipLaunchKernelGGL(k_fgrid<KIND_HELMHOLTZ>, g, dim3(256), 0, s, k, Mf[0], Mf[1], Mf[2], d_tab, F);
  NPME_HIP_CHECK(hipGetLastError());
}

// G = F lam1 lam2 lam3 / (N1 N2 N3) on the (N1/2+1)(N2/2+1)(N3/2+1) grid,
// stored in blocks of 8 along axis 3 (see k_make_G)
template <typename Real, bool CX>
__global__ void k_make_G(long Mf0, long Mf1, long Mf2, const double *__restrict__ F,
                         const double *__restrict__ lam, double scale, Real *__restrict__ G)
{
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = Mf0 * Mf1 * Mf2;
  if (id >= tot) return;
  const long i3 = id % Mf2, i2 = (id / Mf2) % Mf1, i1 = id / (Mf2 * Mf1);
  const double l = lam[i1] * lam[Mf0 + i2] * lam[Mf0 + Mf1 + i3] * scale;
  // G[i2][i3 / 8][i1][i3 % 8]: the x-pass workgroup of 8 consecutive i3
  // reads one contiguous block of (N1/2+1) x 8 values
  const long nb3 = (Mf2 + 7) / 8;
  const long o = ((i2 * nb3 + (i3 >> 3)) * Mf0 + i1) * 8 + (i3 & 7);
  if (CX) {  // F: double2 per point
    G[2 * o] = (Real)(F[2 * id] * l);
    G[2 * o + 1] = (Real)(F[2 * id + 1] * l);
  } else {   // F: double per point
    G[o] = (Real)(F[id] * l);
  }
}

void launch_make_G(bool fp64, bool complexK, const long Mf[3], long K3, const double *F,
                   const double *lam, void *G, hipStream_t s)
{
  (void)K3;
  const long tot = Mf[0] * Mf[1] * Mf[2];
  // N_d = 2 (Mf_d - 1)
  const double scale = 1.0 / ((double)(2 * (Mf[0] - 1)) * (double)(2 * (Mf[1] - 1)) *
                              (double)(2 * (Mf[2] - 1)));
  const dim3 g((unsigned)((tot + 255) / 256));
#define NPME_G(R, CXV)                                                                          \
  hipLaunchKernelGGL((k_make_G<R, CXV>), g, dim3(256), 0, s, Mf[0], Mf[1], Mf[2], F, lam,       \
                     scale, (R *)G)
  if (fp64) {
    if (complexK) NPME_G(double, true);
    else NPME_G(double, false);
  } else {
    if (complexK) NPME_G(float, true);
    else NPME_G(float, false);
  }
#undef NPME_G
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------ translated kernel grid
// The full kernel (not its long-range part) at r2 = |r|^2
template <int KIND>
__device__ __forceinline__ void full_eval(const LRDev &k, double r2, double &fr, double &fi)
{
  const double r = sqrt(r2);
  fi = 0;
  if (KIND == KIND_RALPHA) {
    fr = pow(r2, 0.5 * k.alpha);
  } else if (KIND == KIND_HELMHOLTZ) {
    double sn, cn;
    sincos(k.k0r * r, &sn, &cn);
    const double d = exp(-k.k0i * r) / r;
    fr = d * cn;
    fi = d * sn;
  } else {  // KIND_LAPLACE, KIND_LAPLACE_ERF
    fr = 1.0 / r;
  }
}

// f(r) = func(|r + D|) on the full N0 x N1 x N2 grid as the sum of the 8
// overlapped images with their top-hat weights; tab per axis d holds
// [x1 + D | x2 + D | T1 | T2] of N_d values (axes consecutive).  Images of
// zero weight are skipped.  Output (re, im) per point.
template <int KIND>
__global__ __launch_bounds__(256) void k_fgrid_trans(LRDev k, long N0, long N1, long N2,
                                                     const double *__restrict__ tab,
                                                     double2 *__restrict__ Fout)
{
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= N0 * N1 * N2) return;
  const long i3 = id % N2, i2 = (id / N2) % N1, i1 = id / (N2 * N1);
  const double *t0 = tab, *t1 = tab + 4 * N0, *t2 = t1 + 4 * N1;
  const double X[2] = {t0[i1], t0[N0 + i1]}, TX[2] = {t0[2 * N0 + i1], t0[3 * N0 + i1]};
  const double Y[2] = {t1[i2], t1[N1 + i2]}, TY[2] = {t1[2 * N1 + i2], t1[3 * N1 + i2]};
  const double Z[2] = {t2[i3], t2[N2 + i3]}, TZ[2] = {t2[2 * N2 + i3], t2[3 * N2 + i3]};
  double sr = 0, si = 0;
#pragma unroll
  for (int ax = 0; ax < 2; ax++)
#pragma unroll
    for (int ay = 0; ay < 2; ay++)
#pragma unroll
      for (int az = 0; az < 2; az++) {
        const double w = TX[ax] * TY[ay] * TZ[az];
        if (w == 0.0) continue;
        double fr, fi;
        full_eval<KIND>(k, X[ax] * X[ax] + Y[ay] * Y[ay] + Z[az] * Z[az], fr, fi);
        sr += w * fr;
        si += w * fi;
      }
  Fout[id] = make_double2(sr, si);
}

void launch_fgrid_trans(const DMKernel &km, const long N[3], const double *d_tab, double2 *F,
                        hipStream_t s)
{
  LRDev k{};
  k.kind = km.kind;
  k.alpha = km.alpha;
  k.k0r = km.k0r;
  k.k0i = km.k0i;
  const long tot = N[0] * N[1] * N[2];
  const dim3 g((unsigned)((tot + 255) / 256));
  if (k.kind == KIND_RALPHA)
    hipLaunchKernelGGL(k_fgrid_trans<KIND_RALPHA>, g, dim3(256), 0, s, k, N[0], N[1], N[2], d_tab, F);
  else if (k.kind == KIND_HELMHOLTZ)
    hipLaunchKernelGGL(k_fgrid_trans<KIND_HELMHOLTZ>, g, dim3(256), 0, s, k, N[0], N[1], N[2], d_tab, F);
  else if (k.kind == KIND_LAPLACE || k.kind == KIND_LAPLACE_ERF)
    hipLaunchKernelGGL(k_fgrid_trans<KIND_LAPLACE>, g, dim3(256), 0, s, k, N[0], N[1], N[2], d_tab, F);
  else
    throw Error{NPME_ERR_UNSUPPORTED, "no device form of this kernel's translated grid"};
  NPME_HIP_CHECK(hipGetLastError());
}

// G = F lam1 lam2 lam3 scale over N0 x N1 x G3 (a3 < G3 <= N2 of F), F
// complex [N0][N1][N2]; lam holds the full |lambda|^2 arrays of the three
// axes (N0, N1, N2 values); layout G[a2][a3 / 8][a1][a3 % 8]
template <typename Real>
__global__ void k_make_G_full(long N0, long N1, long N2, long G3, const double2 *__restrict__ F,
                              const double *__restrict__ lam, double scale, Real *__restrict__ G)
{
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= N0 * N1 * G3) return;
  const long a3 = id % G3, a2 = (id / G3) % N1, a1 = id / (G3 * N1);
  const double l = lam[a1] * lam[N0 + a2] * lam[N0 + N1 + a3] * scale;
  const double2 f = F[(a1 * N1 + a2) * N2 + a3];
  const long o = ((a2 * ((G3 + 7) / 8) + (a3 >> 3)) * N0 + a1) * 8 + (a3 & 7);
  G[2 * o] = (Real)(f.x * l);
  G[2 * o + 1] = (Real)(f.y * l);
}

void launch_make_G_full(bool fp64, const long N[3], long G3, const double2 *F, const double *lam,
                        void *G, hipStream_t s)
{
  const long tot = N[0] * N[1] * G3;
  const double scale = 1.0 / ((double)N[0] * (double)N[1] * (double)N[2]);
  const dim3 g((unsigned)((tot + 255) / 256));
  if (fp64)
    hipLaunchKernelGGL(k_make_G_full<double>, g, dim3(256), 0, s, N[0], N[1], N[2], G3, F, lam, scale,
                       (double *)G);
  else
    hipLaunchKernelGGL(k_make_G_full<float>, g, dim3(256), 0, s, N[0], N[1], N[2], G3, F, lam, scale,
                       (float *)G);
  NPME_HIP_CHECK(hipGetLastError());
}

}  // namespace npme
