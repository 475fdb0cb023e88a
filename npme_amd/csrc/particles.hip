// Particle-side kernels of the NPME CalcV pipeline on gfx950:
//   gather   charges into cell-sorted order
//   direct   short-range DM sum over the cell list (one target per lane,
//            per-target cell pruning, FP32 pair math with FP64 per-cell
//            accumulation in the mixed form)
//   spread   B-spline charge spreading, one 8x8x8 grid tile per wavefront,
//            the tile accumulated in registers (no atomics; every grid point
//            sums its charges in list order, so results are deterministic)
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
  double a[62], b[62];
  int nb0, nb1, nb2;
};

// ------------------------------------------------------------------ gather
template <typename Real, bool CX>
__global__ void k_gather(size_t n, const int *__restrict__ perm, const double *__restrict__ q,
                         typename V4t<Real>::T *__restrict__ src, Real *__restrict__ qi)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int o = perm[s];
  if (CX) {
    src[s].w = (Real)q[2 * (size_t)o];
    qi[s] = (Real)q[2 * (size_t)o + 1];
  } else {
    src[s].w = (Real)q[o];
  }
}

void launch_gather(bool fp64, bool complexK, size_t n, const int *perm, const double *q_in,
                   void *src, void *qi, hipStream_t s)
{
  const dim3 grid((unsigned)((n + 255) / 256));
  if (fp64) {
    if (complexK)
      hipLaunchKernelGGL((k_gather<double, true>), grid, dim3(256), 0, s, n, perm, q_in,
                         (double4 *)src, (double *)qi);
    else
      hipLaunchKernelGGL((k_gather<double, false>), grid, dim3(256), 0, s, n, perm, q_in,
                         (double4 *)src, (double *)qi);
  } else {
    if (complexK)
      hipLaunchKernelGGL((k_gather<float, true>), grid, dim3(256), 0, s, n, perm, q_in,
                         (float4 *)src, (float *)qi);
    else
      hipLaunchKernelGGL((k_gather<float, false>), grid, dim3(256), 0, s, n, perm, q_in,
                         (float4 *)src, (float *)qi);
  }
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ direct
template <typename Real> __device__ __forceinline__ Real rsqrt_r(Real x);
__device__ __forceinline__ void sincos_r(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ void sincos_r(double x, double *s, double *c) { sincos(x, s, c); }
template <> __device__ __forceinline__ float rsqrt_r<float>(float x) { return rsqrtf(x); }
template <> __device__ __forceinline__ double rsqrt_r<double>(double x) { return 1.0 / sqrt(x); }

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

// one target per lane; cells of the reference's cell grid (size Rdir/nNeigh)
// within +-nNeigh are visited unless their box is farther than Rdir from the
// target.  Coordinates are local to each cell, so the FP32 differences stay
// accurate for any box size.
template <typename Real, int KIND>
__global__ __launch_bounds__(256) void k_direct(DirectDev p, size_t n,
                                                const typename V4t<Real>::T *__restrict__ src,
                                                const Real *__restrict__ qi,
                                                const int4 *__restrict__ cell,
                                                const int2 *__restrict__ cellSE,
                                                double *__restrict__ Vdir)
{
  using V4 = typename V4t<Real>::T;
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  constexpr bool CX = (KIND == KIND_HELMHOLTZ);
  const V4 me = src[s];
  const int4 c = cell[s];
  const Real cs = (Real)p.cs;
  const Real R2 = (Real)(p.Rdir * p.Rdir);
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int nN = p.nNeigh;
  for (int dx = -nN; dx <= nN; dx++) {
    const int gx = c.x + dx;
    if (gx < 0 || gx >= p.nb0) continue;
    const Real ox = (Real)(dx * p.cs);
    const Real ex = fmax(fmax(ox - me.x, me.x - (ox + cs)), (Real)0);
    for (int dy = -nN; dy <= nN; dy++) {
      const int gy = c.y + dy;
      if (gy < 0 || gy >= p.nb1) continue;
      const Real oy = (Real)(dy * p.cs);
      const Real ey = fmax(fmax(oy - me.y, me.y - (oy + cs)), (Real)0);
      if (ex * ex + ey * ey > R2) continue;
      for (int dz = -nN; dz <= nN; dz++) {
        const int gz = c.z + dz;
        if (gz < 0 || gz >= p.nb2) continue;
        const Real oz = (Real)(dz * p.cs);
        const Real ez = fmax(fmax(oz - me.z, me.z - (oz + cs)), (Real)0);
        if (ex * ex + ey * ey + ez * ez > R2) continue;
        const int2 se = cellSE[((long)gx * p.nb1 + gy) * p.nb2 + gz];
        Real part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = se.x; j < se.x + se.y; j++) {
          if ((size_t)j == s) continue;
          const V4 o = src[j];
          const Real rx = me.x - o.x - ox;
          const Real ry = me.y - o.y - oy;
          const Real rz = me.z - o.z - oz;
          const Real r2 = rx * rx + ry * ry + rz * rz;
          if (!(r2 < R2)) continue;
          if constexpr (KIND == KIND_LAPLACE) {
            const Real ri = rsqrt_r<Real>(r2);
            const Real f0 = ri - horner_re<Real>(p.a, p.Nder, r2);
            const Real f1 = -ri * ri * ri - horner_re<Real>(p.b, p.Nder - 1, r2);
            const Real q = o.w, fq = f1 * q;
            part[0] += f0 * q;
            part[1] += rx * fq;
            part[2] += ry * fq;
            part[3] += rz * fq;
          } else if constexpr (KIND == KIND_RALPHA) {
            const Real fe = pow(r2, (Real)(0.5 * p.alpha));
            const Real f0 = fe - horner_re<Real>(p.a, p.Nder, r2);
            const Real f1 = (Real)p.alpha * fe / r2 - horner_re<Real>(p.b, p.Nder - 1, r2);
            const Real q = o.w, fq = f1 * q;
            part[0] += f0 * q;
            part[1] += rx * fq;
            part[2] += ry * fq;
            part[3] += rz * fq;
          } else {
            // exp(i k0 r)/r with the phase k0r r reduced modulo 2 pi in FP64
            const Real r = sqrt(r2);
            const Real ri = (Real)1 / r;
            const double ph = p.k0r * (double)r;
            const double red = ph - 6.283185307179586477 * rint(ph * 0.15915494309189533577);
            Real sn, cn;
            sincos_r((Real)red, &sn, &cn);
            const Real damp = exp((Real)(-p.k0i) * r) * ri;
            const Real fr = damp * cn, fi = damp * sn;
            // f1 = (i k0 f - f/r)/r
            const Real k0r = (Real)p.k0r, k0i = (Real)p.k0i;
            const Real ikr = -k0i - ri, iki = k0r;  // i k0 - 1/r
            const Real f1r = (ikr * fr - iki * fi) * ri, f1i = (ikr * fi + iki * fr) * ri;
            Real pr, pi, br, bi;
            horner_cx<Real>(p.a, p.Nder, r2, pr, pi);
            horner_cx<Real>(p.b, p.Nder - 1, r2, br, bi);
            const Real s0r = fr - pr, s0i = fi - pi, s1r = f1r - br, s1i = f1i - bi;
            const Real qr = o.w, qim = qi[j];
            const Real vr = s0r * qr - s0i * qim, vi = s0r * qim + s0i * qr;
            const Real gr = s1r * qr - s1i * qim, gi = s1r * qim + s1i * qr;
            part[0] += vr;
            part[1] += vi;
            part[2] += rx * gr;
            part[3] += rx * gi;
            part[4] += ry * gr;
            part[5] += ry * gi;
            part[6] += rz * gr;
            part[7] += rz * gi;
          }
        }
#pragma unroll
        for (int k = 0; k < (CX ? 8 : 4); k++) acc[k] += (double)part[k];
      }
    }
  }
  if (CX) {
#pragma unroll
    for (int k = 0; k < 8; k++) Vdir[8 * s + k] = acc[k];
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) Vdir[4 * s + k] = acc[k];
  }
}

void launch_direct(bool fp64, const DirectParams &p, size_t n, const void *src, const void *qi,
                   const int4 *cell, const int *cellStart, double *Vdir, hipStream_t s)
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
  for (int i = 0; i < 62; i++) {
    d.a[i] = p.a[i];
    d.b[i] = p.b[i];
  }
  d.nb0 = p.nb[0];
  d.nb1 = p.nb[1];
  d.nb2 = p.nb[2];
  const dim3 grid((unsigned)((n + 255) / 256));
  const int2 *se = (const int2 *)cellStart;
#define NPME_DIRECT(R, K)                                                                  \
  hipLaunchKernelGGL((k_direct<R, K>), grid, dim3(256), 0, s, d, n,                        \
                     (const typename V4t<R>::T *)src, (const R *)qi, cell, se, Vdir)
  if (fp64) {
    if (p.kind == KIND_LAPLACE) NPME_DIRECT(double, KIND_LAPLACE);
    else if (p.kind == KIND_RALPHA) NPME_DIRECT(double, KIND_RALPHA);
    else NPME_DIRECT(double, KIND_HELMHOLTZ);
  } else {
    if (p.kind == KIND_LAPLACE) NPME_DIRECT(float, KIND_LAPLACE);
    else if (p.kind == KIND_RALPHA) NPME_DIRECT(float, KIND_RALPHA);
    else NPME_DIRECT(float, KIND_HELMHOLTZ);
  }
#undef NPME_DIRECT
  NPME_HIP_CHECK(hipGetLastError());
}

// number of ordered pairs i != j with r < Rdir (FP64 input coordinates),
// visiting the same cells as the direct sum
__global__ __launch_bounds__(256) void k_count_pairs(DirectDev p, size_t n,
                                                     const int4 *__restrict__ cell,
                                                     const int2 *__restrict__ cellSE,
                                                     const int *__restrict__ perm,
                                                     const double *__restrict__ coord,
                                                     unsigned long long *__restrict__ total)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt = 0;
  if (s < n) {
    const int4 c = cell[s];
    const long i = perm[s];
    const double xi = coord[3 * i], yi = coord[3 * i + 1], zi = coord[3 * i + 2];
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
            const long o = perm[j];
            const double rx = xi - coord[3 * o], ry = yi - coord[3 * o + 1], rz = zi - coord[3 * o + 2];
            if (rx * rx + ry * ry + rz * rz < R2) cnt++;
          }
        }
  }
  // integer sum: the total does not depend on the order of the additions
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(total, cnt);
}

void launch_count_pairs(const DirectParams &p, size_t n, const int4 *cell, const int *cellStart,
                        const int *perm, const double *coord, unsigned long long *total,
                        hipStream_t s)
{
  DirectDev d;
  d.Rdir = p.Rdir;
  d.nNeigh = p.nNeigh;
  d.nb0 = p.nb[0];
  d.nb1 = p.nb[1];
  d.nb2 = p.nb[2];
  hipLaunchKernelGGL(k_count_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, n, cell,
                     (const int2 *)cellStart, perm, coord, total);
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ B-splines
// M_n(t + j), j = 0..n-1 (Cox-de Boor over all shifts), n <= 16
template <typename Real> __device__ __forceinline__ void bsp_shifts(int n, Real t, Real *m)
{
  if (n == 1) {
    m[0] = (Real)1;
    return;
  }
  m[0] = t;
  m[1] = (Real)1 - t;
  for (int k = 3; k <= n; k++) {
    const Real inv = (Real)1 / (Real)(k - 1);
    Real prev = (Real)0;  // M_{k-1}(t + j - 1) for j = 0
    for (int j = 0; j < k; j++) {
      const Real left = (j < k - 1) ? m[j] : (Real)0;
      const Real val = ((t + j) * left + ((Real)k - t - j) * prev) * inv;
      prev = left;
      m[j] = val;
    }
  }
}

// ------------------------------------------------------------------ spread
// tile = 8 x 8 x 8 grid points, 64 lanes; lane (i, j) owns the 8-point
// column (u1 + i, u2 + j, u3 .. u3+7).  Charges whose spline window meets the
// tile are listed per tile (set-up); each chunk of 64 charges stages its
// per-axis weights restricted to the tile in LDS, then every lane adds
// q w1[i] w2[j] w3[k] into its column.
constexpr int TS = 8;

template <typename Real, bool CX>
__global__ __launch_bounds__(64) void k_spread(int Bn, long M1, long M2, long M3, int nt1, int nt2,
                                               int nt3, const int *__restrict__ tileStart,
                                               const int *__restrict__ tileList,
                                               const typename V4t<Real>::T *__restrict__ src,
                                               const Real *__restrict__ qi,
                                               const int4 *__restrict__ ms,
                                               const typename V4t<Real>::T *__restrict__ frac,
                                               Real *__restrict__ grid)
{
  __shared__ Real w[64][3 * TS];
  __shared__ Real qs[64][2];
  const int t = blockIdx.x;
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
      Real sh[16];
      const int mstart[3] = {m0.x, m0.y, m0.z};
      const Real fr[3] = {f.x, f.y, f.z};
      const int ub[3] = {u1, u2, u3};
      for (int d = 0; d < 3; d++) {
        bsp_shifts<Real>(Bn, fr[d], sh);
        for (int k = 0; k < TS; k++) {
          const int idx = ub[d] + k - mstart[d];
          w[threadIdx.x][d * TS + k] = (idx >= 0 && idx < Bn) ? sh[idx] : (Real)0;
        }
      }
      qs[threadIdx.x][0] = src[s].w;
      qs[threadIdx.x][1] = CX ? qi[s] : (Real)0;
    }
    __syncthreads();
    for (int c = 0; c < cnt; c++) {
      const Real w12 = w[c][li] * w[c][TS + lj];
      const Real cr = qs[c][0] * w12;
      if (CX) {
        const Real ci = qs[c][1] * w12;
#pragma unroll
        for (int k = 0; k < TS; k++) {
          acc[2 * k] += (double)(cr * w[c][2 * TS + k]);
          acc[2 * k + 1] += (double)(ci * w[c][2 * TS + k]);
        }
      } else {
#pragma unroll
        for (int k = 0; k < TS; k++) acc[k] += (double)(cr * w[c][2 * TS + k]);
      }
    }
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

void launch_spread(bool fp64, bool complexK, long Bn, const long M[3], const int T[3],
                   const int nt[3], const int *tileStart, const int *tileList, const void *src,
                   const void *qi, const int4 *ms, const void *frac, void *grid, hipStream_t s)
{
  (void)T;
  const dim3 g((unsigned)(nt[0] * nt[1] * nt[2]));
#define NPME_SPREAD(R, CXV)                                                                     \
  hipLaunchKernelGGL((k_spread<R, CXV>), g, dim3(64), 0, s, (int)Bn, M[0], M[1], M[2], nt[0],   \
                     nt[1], nt[2], tileStart, tileList, (const typename V4t<R>::T *)src,       \
                     (const R *)qi, ms, (const typename V4t<R>::T *)frac, (R *)grid)
  if (fp64) {
    if (complexK) NPME_SPREAD(double, true);
    else NPME_SPREAD(double, false);
  } else {
    if (complexK) NPME_SPREAD(float, true);
    else NPME_SPREAD(float, false);
  }
#undef NPME_SPREAD
  NPME_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ interp
template <typename Real, bool CX>
__global__ __launch_bounds__(256) void k_interp(int Bn, long M1, long M2, long M3, double C1,
                                                double C2, double C3, size_t n,
                                                const int4 *__restrict__ ms,
                                                const typename V4t<Real>::T *__restrict__ frac,
                                                const Real *__restrict__ grid,
                                                const double *__restrict__ Vdir,
                                                const int *__restrict__ perm,
                                                const double *__restrict__ q, double fsr,
                                                double fsi, double *__restrict__ Vout)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int4 m0 = ms[s];
  const typename V4t<Real>::T f = frac[s];
  Real B[3][16], dB[3][16], sh[16];
  const Real fr[3] = {f.x, f.y, f.z};
  for (int d = 0; d < 3; d++) {
    bsp_shifts<Real>(Bn, fr[d], B[d]);
    bsp_shifts<Real>(Bn - 1, fr[d], sh);
    for (int j = 0; j < Bn; j++) dB[d][j] = ((j < Bn - 1) ? sh[j] : (Real)0) - ((j >= 1) ? sh[j - 1] : (Real)0);
  }
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int j1 = 0; j1 < Bn; j1++) {
    for (int j2 = 0; j2 < Bn; j2++) {
      const long base = (((long)m0.x + j1) * M2 + (m0.y + j2)) * M3 + m0.z;
      Real tb = 0, tdb = 0, tbi = 0, tdbi = 0;
      for (int j3 = 0; j3 < Bn; j3++) {
        if (CX) {
          const Real tr = grid[2 * (base + j3)], ti = grid[2 * (base + j3) + 1];
          tb += tr * B[2][j3];
          tdb += tr * dB[2][j3];
          tbi += ti * B[2][j3];
          tdbi += ti * dB[2][j3];
        } else {
          const Real tv = grid[base + j3];
          tb += tv * B[2][j3];
          tdb += tv * dB[2][j3];
        }
      }
      const double A1 = (double)(B[0][j1] * B[1][j2]);
      const double A2 = (double)(dB[0][j1] * B[1][j2]);
      const double A3 = (double)(B[0][j1] * dB[1][j2]);
      v[0] += A1 * tb;
      v[2] += A2 * tb;
      v[4] += A3 * tb;
      v[6] += A1 * tdb;
      if (CX) {
        v[1] += A1 * tbi;
        v[3] += A2 * tbi;
        v[5] += A3 * tbi;
        v[7] += A1 * tdbi;
      }
    }
  }
  const int o = perm[s];
  if (CX) {
    const double qr = q[2 * (size_t)o], qim = q[2 * (size_t)o + 1];
    const double C[4] = {1.0, C1, C2, C3};
    double out[8];
    for (int k = 0; k < 4; k++) {
      out[2 * k] = v[2 * k] * C[k] + Vdir[8 * s + 2 * k];
      out[2 * k + 1] = v[2 * k + 1] * C[k] + Vdir[8 * s + 2 * k + 1];
    }
    out[0] -= fsr * qr - fsi * qim;
    out[1] -= fsr * qim + fsi * qr;
    for (int k = 0; k < 8; k++) Vout[8 * (size_t)o + k] = out[k];
  } else {
    Vout[4 * (size_t)o + 0] = v[0] + Vdir[4 * s + 0] - fsr * q[o];
    Vout[4 * (size_t)o + 1] = v[2] * C1 + Vdir[4 * s + 1];
    Vout[4 * (size_t)o + 2] = v[4] * C2 + Vdir[4 * s + 2];
    Vout[4 * (size_t)o + 3] = v[6] * C3 + Vdir[4 * s + 3];
  }
}

void launch_interp(bool fp64, bool complexK, long Bn, const long M[3], const double C[3],
                   size_t n, const int4 *ms, const void *frac, const void *grid,
                   const double *Vdir, const int *perm, const double *q_in, const double fself[2],
                   double *Vout, hipStream_t s)
{
  const dim3 g((unsigned)((n + 255) / 256));
#define NPME_INTERP(R, CXV)                                                                    \
  hipLaunchKernelGGL((k_interp<R, CXV>), g, dim3(256), 0, s, (int)Bn, M[0], M[1], M[2], C[0],  \
                     C[1], C[2], n, ms, (const typename V4t<R>::T *)frac, (const R *)grid,     \
                     Vdir, perm, q_in, fself[0], fself[1], Vout)
  if (fp64) {
    if (complexK) NPME_INTERP(double, true);
    else NPME_INTERP(double, false);
  } else {
    if (complexK) NPME_INTERP(float, true);
    else NPME_INTERP(float, false);
  }
#undef NPME_INTERP
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
  constexpr bool CX = (KIND == KIND_HELMHOLTZ);
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
  if (kind == KIND_LAPLACE)
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
  double Rdir, alpha, k0r, k0i;
  double a[62], b[62];
};

__device__ void lr_eval(const LRDev &k, double r2, double &fr, double &fi)
{
  const double r = sqrt(fabs(r2));
  const bool inside = (k.kind == KIND_RALPHA) ? !(r2 > k.Rdir * k.Rdir) : !(r > k.Rdir);
  if (inside) {
    double sr = k.a[2 * k.Nder], si = k.a[2 * k.Nder + 1];
    for (int i = k.Nder - 1; i >= 0; i--) {
      sr = sr * r2 + k.a[2 * i];
      si = si * r2 + k.a[2 * i + 1];
    }
    fr = sr;
    fi = si;
  } else if (k.kind == KIND_LAPLACE) {
    fr = 1.0 / r;
    fi = 0;
  } else if (k.kind == KIND_RALPHA) {
    fr = pow(r2, 0.5 * k.alpha);
    fi = 0;
  } else {
    double sn, cn;
    sincos(k.k0r * r, &sn, &cn);
    const double d = exp(-k.k0i * r) / r;
    fr = d * cn;
    fi = d * sn;
  }
}

// tab layout per axis d: [x1 | x2 | T1 | T2] each Mf[d] long, axes consecutive
__global__ void k_fgrid(LRDev k, long Mf0, long Mf1, long Mf2, const double *__restrict__ tab,
                        double2 *__restrict__ F)
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
  for (int ax = 0; ax < 2; ax++)
    for (int ay = 0; ay < 2; ay++)
      for (int az = 0; az < 2; az++) {
        const double w = TX[ax] * TY[ay] * TZ[az];
        double fr, fi;
        lr_eval(k, X[ax] * X[ax] + Y[ay] * Y[ay] + Z[az] * Z[az], fr, fi);
        sr += w * fr;
        si += w * fi;
      }
  double2 o;
  o.x = sr;
  o.y = si;
  F[id] = o;
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
  for (int i = 0; i < 62; i++) {
    k.a[i] = km.a[i];
    k.b[i] = km.b[i];
  }
  const long tot = Mf[0] * Mf[1] * Mf[2];
  hipLaunchKernelGGL(k_fgrid, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, k, Mf[0],
                     Mf[1], Mf[2], d_tab, (double2 *)F);
  NPME_HIP_CHECK(hipGetLastError());
}

// G = F lam1 lam2 lam3 / (N1 N2 N3) on the (N1/2+1)(N2/2+1)(N3/2+1) grid
template <typename Real, bool CX>
__global__ void k_make_G(long Mf0, long Mf1, long Mf2, const double2 *__restrict__ F,
                         const double *__restrict__ lam, double scale, Real *__restrict__ G)
{
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = Mf0 * Mf1 * Mf2;
  if (id >= tot) return;
  const long i3 = id % Mf2, i2 = (id / Mf2) % Mf1, i1 = id / (Mf2 * Mf1);
  const double l = lam[i1] * lam[Mf0 + i2] * lam[Mf0 + Mf1 + i3] * scale;
  if (CX) {
    G[2 * id] = (Real)(F[id].x * l);
    G[2 * id + 1] = (Real)(F[id].y * l);
  } else {
    G[id] = (Real)(F[id].x * l);
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
  hipLaunchKernelGGL((k_make_G<R, CXV>), g, dim3(256), 0, s, Mf[0], Mf[1], Mf[2],              \
                     (const double2 *)F, lam, scale, (R *)G)
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

}  // namespace npme
