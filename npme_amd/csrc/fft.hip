// Batched 1-D Stockham FFT passes in LDS for gfx950.
//
// The reciprocal-sum convolution theta = IDFT(G . DFT(Q)) is computed on the
// zero-padded full grid (N1 x N2 x N3) but only the compact quarter-volume
// input and output (M_d = N_d/2 per axis) are ever stored:
//   A  real rows (M3) -> R2C along axis 3, length N3, K3 = N3/2+1 outputs
//   B  columns along axis 2, M2 inputs -> N2 outputs
//   C  columns along axis 1: M1 inputs, forward N1, multiply by G,
//      backward N1, M1 outputs (in place)
//   D  columns along axis 2, N2 inputs -> backward -> M2 outputs
//   E  C2R rows along axis 3, M3 real outputs
// Backward transforms use conj(DFT(conj(x))), so only forward butterflies
// exist.  The twiddle table holds exp(-2 pi i t/N) computed in FP64 on the
// host.
//
// A workgroup of 256 threads transforms B columns (B a power of two) held in
// LDS; every thread owns one column and a strided subset of its butterflies,
// so all per-column index arithmetic is done once, with shifts.  Strided
// passes keep LDS as [n][c] (B consecutive columns side by side: global rows
// of B complex values, 128 B for B = 16 in FP32, map to contiguous LDS);
// row passes keep [c][n] with one element of padding per column.  Between
// the read and the write barrier of a stage every thread holds its
// butterflies in registers, so one LDS buffer suffices.
#include <type_traits>

#include "npme_internal.h"

namespace npme {

template <typename R> struct Cx;
template <> struct Cx<float> { using T = float2; };
template <> struct Cx<double> { using T = double2; };

template <typename T> __device__ __forceinline__ T mk(decltype(T::x) a, decltype(T::x) b)
{
  T r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T> __device__ __forceinline__ T cadd(T a, T b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ T csub(T a, T b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ T cmul(T a, T b)
{
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ T cconj(T a) { return mk<T>(a.x, -a.y); }
template <typename T> __device__ __forceinline__ T mul_mi(T a) { return mk<T>(a.y, -a.x); }  // * -i

// p[idx] when ok, else zero, with no branch around the load: the index is
// clamped to 0 (p[0] must be readable) and the value selected afterwards.
// A load under a per-lane condition becomes an exec-masked block whose
// result is copied out at once, which waits for that load before the next
// one is issued; with the select, a thread's loads are all in flight together
template <typename T> __device__ __forceinline__ T load_or_zero(const T *p, long idx, bool ok)
{
  const T x = p[ok ? idx : 0];
  return ok ? x : mk<T>(0, 0);
}
__device__ __forceinline__ float load_or_zero(const float *p, long idx, bool ok)
{
  const float x = p[ok ? idx : 0];
  return ok ? x : 0.f;
}
__device__ __forceinline__ double load_or_zero(const double *p, long idx, bool ok)
{
  const double x = p[ok ? idx : 0];
  return ok ? x : 0.0;
}

constexpr int NT = 256;  // threads per workgroup

struct PlanDev {
  int N, nst;
  int radix[24];
};

// ------------------------------------------------------------ small DFTs
// forward: y_k = sum_r x_r exp(-2 pi i r k / R)
template <typename T, int R> struct DFT;

template <typename T> struct DFT<T, 2> {
  __device__ static void run(T *x)
  {
    T a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};

template <typename T> struct DFT<T, 4> {
  __device__ static void run(T *x)
  {
    T a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    T c = cadd(x[1], x[3]), d = mul_mi(csub(x[1], x[3]));
    x[0] = cadd(a, c);
    x[2] = csub(a, c);
    x[1] = cadd(b, d);
    x[3] = csub(b, d);
  }
};

template <typename T> struct DFT<T, 8> {
  __device__ static void run(T *x)
  {
    using S = decltype(T::x);
    const S c = (S)0.70710678118654752440;
    T e[4] = {x[0], x[2], x[4], x[6]};
    T o[4] = {x[1], x[3], x[5], x[7]};
    DFT<T, 4>::run(e);
    DFT<T, 4>::run(o);
    T o1 = mk<T>(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));   // * exp(-i pi/4)
    T o2 = mul_mi(o[2]);                                          // * -i
    T o3 = mk<T>(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));  // * exp(-3i pi/4)
    x[0] = cadd(e[0], o[0]);
    x[4] = csub(e[0], o[0]);
    x[1] = cadd(e[1], o1);
    x[5] = csub(e[1], o1);
    x[2] = cadd(e[2], o2);
    x[6] = csub(e[2], o2);
    x[3] = cadd(e[3], o3);
    x[7] = csub(e[3], o3);
  }
};

template <typename T> struct DFT<T, 3> {
  __device__ static void run(T *x)
  {
    using S = decltype(T::x);
    const S s3 = (S)0.86602540378443864676;
    T t1 = cadd(x[1], x[2]);
    T t2 = mk<T>(x[0].x - (S)0.5 * t1.x, x[0].y - (S)0.5 * t1.y);
    T t3 = mk<T>(s3 * (x[1].x - x[2].x), s3 * (x[1].y - x[2].y));
    x[0] = cadd(x[0], t1);
    x[1] = cadd(t2, mul_mi(t3));
    x[2] = csub(t2, mul_mi(t3));
  }
};

// odd radix R = 5, 7: pairwise symmetric form with constant cos/sin
template <int R> struct OddTab;
template <> struct OddTab<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410,
                                  -0.80901699437494742410, 0.30901699437494742410};
  static constexpr double s[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917,
                                  -0.58778525229247312917, -0.95105651629515357212};
};
template <> struct OddTab<7> {
  static constexpr double c[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429,
                                  -0.90096886790241912624, -0.90096886790241912624,
                                  -0.22252093395631440429, 0.62348980185873353053};
  static constexpr double s[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702,
                                  0.43388373911755812048, -0.43388373911755812048,
                                  -0.97492791218182360702, -0.78183148246802980871};
};

template <typename T, int R> struct DFTOdd {
  __device__ static void run(T *x)
  {
    using S = decltype(T::x);
    constexpr int H = (R - 1) / 2;
    T Sj[H], Dj[H];
#pragma unroll
    for (int j = 1; j <= H; j++) {
      Sj[j - 1] = cadd(x[j], x[R - j]);
      Dj[j - 1] = csub(x[j], x[R - j]);
    }
    T y[R];
    T y0 = x[0];
#pragma unroll
    for (int j = 0; j < H; j++) y0 = cadd(y0, Sj[j]);
    y[0] = y0;
#pragma unroll
    for (int k = 1; k <= H; k++) {
      T A = x[0], B = mk<T>(0, 0);
#pragma unroll
      for (int j = 1; j <= H; j++) {
        const S c = (S)OddTab<R>::c[(j * k) % R], s = (S)OddTab<R>::s[(j * k) % R];
        A.x += c * Sj[j - 1].x;
        A.y += c * Sj[j - 1].y;
        B.x += s * Dj[j - 1].x;
        B.y += s * Dj[j - 1].y;
      }
      y[k] = mk<T>(A.x + B.y, A.y - B.x);      // A - iB
      y[R - k] = mk<T>(A.x - B.y, A.y + B.x);  // A + iB
    }
#pragma unroll
    for (int k = 0; k < R; k++) x[k] = y[k];
  }
};

template <typename T, int R> __device__ __forceinline__ void dft_any(T *x)
{
  if constexpr (R == 5 || R == 7)
    DFTOdd<T, R>::run(x);
  else
    DFT<T, R>::run(x);
}

// ------------------------------------------------------------ column access
// element n of the thread's column: [c][n] (row layout, stride ld = N + 1)
// or [n][c] (strided layout, B columns side by side)
struct ColRef {
  int base, sh;  // lds index = base + n * 2^sh (the step is 1 or the column count)
  __device__ __forceinline__ int at(int n) const { return base + (n << sh); }
};

// ------------------------------------------------------------ Stockham stage
// butterfly j of a column reads x[j + r N/R], twiddles by w^(r k N/(Ns R)),
// k = j mod Ns, and writes y[(j - k) R + k + r Ns].  POW2: N and Ns are
// powers of two, so the modulo is a mask (the generic form divides).
template <typename T, int R, int BPT, bool POW2>
__device__ __forceinline__ void stage(T *lds, ColRef col, int tj, int tpc, int N, int Ns,
                                      const T *__restrict__ tw)
{
  const int nb = N / R;
  const int stride = N / (Ns * R);
  T v[BPT][R];
#pragma unroll
  for (int u = 0; u < BPT; u++) {
    const int j = tj + u * tpc;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; r++) v[u][r] = lds[col.at(j + r * nb)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < BPT; u++) {
    const int j = tj + u * tpc;
    if (j < nb) {
      const int k = (Ns == 1) ? 0 : (POW2 ? (j & (Ns - 1)) : j % Ns);
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < R; r++) v[u][r] = cmul(v[u][r], tw[r * k * stride]);
      }
      dft_any<T, R>(v[u]);
      const int idx = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) lds[col.at(idx + r * Ns)] = v[u][r];
    }
  }
  __syncthreads();
}

// radix classes of a plan: the kernels are compiled for the radices a plan
// can contain, since every stage variant in the switch adds registers
// (radix 8 only: ~80 VGPRs, 8/4/2: ~110, all six: ~170)
enum { RC_ALL = 0, RC_POW2 = 1, RC_R8 = 2 };

// all stages of the plan on every column of the workgroup; EMAX bounds the
// elements one thread holds (N / tpc <= EMAX)
template <typename T, int EMAX, int RC>
__device__ void fft_columns(T *lds, ColRef col, int tj, int tpc, const PlanDev &p,
                            const T *__restrict__ tw)
{
  constexpr bool P2 = RC != RC_ALL;
  int Ns = 1;
  for (int s = 0; s < p.nst; s++) {
    const int R = p.radix[s];
    switch (R) {
      case 8: stage<T, 8, (EMAX + 7) / 8, P2>(lds, col, tj, tpc, p.N, Ns, tw); break;
      case 4:
        if constexpr (RC != RC_R8) stage<T, 4, (EMAX + 3) / 4, P2>(lds, col, tj, tpc, p.N, Ns, tw);
        break;
      case 2:
        if constexpr (RC != RC_R8) stage<T, 2, (EMAX + 1) / 2, P2>(lds, col, tj, tpc, p.N, Ns, tw);
        break;
      case 3:
        if constexpr (RC == RC_ALL) stage<T, 3, (EMAX + 2) / 3, P2>(lds, col, tj, tpc, p.N, Ns, tw);
        break;
      case 5:
        if constexpr (RC == RC_ALL) stage<T, 5, (EMAX + 4) / 5, P2>(lds, col, tj, tpc, p.N, Ns, tw);
        break;
      case 7:
        if constexpr (RC == RC_ALL) stage<T, 7, (EMAX + 6) / 7, P2>(lds, col, tj, tpc, p.N, Ns, tw);
        break;
      default: break;
    }
    Ns *= R;
  }
}

// at most 4096 complex values (32 KB FP32, 64 KB FP64) per workgroup, so
// several workgroups share a CU and overlap their load, LDS and store phases
template <typename Real> constexpr int emax() { return 16; }

// workgroups are dealt round-robin to the 8 XCDs; remap so each XCD walks a
// contiguous range of columns (adjacent column groups share L2 lines)
__device__ __forceinline__ int xcd_wg()
{
  const int wg = blockIdx.x, nwg = gridDim.x;
  const int x = wg & 7, local = wg >> 3, q = nwg >> 3, r = nwg & 7;
  return x * q + min(x, r) + local;
}

// thread -> (column, index within column) for the two layouts
struct Lane {
  int c, tj, tpc;
  ColRef col;
};
__device__ __forceinline__ Lane lane_strided(int lgB, int N)
{
  Lane L;
  const int B = 1 << lgB;
  L.c = threadIdx.x & (B - 1);
  L.tj = threadIdx.x >> lgB;
  L.tpc = NT >> lgB;
  L.col.base = L.c;
  L.col.sh = lgB;
  (void)N;
  return L;
}
__device__ __forceinline__ Lane lane_rows(int lgB, int N)
{
  Lane L;
  L.tpc = NT >> lgB;
  L.c = threadIdx.x >> (8 - lgB);  // NT = 2^8
  L.tj = threadIdx.x & (L.tpc - 1);
  L.col.base = L.c * (N + 1);
  L.col.sh = 0;
  return L;
}

// loads of one column into LDS: all of a thread's (at most EMAX) elements
// are fetched into registers first, so its global loads are in flight
// together, then stored; get(n) returns element n (n < N) with no branch
// around its loads
template <typename T, int EMAX, typename F>
__device__ __forceinline__ void stage_in(T *lds, const Lane &L, int N, F &&get)
{
  T v[EMAX];
  // get runs for every slot (element 0 in place of those past N), so no
  // load sits behind a per-lane branch
#pragma unroll
  for (int u = 0; u < EMAX; u++) {
    const int n = L.tj + u * L.tpc;
    v[u] = get(n < N ? n : 0);
  }
#pragma unroll
  for (int u = 0; u < EMAX; u++) {
    const int n = L.tj + u * L.tpc;
    if (n < N) lds[L.col.at(n)] = v[u];
  }
}

// twiddle table of the pass: copied into LDS behind the column data in FP32
// (data <= 32 KB plus N <= 4096 entries stay within 64 KB; FP64 columns
// already fill 64 KB), so the butterflies read it with ds_read instead of a
// global load per twiddled element; the copy is ordered by the barrier that
// follows stage_in
template <typename Real> constexpr bool twl_lds() { return sizeof(Real) == 4; }

template <typename T, bool TWL>
__device__ __forceinline__ const T *twiddles_in(T *lds, int ldsElems, int N, const T *tw)
{
  if constexpr (TWL) {
    T *tl = lds + ldsElems;
    for (int i = threadIdx.x; i < N; i += NT) tl[i] = tw[i];
    return tl;
  } else {
    return tw;
  }
}

// ------------------------------------------------------------ pass kernels
// generic complex pass.  Column col = o * nInner + i;
// element n at o * outer + i * inner + n * pos (all element counts < 2^31).
struct CGeom {
  int N, Lin, Lout;
  int nCols, nInner, validInner;  // inner index >= validInner: padding column
  int inOuter, inInner, inPos, outOuter, outInner, outPos;
};

template <typename Real, bool ROWS, int RC>
__global__ __launch_bounds__(NT) void k_fft_complex(CGeom g, PlanDev p, const typename Cx<Real>::T *tw,
                                                    const typename Cx<Real>::T *__restrict__ in,
                                                    typename Cx<Real>::T *__restrict__ out, int lgB,
                                                    int backward)
{
  using T = typename Cx<Real>::T;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int N = g.N;
  const Lane L = ROWS ? lane_rows(lgB, N) : lane_strided(lgB, N);
  const int colg = (xcd_wg() << lgB) + L.c;
  const int o0 = colg / g.nInner;
  const bool valid = colg < g.nCols && colg - o0 * g.nInner < g.validInner;
  const int o = valid ? o0 : 0, i = valid ? colg - o0 * g.nInner : 0;
  const T *src = in + ((long)o * g.inOuter + (long)i * g.inInner);
  const T *twl = twiddles_in<T, twl_lds<Real>()>(lds, (1 << lgB) * (ROWS ? N + 1 : N), N, tw);
  stage_in<T, emax<Real>()>(lds, L, N, [&](int n) {
    const T v = load_or_zero(src, (long)n * g.inPos, valid && n < g.Lin);
    return backward ? cconj(v) : v;
  });
  __syncthreads();
  fft_columns<T, emax<Real>(), RC>(lds, L.col, L.tj, L.tpc, p, twl);
  if (!valid) return;
  T *dst = out + ((long)o * g.outOuter + (long)i * g.outInner);
  for (int n = L.tj; n < g.Lout; n += L.tpc) {
    T v = lds[L.col.at(n)];
    if (backward) v = cconj(v);
    dst[(long)n * g.outPos] = v;
  }
}

// pass A: two real rows per complex column, R2C by the two-for-one split
template <typename Real, int RC>
__global__ __launch_bounds__(NT) void k_fft_r2c(int N, int Lin, int nPairs, int inRow, int outRow,
                                                PlanDev p, const typename Cx<Real>::T *tw,
                                                const Real *__restrict__ in,
                                                typename Cx<Real>::T *__restrict__ out, int lgB)
{
  using T = typename Cx<Real>::T;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const Lane L = lane_rows(lgB, N);
  const int pr = (xcd_wg() << lgB) + L.c;
  const bool valid = pr < nPairs;
  const Real *ra = in + (valid ? (long)(2 * pr) * inRow : 0), *rb = ra + (valid ? inRow : 0);
  const T *twl = twiddles_in<T, twl_lds<Real>()>(lds, (1 << lgB) * (N + 1), N, tw);
  stage_in<T, emax<Real>()>(lds, L, N, [&](int n) {
    const bool ok = valid && n < Lin;
    return mk<T>(load_or_zero(ra, n, ok), load_or_zero(rb, n, ok));
  });
  __syncthreads();
  fft_columns<T, emax<Real>(), RC>(lds, L.col, L.tj, L.tpc, p, twl);
  if (!valid) return;
  const int K = N / 2 + 1;
  const Real h = (Real)0.5;
  T *oa = out + (long)(2 * pr) * outRow, *ob = oa + outRow;
  for (int k = L.tj; k < K; k += L.tpc) {
    const T z = lds[L.col.at(k)];
    const T w = cconj(lds[L.col.at(k == 0 ? 0 : N - k)]);
    const T s = cadd(z, w), d = csub(z, w);
    oa[k] = mk<T>(h * s.x, h * s.y);
    ob[k] = mk<T>(h * d.y, -h * d.x);  // -i d / 2
  }
}

// pass E: Hermitian half rows -> two real rows (backward transform)
template <typename Real, int RC>
__global__ __launch_bounds__(NT) void k_fft_c2r(int N, int Lout, int nPairs, int inRow, int outRow,
                                                PlanDev p, const typename Cx<Real>::T *tw,
                                                const typename Cx<Real>::T *__restrict__ in,
                                                Real *__restrict__ out, int lgB)
{
  using T = typename Cx<Real>::T;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const Lane L = lane_rows(lgB, N);
  const int pr = (xcd_wg() << lgB) + L.c;
  const bool valid = pr < nPairs;
  const int h = N / 2;
  const T *ia = in + (valid ? (long)(2 * pr) * inRow : 0), *ib = ia + (valid ? inRow : 0);
  const T *twl = twiddles_in<T, twl_lds<Real>()>(lds, (1 << lgB) * (N + 1), N, tw);
  {
    // each stored element m = 0 .. N/2 of the two Hermitian rows is loaded
    // once and gives both rows n = m and n = N - m of the column:
    // Z = Xa + i Xb (n <= N/2), conj(Xa) + i conj(Xb) above; stored conj(Z).
    // The thread's loads are all issued (index clamped, no branch around
    // them) before the LDS stores; nu, the slots a thread needs, is uniform
    constexpr int EH = emax<Real>() / 2 + 1;
    const int nu = (h + L.tpc) / L.tpc;
    T a[EH], b[EH];
#pragma unroll
    for (int u = 0; u < EH; u++) {
      if (u < nu) {
        const int m = L.tj + u * L.tpc;
        const bool ok = valid && m <= h;
        a[u] = load_or_zero(ia, m <= h ? m : 0, ok);
        b[u] = load_or_zero(ib, m <= h ? m : 0, ok);
      }
    }
#pragma unroll
    for (int u = 0; u < EH; u++) {
      const int m = L.tj + u * L.tpc;
      if (u < nu && m <= h) {
        lds[L.col.at(m)] = mk<T>(a[u].x - b[u].y, -(a[u].y + b[u].x));
        if (m > 0 && m < h) lds[L.col.at(N - m)] = mk<T>(a[u].x + b[u].y, a[u].y - b[u].x);
      }
    }
  }
  __syncthreads();
  fft_columns<T, emax<Real>(), RC>(lds, L.col, L.tj, L.tpc, p, twl);
  if (!valid) return;
  Real *oa = out + (long)(2 * pr) * outRow, *ob = oa + outRow;
  for (int n = L.tj; n < Lout; n += L.tpc) {
    const T z = lds[L.col.at(n)];  // conj of the backward transform
    oa[n] = z.x;
    ob[n] = -z.y;
  }
}

// G of pass C.  GM 0: real G at (|k1|, |k2|, k3); GM 1: complex G at
// (|k1|, |k2|, k3 or |k3|) -- the radial kernels, whose grid is even; GM 2:
// complex G at (k1, k2, k3) over the full N1 x N2 x G3 grid (translated
// kernels), where gfl selects k -> -k per axis (bits 1, 2, 4) and the
// conjugate (bit 8).  Layout G[a2][a3 / 8][a1][a3 % 8], a1 over A1 values.
struct GIndex {
  long base;
  int A1;
};
__device__ __forceinline__ int g_neg(int k, int N) { return k ? N - k : 0; }
template <int GM>
__device__ __forceinline__ GIndex g_index(int N, int N2, int Kv, int N3fold, int gfl, bool valid,
                                          int colg, int C3)
{
  const int k2 = valid ? colg / C3 : 0, k3 = valid ? colg - k2 * C3 : 0;
  int a2, a3, G3, A1;
  if constexpr (GM == 2) {
    a2 = (gfl & 2) ? g_neg(k2, N2) : k2;
    a3 = (gfl & 4) ? g_neg(k3, N3fold) : k3;
    G3 = Kv;
    A1 = N;
  } else {
    a2 = (k2 <= N2 / 2) ? k2 : N2 - k2;
    a3 = (N3fold > 0 && k3 > N3fold / 2) ? N3fold - k3 : k3;
    G3 = N3fold > 0 ? (N3fold / 2 + 1) : Kv;
    A1 = N / 2 + 1;
  }
  GIndex r;
  r.base = (((long)a2 * ((G3 + 7) >> 3) + (a3 >> 3)) * A1) * 8 + (a3 & 7);
  r.A1 = A1;
  return r;
}
template <int GM> __device__ __forceinline__ int g_a1(int n, int N, int gfl)
{
  if constexpr (GM == 2) return (gfl & 1) ? g_neg(n, N) : n;
  else return (n <= N / 2) ? n : N - n;
}

// pass C: forward along axis 1, X <- X G, backward, keep the first Lin
// outputs.  Columns c = k2 * C3 + k3.
template <typename Real, int GM, int RC>
__global__ __launch_bounds__(NT) void k_fft_conv(int N, int Lin, int nCols, int pos, int N2, int gfl,
                                                 int C3, int Kv, int N3fold, PlanDev p,
                                                 const typename Cx<Real>::T *tw, const void *Gv,
                                                 typename Cx<Real>::T *__restrict__ data, int lgB)
{
  constexpr bool CG = GM > 0;
  using T = typename Cx<Real>::T;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const Lane L = lane_strided(lgB, N);
  const int colg = (xcd_wg() << lgB) + L.c;
  // C3: row length of the data (padded); Kv: valid entries of a row
  const bool valid = colg < nCols && colg % C3 < Kv;
  T *col = data + (valid ? colg : 0);
  const T *twl = twiddles_in<T, twl_lds<Real>()>(lds, (1 << lgB) * N, N, tw);
  stage_in<T, emax<Real>()>(lds, L, N, [&](int n) {
    return load_or_zero(col, (long)n * pos, valid && n < Lin);
  });
  __syncthreads();
  fft_columns<T, emax<Real>(), RC>(lds, L.col, L.tj, L.tpc, p, twl);
  // multiply by G and conjugate (for the backward transform)
  const GIndex gi = g_index<GM>(N, N2, Kv, N3fold, gfl, valid, colg, C3);
  constexpr int E = emax<Real>();
  using GT = typename std::conditional<CG, T, Real>::type;
  GT gv[E];
#pragma unroll
  for (int u = 0; u < E; u++) {  // the G loads of the thread, in flight together
    const int n = L.tj + u * L.tpc;
    const int nn = n < N ? n : 0;  // unconditional load (the value past N is unused)
    gv[u] = reinterpret_cast<const GT *>(Gv)[(long)g_a1<GM>(nn, N, gfl) * 8 + gi.base];
    if constexpr (GM == 2)
      if (gfl & 8) gv[u] = cconj(gv[u]);
  }
#pragma unroll
  for (int u = 0; u < E; u++) {
    const int n = L.tj + u * L.tpc;
    if (n < N) {
      T x = lds[L.col.at(n)];
      if constexpr (CG) x = cconj(cmul(x, gv[u]));
      else x = mk<T>(gv[u] * x.x, -gv[u] * x.y);
      lds[L.col.at(n)] = x;
    }
  }
  __syncthreads();
  fft_columns<T, emax<Real>(), RC>(lds, L.col, L.tj, L.tpc, p, twl);
  if (!valid) return;
  for (int n = L.tj; n < Lin; n += L.tpc) col[(long)n * pos] = cconj(lds[L.col.at(n)]);
}

// ------------------------------------------------------------ N = 8^L passes
// Strided passes (B, C, D) for N = 8^L (64, 512) with every index known at
// compile time: 4096 / N columns per workgroup, each thread owns two
// radix-8 butterflies of one column per stage.  The first stage reads its
// eight inputs straight from global memory (rows j + r N/8; with HIN only
// the first N/2 rows are non-zero, so four of them) and the last stage
// writes straight to global memory, so LDS carries only the L-1 exchanges
// between stages.  In the convolution the outputs of the forward transform's
// last stage are exactly the inputs of the backward transform's first stage
// for the same thread (rows j + r N/8), so the k-space multiply and the
// conjugation happen in registers between the two transforms.
template <typename T> __device__ __forceinline__ void dft8_half(T *x)
{
  // DFT8 of (x0, x1, x2, x3, 0, 0, 0, 0)
  using S = decltype(T::x);
  const S c = (S)0.70710678118654752440;
  // even part DFT4(x0, x2, 0, 0) = (x0 + x2, x0 - i x2, x0 - x2, x0 + i x2)
  const T e0 = cadd(x[0], x[2]), e2 = csub(x[0], x[2]);
  const T e1 = cadd(x[0], mul_mi(x[2])), e3 = csub(x[0], mul_mi(x[2]));
  const T o0 = cadd(x[1], x[3]), o2 = csub(x[1], x[3]);
  const T o1 = cadd(x[1], mul_mi(x[3])), o3 = csub(x[1], mul_mi(x[3]));
  const T t1 = mk<T>(c * (o1.x + o1.y), c * (o1.y - o1.x));   // * exp(-i pi/4)
  const T t2 = mul_mi(o2);                                    // * -i
  const T t3 = mk<T>(c * (o3.y - o3.x), -c * (o3.x + o3.y));  // * exp(-3i pi/4)
  x[0] = cadd(e0, o0);
  x[4] = csub(e0, o0);
  x[1] = cadd(e1, t1);
  x[5] = csub(e1, t1);
  x[2] = cadd(e2, t2);
  x[6] = csub(e2, t2);
  x[3] = cadd(e3, t3);
  x[7] = csub(e3, t3);
}

template <int L> struct P8 {
  static constexpr int N = L == 2 ? 64 : 512;
  static constexpr int NB = N / 8;          // butterflies per column and stage
  static constexpr int LGB = L == 2 ? 6 : 3;  // columns per workgroup: 4096 / N
  static constexpr int B = 1 << LGB;
  static constexpr int TPC = NT / B;        // threads per column
  static constexpr int BPT = NB / TPC;      // butterflies per thread
};

// LDS slot of (row, column c) of a radix-8 workgroup: rows of B columns.
// For the 512-point FP32 kernels a row (8 columns of 8 bytes) covers half of
// the LDS banks, and which half follows the row's parity; the first stage's
// writes of one wavefront go to rows 8 apart, all of one parity, so the rows
// of every odd group of 8 are stored pairwise swapped (row ^ 1): the
// wavefront's writes then alternate halves.  The other stages access whole
// groups of 8 consecutive rows, which the swap only permutes.
template <typename T, int L> constexpr bool lds_swap8() { return L == 3 && sizeof(T) == 8; }
template <typename T, int L> __device__ __forceinline__ int slot8(int row, int c)
{
  using P = P8<L>;
  if constexpr (lds_swap8<T, L>()) return ((row ^ ((row >> 3) & 1)) << P::LGB) + c;
  else return (row << P::LGB) + c;
}
// LDS elements of the columns (the twiddle table follows them)
template <typename T, int L> constexpr int lds_elems8() { return P8<L>::N << P8<L>::LGB; }

// middle stage NS of a column: read x[j + r NB], twiddle, DFT8, write
// y[(j - k) 8 + k + r NS] (LDS index * B + c), with a barrier between the
// reads and the writes
template <typename T, int L, int NS>
__device__ __forceinline__ void mid8(T *lds, int c, int tj, const T *tw)
{
  using P = P8<L>;
  T v[P::BPT][8];
#pragma unroll
  for (int u = 0; u < P::BPT; u++) {
    const int j = tj + u * P::TPC;
#pragma unroll
    for (int r = 0; r < 8; r++) v[u][r] = lds[slot8<T, L>(j + r * P::NB, c)];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < P::BPT; u++) {
    const int j = tj + u * P::TPC, k = j & (NS - 1);
    constexpr int stride = P::N / (8 * NS);
#pragma unroll
    for (int r = 1; r < 8; r++) v[u][r] = cmul(v[u][r], tw[r * k * stride]);
    DFT<T, 8>::run(v[u]);
    const int o = (j - k) * 8 + k;
#pragma unroll
    for (int r = 0; r < 8; r++) lds[slot8<T, L>(o + r * NS, c)] = v[u][r];
  }
  __syncthreads();
}

// last stage (NS = NB): k = j, outputs rows j + r NB, kept in registers
template <typename T, int L>
__device__ __forceinline__ void last8(T *lds, int c, int tj, const T *tw, T (*v)[8])
{
  using P = P8<L>;
#pragma unroll
  for (int u = 0; u < P::BPT; u++) {
    const int j = tj + u * P::TPC;
#pragma unroll
    for (int r = 0; r < 8; r++) v[u][r] = lds[slot8<T, L>(j + r * P::NB, c)];
  }
#pragma unroll
  for (int u = 0; u < P::BPT; u++) {
    const int j = tj + u * P::TPC;
#pragma unroll
    for (int r = 1; r < 8; r++) v[u][r] = cmul(v[u][r], tw[r * j]);
    DFT<T, 8>::run(v[u]);
  }
}

// first stage (NS = 1) from registers holding rows j + r NB: DFT8, write
// y[8 j + r]
template <typename T, int L, bool HALF>
__device__ __forceinline__ void first8(T *lds, int c, int tj, T (*v)[8])
{
  using P = P8<L>;
#pragma unroll
  for (int u = 0; u < P::BPT; u++) {
    const int j = tj + u * P::TPC;
    if constexpr (HALF) dft8_half(v[u]);
    else DFT<T, 8>::run(v[u]);
#pragma unroll
    for (int r = 0; r < 8; r++) lds[slot8<T, L>(8 * j + r, c)] = v[u][r];
  }
  __syncthreads();
}

// the transform of one column after its first stage: stages 1 .. L-1
template <typename T, int L>
__device__ __forceinline__ void rest8(T *lds, int c, int tj, const T *tw, T (*v)[8])
{
  if constexpr (L == 3) mid8<T, L, 8>(lds, c, tj, tw);
  last8<T, L>(lds, c, tj, tw, v);
}

// passes B and D: column o * nInner + i, element n at o * outer + i * inner + n * pos
template <typename Real, int L, bool HIN>
__global__ __launch_bounds__(NT) void k_fft8_cols(CGeom g, const typename Cx<Real>::T *tw,
                                                  const typename Cx<Real>::T *__restrict__ in,
                                                  typename Cx<Real>::T *__restrict__ out,
                                                  int backward)
{
  using T = typename Cx<Real>::T;
  using P = P8<L>;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int c = threadIdx.x & (P::B - 1), tj = threadIdx.x >> P::LGB;
  const int colg = (xcd_wg() << P::LGB) + c;
  const int o0 = colg / g.nInner;
  const bool valid = colg < g.nCols && colg - o0 * g.nInner < g.validInner;
  const int o = valid ? o0 : 0, i = valid ? colg - o0 * g.nInner : 0;
  const T *twl = twiddles_in<T, twl_lds<Real>()>(lds, lds_elems8<T, L>(), P::N, tw);
  const T *src = in + ((long)o * g.inOuter + (long)i * g.inInner);
  T v[P::BPT][8];
#pragma unroll
  for (int u = 0; u < P::BPT; u++) {
    const int j = tj + u * P::TPC;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      // the rows loaded are all below Lin (Lin = N, or N / 2 with HIN), and
      // a column past the last reads column 0 and is not stored, so the
      // loads need no condition and are all in flight together
      const int n = j + r * P::NB;
      if (HIN && r >= 4) {
        v[u][r] = mk<T>(0, 0);
      } else {
        const T x = src[(long)n * g.inPos];
        v[u][r] = backward ? cconj(x) : x;
      }
    }
  }
  // the barrier of first8 also orders the twiddle copy before its use
  first8<T, L, HIN>(lds, c, tj, v);
  rest8<T, L>(lds, c, tj, twl, v);
  if (!valid) return;
  T *dst = out + ((long)o * g.outOuter + (long)i * g.outInner);
#pragma unroll
  for (int u = 0; u < P::BPT; u++) {
    const int j = tj + u * P::TPC;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int n = j + r * P::NB;
      if (n < g.Lout) dst[(long)n * g.outPos] = backward ? cconj(v[u][r]) : v[u][r];
    }
  }
}

// pass C: forward along axis 1 (the first N/2 rows non-zero), times G,
// backward, the first N/2 rows kept; columns c = k2 * C3 + k3, Kv valid k3
template <typename Real, int L, int GM>
__global__ __launch_bounds__(NT) void k_fft8_conv(int nCols, int pos, int N2, int gfl, int C3, int Kv,
                                                  int N3fold, const typename Cx<Real>::T *tw,
                                                  const void *Gv,
                                                  typename Cx<Real>::T *__restrict__ data)
{
  constexpr bool CG = GM > 0;
  using T = typename Cx<Real>::T;
  using P = P8<L>;
  constexpr int N = P::N;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int c = threadIdx.x & (P::B - 1), tj = threadIdx.x >> P::LGB;
  const int colg = (xcd_wg() << P::LGB) + c;
  const bool valid = colg < nCols && colg % C3 < Kv;
  T *col = data + (valid ? colg : 0);
  const T *twl = twiddles_in<T, twl_lds<Real>()>(lds, lds_elems8<T, L>(), N, tw);
  // the G values of the thread's rows, fetched up front
  const GIndex gi = g_index<GM>(N, N2, Kv, N3fold, gfl, valid, colg, C3);
  using GT = typename std::conditional<CG, T, Real>::type;
  GT gv[P::BPT][8];
  T v[P::BPT][8];
#pragma unroll
  for (int u = 0; u < P::BPT; u++) {
    const int j = tj + u * P::TPC;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int n = j + r * P::NB;
      gv[u][r] = reinterpret_cast<const GT *>(Gv)[(long)g_a1<GM>(n, N, gfl) * 8 + gi.base];
      if constexpr (GM == 2)
        if (gfl & 8) gv[u][r] = cconj(gv[u][r]);
      // rows below N / 2 hold data; an invalid column reads column 0 and is
      // not stored, so the loads need no condition
      v[u][r] = r < 4 ? col[(long)n * pos] : mk<T>(0, 0);
    }
  }
  first8<T, L, true>(lds, c, tj, v);
  rest8<T, L>(lds, c, tj, twl, v);
  // times G, conjugated for the backward transform (rows j + r NB)
#pragma unroll
  for (int u = 0; u < P::BPT; u++)
#pragma unroll
    for (int r = 0; r < 8; r++) {
      if constexpr (CG) v[u][r] = cconj(cmul(v[u][r], gv[u][r]));
      else v[u][r] = mk<T>(gv[u][r] * v[u][r].x, -gv[u][r] * v[u][r].y);
    }
  __syncthreads();  // every thread has read the last forward stage
  first8<T, L, false>(lds, c, tj, v);
  rest8<T, L>(lds, c, tj, twl, v);
  if (!valid) return;
#pragma unroll
  for (int u = 0; u < P::BPT; u++) {
    const int j = tj + u * P::TPC;
#pragma unroll
    for (int r = 0; r < 4; r++) col[(long)(j + r * P::NB) * pos] = cconj(v[u][r]);
  }
}

// set-up: even extension of a compact line (length N/2+1) to length N,
// forward DFT, folded back to the compact line, in place (FP64)
__global__ __launch_bounds__(NT) void k_fft_even(int N, int nCols, int nInner, int outer, int inner,
                                                 int pos, PlanDev p, const double2 *tw,
                                                 double2 *__restrict__ data, int lgB)
{
  extern __shared__ char smem_raw[];
  double2 *lds = reinterpret_cast<double2 *>(smem_raw);
  const int h = N / 2;
  const Lane L = lane_strided(lgB, N);
  const int colg = (xcd_wg() << lgB) + L.c;
  const bool valid = colg < nCols;
  const int o = valid ? colg / nInner : 0, i = valid ? colg - o * nInner : 0;
  double2 *line = data + ((long)o * outer + (long)i * inner);
  stage_in<double2, 16>(lds, L, N, [&](int n) {
    return valid ? line[(long)((n <= h) ? n : N - n) * pos] : mk<double2>(0, 0);
  });
  __syncthreads();
  fft_columns<double2, 16, RC_ALL>(lds, L.col, L.tj, L.tpc, p, tw);
  if (!valid) return;
  for (int n = L.tj; n <= h; n += L.tpc) {
    double2 v = lds[L.col.at(n)];
    if (n != 0 && n != h) {
      const double2 w = lds[L.col.at(N - n)];
      v = mk<double2>(0.5 * (v.x + w.x), 0.5 * (v.y + w.y));
    }
    line[(long)n * pos] = v;
  }
}

// set-up, real even lines (the kernels 1/r and r^alpha): lines 2c and
// 2c + 1 of a pass share one complex transform z = a + i b.  The transform
// of a real even line is real, so A = Re Z and B = Im Z; folded back to the
// compact line as in k_fft_even, in place (FP64)
__global__ __launch_bounds__(NT) void k_fft_even_pair(int N, int nCols, int nInner, int outer,
                                                      int inner, int pos, PlanDev p,
                                                      const double2 *tw, double *__restrict__ data,
                                                      int lgB)
{
  extern __shared__ char smem_raw[];
  double2 *lds = reinterpret_cast<double2 *>(smem_raw);
  const int h = N / 2;
  const Lane L = lane_strided(lgB, N);
  const int pair = (xcd_wg() << lgB) + L.c;
  const int ca = 2 * pair, cb = 2 * pair + 1;
  const bool va = ca < nCols, vb = cb < nCols;
  auto line_of = [&](int c) {
    const int o = c / nInner, i = c - o * nInner;
    return data + ((long)o * outer + (long)i * inner);
  };
  double *la = line_of(va ? ca : 0), *lb = line_of(vb ? cb : 0);
  stage_in<double2, 16>(lds, L, N, [&](int n) {
    const long k = (long)((n <= h) ? n : N - n) * pos;
    return mk<double2>(va ? la[k] : 0.0, vb ? lb[k] : 0.0);
  });
  __syncthreads();
  fft_columns<double2, 16, RC_ALL>(lds, L.col, L.tj, L.tpc, p, tw);
  if (!va) return;
  for (int n = L.tj; n <= h; n += L.tpc) {
    double2 v = lds[L.col.at(n)];
    if (n != 0 && n != h) {
      const double2 w = lds[L.col.at(N - n)];
      v = mk<double2>(0.5 * (v.x + w.x), 0.5 * (v.y + w.y));
    }
    la[(long)n * pos] = v.x;
    if (vb) lb[(long)n * pos] = v.y;
  }
}

// ------------------------------------------------------------ host side
static PlanDev to_dev(const FFTPlan1D &p)
{
  PlanDev d;
  d.N = p.N;
  d.nst = p.nst;
  for (int i = 0; i < 24; i++) d.radix[i] = p.radix[i];
  return d;
}

// columns per workgroup: a power of two with B * N <= EMAX * NT (strided
// layout) or B * (N + 1) <= EMAX * NT (row layout), at most 64
static int lg_batch(int N, bool fp64, bool rows)
{
  const int cap = 16 * NT;
  (void)fp64;
  const int per = rows ? N + 1 : N;
  int lg = 0;
  while (lg < 6 && (2 << lg) * per <= cap) lg++;
  return lg;
}

static size_t lds_bytes(int N, int lgB, bool fp64, bool rows)
{
  return (size_t)(1 << lgB) * (rows ? N + 1 : N) * (fp64 ? 16 : 8);
}

// LDS of a pass: the columns, plus the twiddle table in FP32
static size_t pass_lds(int N, int lgB, bool fp64, bool rows)
{
  return lds_bytes(N, lgB, fp64, rows) + (fp64 ? 0 : (size_t)N * 8);
}

// LDS of a radix-8 pass (N = 64 or 512): the columns, plus the twiddle
// table in FP32
static size_t pass_lds8(int N, bool fp64) { return pass_lds(N, N == 64 ? 6 : 3, fp64, false); }

// L when the plan is L radix-8 stages of N = 8^L with L in {2, 3}, else 0
static int pow8_level(const FFTPlan1D &p)
{
  if ((p.N != 64 && p.N != 512) || p.nst != (p.N == 64 ? 2 : 3)) return 0;
  for (int i = 0; i < p.nst; i++)
    if (p.radix[i] != 8) return 0;
  return p.nst;
}

static int radix_class(const FFTPlan1D &p)
{
  bool r8 = true, pow2 = true;
  for (int i = 0; i < p.nst; i++) {
    const int r = p.radix[i];
    if (r != 8) r8 = false;
    if (r != 8 && r != 4 && r != 2) pow2 = false;
  }
  return r8 ? RC_R8 : pow2 ? RC_POW2 : RC_ALL;
}

static void check_int(long v, const char *what)
{
  if (v < 0 || v > 0x7fffffffL) throw Error{NPME_ERR_UNSUPPORTED, std::string(what) + " exceeds 2^31 elements"};
}

void fft_pass_complex(bool fp64, const PassGeom &g, const FFTPlan1D &plan, const void *tw,
                      const void *in, void *out, bool backward, hipStream_t s, bool pow8)
{
  const bool rows = (g.inPos == 1 && g.outPos == 1);
  const int lgB = lg_batch(g.N, fp64, rows);
  CGeom c;
  c.N = g.N;
  c.Lin = g.Lin;
  c.Lout = g.Lout;
  c.nCols = (int)g.nCols;
  c.nInner = (int)g.nInner;
  c.validInner = g.validInner < 0 ? (int)g.nInner : (int)g.validInner;
  c.inOuter = (int)g.inOuter;
  c.inInner = (int)g.inInner;
  c.inPos = (int)g.inPos;
  c.outOuter = (int)g.outOuter;
  c.outInner = (int)g.outInner;
  c.outPos = (int)g.outPos;
  check_int(g.nCols * std::max<long>(1, g.N), "FFT pass");
  const dim3 grid((unsigned)((g.nCols + (1 << lgB) - 1) >> lgB));
  const int L8 = pow8 ? pow8_level(plan) : 0;
  if (!rows && L8 > 0 && (g.Lin == g.N || 2 * g.Lin == g.N) && g.Lout <= g.N) {
    // N = 8^L: the compile-time kernel (4096 / N columns per workgroup)
    const size_t sh8 = pass_lds8(g.N, fp64);
    const bool half = 2 * g.Lin == g.N;
#define NPME_F8(R, LV, H)                                                                      \
  hipLaunchKernelGGL((k_fft8_cols<R, LV, H>), grid, dim3(NT), sh8, s, c,                       \
                     (const typename Cx<R>::T *)tw, (const typename Cx<R>::T *)in,            \
                     (typename Cx<R>::T *)out, (int)backward)
#define NPME_F8L(R, LV)                                                                        \
  do {                                                                                         \
    if (half) NPME_F8(R, LV, true);                                                            \
    else NPME_F8(R, LV, false);                                                                \
  } while (0)
    if (fp64) {
      if (L8 == 2) NPME_F8L(double, 2);
      else NPME_F8L(double, 3);
    } else {
      if (L8 == 2) NPME_F8L(float, 2);
      else NPME_F8L(float, 3);
    }
#undef NPME_F8
#undef NPME_F8L
    NPME_HIP_CHECK(hipGetLastError());
    return;
  }
  const int rc = radix_class(plan);
  const size_t sh = pass_lds(g.N, lgB, fp64, rows);
#define NPME_FFTC(R, ROWS, RCV)                                                                \
  hipLaunchKernelGGL((k_fft_complex<R, ROWS, RCV>), grid, dim3(NT), sh, s, c, to_dev(plan),    \
                     (const typename Cx<R>::T *)tw, (const typename Cx<R>::T *)in,            \
                     (typename Cx<R>::T *)out, lgB, (int)backward)
#define NPME_FFTC2(R, ROWS)                                                                    \
  do {                                                                                         \
    if (rc == RC_R8) NPME_FFTC(R, ROWS, RC_R8);                                                \
    else if (rc == RC_POW2) NPME_FFTC(R, ROWS, RC_POW2);                                       \
    else NPME_FFTC(R, ROWS, RC_ALL);                                                           \
  } while (0)
  if (fp64) {
    if (rows) NPME_FFTC2(double, true);
    else NPME_FFTC2(double, false);
  } else {
    if (rows) NPME_FFTC2(float, true);
    else NPME_FFTC2(float, false);
  }
#undef NPME_FFTC
#undef NPME_FFTC2
  NPME_HIP_CHECK(hipGetLastError());
}

void fft_pass_r2c_rows(bool fp64, int N, int Lin, long nRows, long inRowStride,
                       long outRowStride, const FFTPlan1D &plan, const void *tw, const void *in,
                       void *out, hipStream_t s)
{
  check_int(nRows * std::max(inRowStride, outRowStride), "FFT pass");
  const int nPairs = (int)(nRows / 2);
  const int lgB = lg_batch(N, fp64, true);
  const dim3 grid((unsigned)((nPairs + (1 << lgB) - 1) >> lgB));
  const int rc = radix_class(plan);
  const size_t sh = pass_lds(N, lgB, fp64, true);
#define NPME_R2C(R, RCV)                                                                       \
  hipLaunchKernelGGL((k_fft_r2c<R, RCV>), grid, dim3(NT), sh, s, N, Lin, nPairs, (int)inRowStride, \
                     (int)outRowStride, to_dev(plan), (const typename Cx<R>::T *)tw,           \
                     (const R *)in, (typename Cx<R>::T *)out, lgB)
#define NPME_R2C2(R)                                                                           \
  do {                                                                                         \
    if (rc == RC_R8) NPME_R2C(R, RC_R8);                                                       \
    else if (rc == RC_POW2) NPME_R2C(R, RC_POW2);                                              \
    else NPME_R2C(R, RC_ALL);                                                                  \
  } while (0)
  if (fp64) NPME_R2C2(double);
  else NPME_R2C2(float);
#undef NPME_R2C
#undef NPME_R2C2
  NPME_HIP_CHECK(hipGetLastError());
}

void fft_pass_c2r_rows(bool fp64, int N, int Lout, long nRows, long inRowStride,
                       long outRowStride, const FFTPlan1D &plan, const void *tw, const void *in,
                       void *out, hipStream_t s)
{
  check_int(nRows * std::max(inRowStride, outRowStride), "FFT pass");
  const int nPairs = (int)(nRows / 2);
  const int lgB = lg_batch(N, fp64, true);
  const dim3 grid((unsigned)((nPairs + (1 << lgB) - 1) >> lgB));
  const int rc = radix_class(plan);
  const size_t sh = pass_lds(N, lgB, fp64, true);
#define NPME_C2R(R, RCV)                                                                       \
  hipLaunchKernelGGL((k_fft_c2r<R, RCV>), grid, dim3(NT), sh, s, N, Lout, nPairs, (int)inRowStride, \
                     (int)outRowStride, to_dev(plan), (const typename Cx<R>::T *)tw,           \
                     (const typename Cx<R>::T *)in, (R *)out, lgB)
#define NPME_C2R2(R)                                                                           \
  do {                                                                                         \
    if (rc == RC_R8) NPME_C2R(R, RC_R8);                                                       \
    else if (rc == RC_POW2) NPME_C2R(R, RC_POW2);                                              \
    else NPME_C2R(R, RC_ALL);                                                                  \
  } while (0)
  if (fp64) NPME_C2R2(double);
  else NPME_C2R2(float);
#undef NPME_C2R
#undef NPME_C2R2
  NPME_HIP_CHECK(hipGetLastError());
}

void fft_pass_convolve(bool fp64, bool complexG, int N, int Lin, long nCols, long pos_stride,
                       int N2, int G2, long C3, long Kv, const FFTPlan1D &plan, const void *tw,
                       const void *G, void *data, hipStream_t s, bool pow8, bool fullG, int gflags)
{
  // C3: entries per axis-2 row in the data (padded), Kv of them valid.
  // complexG: complex charges and kernel, the axis-3 index runs over
  // N3 = Kv (folded to |k3| for a radial kernel's G); else the R2C data of
  // real charges, k3 in [0, Kv).  fullG: G is complex over the full grid
  // (GM 2), for either data; the k3 reversal (gflags bit 4) needs complex
  // data.
  (void)G2;
  check_int(nCols * (long)Lin, "FFT pass");
  const int gm = fullG ? 2 : complexG ? 1 : 0;
  if (fullG && !complexG && (gflags & 4))
    throw Error{NPME_ERR_UNSUPPORTED, "axis-3 reversal of G needs complex data"};
  const int lgB = lg_batch(N, fp64, false);
  const dim3 grid((unsigned)((nCols + (1 << lgB) - 1) >> lgB));
  const int rc = radix_class(plan);
  const size_t sh = pass_lds(N, lgB, fp64, false);
  // the axis-3 length of complex data (0 for the R2C data of real kernels)
  const int N3fold = complexG ? (int)Kv : 0;
  const int L8 = pow8 ? pow8_level(plan) : 0;
  if (L8 > 0 && 2 * Lin == N) {
    const size_t sh8 = pass_lds8(N, fp64);
#define NPME_C8(R, LV, GMV)                                                                    \
  hipLaunchKernelGGL((k_fft8_conv<R, LV, GMV>), grid, dim3(NT), sh8, s, (int)nCols,           \
                     (int)pos_stride, N2, gflags, (int)C3, (int)Kv, N3fold,                     \
                     (const typename Cx<R>::T *)tw, G, (typename Cx<R>::T *)data)
#define NPME_C8L(R, GMV)                                                                       \
  do {                                                                                         \
    if (L8 == 2) NPME_C8(R, 2, GMV);                                                           \
    else NPME_C8(R, 3, GMV);                                                                   \
  } while (0)
#define NPME_C8G(R)                                                                            \
  do {                                                                                         \
    if (gm == 2) NPME_C8L(R, 2);                                                               \
    else if (gm == 1) NPME_C8L(R, 1);                                                          \
    else NPME_C8L(R, 0);                                                                       \
  } while (0)
    if (fp64) NPME_C8G(double);
    else NPME_C8G(float);
#undef NPME_C8
#undef NPME_C8L
#undef NPME_C8G
    NPME_HIP_CHECK(hipGetLastError());
    return;
  }
#define NPME_CONV(R, GMV, RCV)                                                                 \
  hipLaunchKernelGGL((k_fft_conv<R, GMV, RCV>), grid, dim3(NT), sh, s, N, Lin, (int)nCols,     \
                     (int)pos_stride, N2, gflags, (int)C3, (int)Kv, N3fold, to_dev(plan),       \
                     (const typename Cx<R>::T *)tw, G, (typename Cx<R>::T *)data, lgB)
#define NPME_CONV2(R, GMV)                                                                     \
  do {                                                                                         \
    if (rc == RC_R8) NPME_CONV(R, GMV, RC_R8);                                                 \
    else if (rc == RC_POW2) NPME_CONV(R, GMV, RC_POW2);                                        \
    else NPME_CONV(R, GMV, RC_ALL);                                                            \
  } while (0)
#define NPME_CONVG(R)                                                                          \
  do {                                                                                         \
    if (gm == 2) NPME_CONV2(R, 2);                                                             \
    else if (gm == 1) NPME_CONV2(R, 1);                                                        \
    else NPME_CONV2(R, 0);                                                                     \
  } while (0)
  if (fp64) NPME_CONVG(double);
  else NPME_CONVG(float);
#undef NPME_CONVG
#undef NPME_CONV
#undef NPME_CONV2
  NPME_HIP_CHECK(hipGetLastError());
}

void fft_pass_even_real(int N, long nCols, long nInner, long outer, long inner, long pos,
                        const FFTPlan1D &plan, const void *tw, double *data, hipStream_t s)
{
  check_int(nCols * (long)(N / 2 + 1), "FFT pass");
  const int lgB = lg_batch(N, true, false);
  const long pairs = (nCols + 1) / 2;
  const dim3 grid((unsigned)((pairs + (1 << lgB) - 1) >> lgB));
  const size_t sh = lds_bytes(N, lgB, true, false);
  hipLaunchKernelGGL(k_fft_even_pair, grid, dim3(NT), sh, s, N, (int)nCols, (int)nInner, (int)outer,
                     (int)inner, (int)pos, to_dev(plan), (const double2 *)tw, data, lgB);
  NPME_HIP_CHECK(hipGetLastError());
}

void fft_pass_even(int N, long nCols, long nInner, long outer, long inner, long pos,
                   const FFTPlan1D &plan, const void *tw, void *data, hipStream_t s)
{
  check_int(nCols * (long)(N / 2 + 1), "FFT pass");
  const int lgB = lg_batch(N, true, false);
  const dim3 grid((unsigned)((nCols + (1 << lgB) - 1) >> lgB));
  const size_t sh = lds_bytes(N, lgB, true, false);
  hipLaunchKernelGGL(k_fft_even, grid, dim3(NT), sh, s, N, (int)nCols, (int)nInner, (int)outer,
                     (int)inner, (int)pos, to_dev(plan), (const double2 *)tw, (double2 *)data, lgB);
  NPME_HIP_CHECK(hipGetLastError());
}

}  // namespace npme
