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
// host.  One workgroup transforms B columns held in LDS (column-major,
// stride N+1 elements); every thread keeps all butterflies of a stage in
// registers between the read and the write barrier, so one LDS buffer
// suffices.
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
// multiply by -i
template <typename T> __device__ __forceinline__ T mul_mi(T a) { return mk<T>(a.y, -a.x); }

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

// odd radix R = 5, 7: pairwise symmetric form
template <typename T, int R> struct DFTOdd {
  __device__ static void run(T *x, const double *cs, const double *sn)
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
        const S c = (S)cs[(j * k) % R], s = (S)sn[(j * k) % R];
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

__device__ __constant__ double c5_cos[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410,
                                            -0.80901699437494742410, 0.30901699437494742410};
__device__ __constant__ double c5_sin[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917,
                                            -0.58778525229247312917, -0.95105651629515357212};
__device__ __constant__ double c7_cos[7] = {1.0,
                                            0.62348980185873353053,
                                            -0.22252093395631440429,
                                            -0.90096886790241912624,
                                            -0.90096886790241912624,
                                            -0.22252093395631440429,
                                            0.62348980185873353053};
__device__ __constant__ double c7_sin[7] = {0.0,
                                            0.78183148246802980871,
                                            0.97492791218182360702,
                                            0.43388373911755812048,
                                            -0.43388373911755812048,
                                            -0.97492791218182360702,
                                            -0.78183148246802980871};

template <typename T, int R> __device__ __forceinline__ void dft_any(T *x)
{
  if constexpr (R == 5)
    DFTOdd<T, 5>::run(x, c5_cos, c5_sin);
  else if constexpr (R == 7)
    DFTOdd<T, 7>::run(x, c7_cos, c7_sin);
  else
    DFT<T, R>::run(x);
}

// ------------------------------------------------------------ Stockham stage
// Columns c in [0, B) of length N at lds[c*ld + n].  Butterfly j of a column
// reads x[j + r N/R], twiddles by w^(r k N/(Ns R)) with k = j mod Ns, and
// writes y[(j/Ns) Ns R + k + r Ns].
template <typename T, int R, int BPT>
__device__ __forceinline__ void stage(T *lds, int ld, int B, int N, int Ns, const T *tw)
{
  const int nb = N / R;
  const int total = B * nb;
  const int stride = N / (Ns * R);
  T v[BPT][R];
  int cj[BPT];
#pragma unroll
  for (int u = 0; u < BPT; u++) {
    const int b = threadIdx.x + u * NT;
    if (b < total) {
      const int c = b / nb, j = b - c * nb;
      cj[u] = c * ld + j;
      const T *col = lds + c * ld;
#pragma unroll
      for (int r = 0; r < R; r++) v[u][r] = col[j + r * nb];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < BPT; u++) {
    const int b = threadIdx.x + u * NT;
    if (b < total) {
      const int c = b / nb, j = b - c * nb;
      const int k = j % Ns;
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < R; r++) v[u][r] = cmul(v[u][r], tw[r * k * stride]);
      }
      dft_any<T, R>(v[u]);
      T *col = lds + c * ld;
      const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) col[idx + r * Ns] = v[u][r];
    }
  }
  (void)cj;
  __syncthreads();
}

template <typename T, int EMAX>
__device__ void fft_columns(T *lds, int ld, int B, const PlanDev &p, const T *tw)
{
  int Ns = 1;
  for (int s = 0; s < p.nst; s++) {
    const int R = p.radix[s];
    switch (R) {
      case 8: stage<T, 8, (EMAX + 7) / 8>(lds, ld, B, p.N, Ns, tw); break;
      case 4: stage<T, 4, (EMAX + 3) / 4>(lds, ld, B, p.N, Ns, tw); break;
      case 2: stage<T, 2, (EMAX + 1) / 2>(lds, ld, B, p.N, Ns, tw); break;
      case 3: stage<T, 3, (EMAX + 2) / 3>(lds, ld, B, p.N, Ns, tw); break;
      case 5: stage<T, 5, (EMAX + 4) / 5>(lds, ld, B, p.N, Ns, tw); break;
      case 7: stage<T, 7, (EMAX + 6) / 7>(lds, ld, B, p.N, Ns, tw); break;
      default: break;
    }
    Ns *= R;
  }
}

template <typename Real> constexpr int emax() { return sizeof(Real) == 4 ? 32 : 16; }

// ------------------------------------------------------------ pass kernels
// generic complex pass along strided or contiguous columns
template <typename Real>
__global__ __launch_bounds__(NT) void k_fft_complex(PassGeom g, PlanDev p, const typename Cx<Real>::T *tw,
                                                    const typename Cx<Real>::T *in,
                                                    typename Cx<Real>::T *out, int B, int backward)
{
  using T = typename Cx<Real>::T;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int N = g.N, ld = N + 1;
  const long col0 = (long)blockIdx.x * B;
  const bool rowMode = (g.inPos == 1);
  for (int e = threadIdx.x; e < B * N; e += NT) {
    int c, n;
    if (rowMode) { c = e / N; n = e - c * N; }
    else { n = e / B; c = e - n * B; }
    const long col = col0 + c;
    T v = mk<T>(0, 0);
    if (n < g.Lin && col < g.nCols) {
      const long o = col / g.nInner, i = col - o * g.nInner;
      v = in[o * g.inOuter + i * g.inInner + (long)n * g.inPos];
      if (backward) v = cconj(v);
    }
    lds[c * ld + n] = v;
  }
  __syncthreads();
  fft_columns<T, emax<Real>()>(lds, ld, B, p, tw);
  const bool rowOut = (g.outPos == 1);
  for (int e = threadIdx.x; e < B * g.Lout; e += NT) {
    int c, n;
    if (rowOut) { c = e / g.Lout; n = e - c * g.Lout; }
    else { n = e / B; c = e - n * B; }
    const long col = col0 + c;
    if (col < g.nCols) {
      const long o = col / g.nInner, i = col - o * g.nInner;
      T v = lds[c * ld + n];
      if (backward) v = cconj(v);
      out[o * g.outOuter + i * g.outInner + (long)n * g.outPos] = v;
    }
  }
}

// pass A: two real rows per complex column, R2C by the two-for-one split
template <typename Real>
__global__ __launch_bounds__(NT) void k_fft_r2c(int N, int Lin, long nPairs, long inRow, long outRow,
                                                PlanDev p, const typename Cx<Real>::T *tw,
                                                const Real *in, typename Cx<Real>::T *out, int B)
{
  using T = typename Cx<Real>::T;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int ld = N + 1;
  const long pr0 = (long)blockIdx.x * B;
  for (int e = threadIdx.x; e < B * N; e += NT) {
    const int c = e / N, n = e - c * N;
    const long pr = pr0 + c;
    T v = mk<T>(0, 0);
    if (n < Lin && pr < nPairs) v = mk<T>(in[(2 * pr) * inRow + n], in[(2 * pr + 1) * inRow + n]);
    lds[c * ld + n] = v;
  }
  __syncthreads();
  fft_columns<T, emax<Real>()>(lds, ld, B, p, tw);
  const int K = N / 2 + 1;
  const Real h = (Real)0.5;
  for (int e = threadIdx.x; e < B * K; e += NT) {
    const int c = e / K, k = e - c * K;
    const long pr = pr0 + c;
    if (pr < nPairs) {
      const T z = lds[c * ld + k];
      const T w = cconj(lds[c * ld + (k == 0 ? 0 : N - k)]);
      const T s = cadd(z, w), d = csub(z, w);
      out[(2 * pr) * outRow + k] = mk<T>(h * s.x, h * s.y);
      out[(2 * pr + 1) * outRow + k] = mk<T>(h * d.y, -h * d.x);  // -i d / 2
    }
  }
}

// pass E: Hermitian half rows -> two real rows (backward transform)
template <typename Real>
__global__ __launch_bounds__(NT) void k_fft_c2r(int N, int Lout, long nPairs, long inRow, long outRow,
                                                PlanDev p, const typename Cx<Real>::T *tw,
                                                const typename Cx<Real>::T *in, Real *out, int B)
{
  using T = typename Cx<Real>::T;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int ld = N + 1;
  const long pr0 = (long)blockIdx.x * B;
  const int h = N / 2;
  for (int e = threadIdx.x; e < B * N; e += NT) {
    const int c = e / N, n = e - c * N;
    const long pr = pr0 + c;
    T v = mk<T>(0, 0);
    if (pr < nPairs) {
      const int m = (n <= h) ? n : N - n;
      const T a = in[(2 * pr) * inRow + m], b = in[(2 * pr + 1) * inRow + m];
      // Z = Xa + i Xb (n <= N/2), conj(Xa) + i conj(Xb) above; store conj(Z)
      if (n <= h) v = mk<T>(a.x - b.y, -(a.y + b.x));
      else v = mk<T>(a.x + b.y, -(-a.y + b.x));
    }
    lds[c * ld + n] = v;
  }
  __syncthreads();
  fft_columns<T, emax<Real>()>(lds, ld, B, p, tw);
  for (int e = threadIdx.x; e < B * Lout; e += NT) {
    const int c = e / Lout, n = e - c * Lout;
    const long pr = pr0 + c;
    if (pr < nPairs) {
      const T z = lds[c * ld + n];  // conj of the backward transform
      out[(2 * pr) * outRow + n] = z.x;
      out[(2 * pr + 1) * outRow + n] = -z.y;
    }
  }
}

// pass C: forward along axis 1, X <- X G, backward, keep the first Lin
// outputs.  Columns c = k2 * C3 + k3; G index (|k1|, |k2|, k3 or |k3|).
template <typename Real, bool CG>
__global__ __launch_bounds__(NT) void k_fft_conv(int N, int Lin, long nCols, long pos, int N2, int G2,
                                                 long C3, int N3fold, PlanDev p,
                                                 const typename Cx<Real>::T *tw, const void *Gv,
                                                 typename Cx<Real>::T *data, int B)
{
  using T = typename Cx<Real>::T;
  extern __shared__ char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int ld = N + 1;
  const long col0 = (long)blockIdx.x * B;
  for (int e = threadIdx.x; e < B * N; e += NT) {
    const int n = e / B, c = e - n * B;
    const long col = col0 + c;
    T v = mk<T>(0, 0);
    if (n < Lin && col < nCols) v = data[col + (long)n * pos];
    lds[c * ld + n] = v;
  }
  __syncthreads();
  fft_columns<T, emax<Real>()>(lds, ld, B, p, tw);
  // multiply by G and conjugate (for the backward transform)
  const long G3 = N3fold > 0 ? (N3fold / 2 + 1) : C3;
  for (int e = threadIdx.x; e < B * N; e += NT) {
    const int n = e / B, c = e - n * B;
    const long col = col0 + c;
    if (col < nCols) {
      const long k2 = col / C3, k3 = col - k2 * C3;
      const long a2 = (k2 <= N2 / 2) ? k2 : N2 - k2;
      const long a3 = (N3fold > 0 && k3 > N3fold / 2) ? N3fold - k3 : k3;
      const long a1 = (n <= N / 2) ? n : N - n;
      const long gi = (a1 * G2 + a2) * G3 + a3;
      T x = lds[c * ld + n];
      if constexpr (CG) {
        const T gg = reinterpret_cast<const T *>(Gv)[gi];
        x = cconj(cmul(x, gg));
      } else {
        const Real gg = reinterpret_cast<const Real *>(Gv)[gi];
        x = mk<T>(gg * x.x, -gg * x.y);
      }
      lds[c * ld + n] = x;
    }
  }
  __syncthreads();
  fft_columns<T, emax<Real>()>(lds, ld, B, p, tw);
  for (int e = threadIdx.x; e < B * Lin; e += NT) {
    const int n = e / B, c = e - n * B;
    const long col = col0 + c;
    if (col < nCols) data[col + (long)n * pos] = cconj(lds[c * ld + n]);
  }
}

// set-up: even extension of a compact line (length N/2+1) to length N,
// forward DFT, folded back to the compact line, in place (FP64)
__global__ __launch_bounds__(NT) void k_fft_even(int N, long nCols, long nInner, long outer, long inner,
                                                 long pos, PlanDev p, const double2 *tw,
                                                 double2 *data, int B)
{
  extern __shared__ char smem_raw[];
  double2 *lds = reinterpret_cast<double2 *>(smem_raw);
  const int ld = N + 1, h = N / 2;
  const long col0 = (long)blockIdx.x * B;
  for (int e = threadIdx.x; e < B * N; e += NT) {
    const int n = e / B, c = e - n * B;
    const long col = col0 + c;
    double2 v = mk<double2>(0, 0);
    if (col < nCols) {
      const long o = col / nInner, i = col - o * nInner;
      const int m = (n <= h) ? n : N - n;
      v = data[o * outer + i * inner + (long)m * pos];
    }
    lds[c * ld + n] = v;
  }
  __syncthreads();
  fft_columns<double2, 16>(lds, ld, B, p, tw);
  for (int e = threadIdx.x; e < B * (h + 1); e += NT) {
    const int n = e / B, c = e - n * B;
    const long col = col0 + c;
    if (col < nCols) {
      const long o = col / nInner, i = col - o * nInner;
      double2 v = lds[c * ld + n];
      if (n != 0 && n != h) {
        const double2 w = lds[c * ld + (N - n)];
        v = mk<double2>(0.5 * (v.x + w.x), 0.5 * (v.y + w.y));
      }
      data[o * outer + i * inner + (long)n * pos] = v;
    }
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

// columns per workgroup: power of two, B (N+1) <= EMAX * NT
static int batch_for(int N, bool fp64)
{
  const int cap = (fp64 ? 16 : 32) * NT;
  int B = 1;
  while (B * 2 * (N + 1) <= cap && B < 64) B *= 2;
  return B;
}

static size_t lds_bytes(int N, int B, bool fp64) { return (size_t)B * (N + 1) * (fp64 ? 16 : 8); }

void fft_pass_complex(bool fp64, const PassGeom &g, const FFTPlan1D &plan, const void *tw,
                      const void *in, void *out, bool backward, hipStream_t s)
{
  const int B = batch_for(g.N, fp64);
  const dim3 grid((unsigned)((g.nCols + B - 1) / B));
  const size_t sh = lds_bytes(g.N, B, fp64);
  if (fp64)
    hipLaunchKernelGGL(k_fft_complex<double>, grid, dim3(NT), sh, s, g, to_dev(plan),
                       (const double2 *)tw, (const double2 *)in, (double2 *)out, B, (int)backward);
  else
    hipLaunchKernelGGL(k_fft_complex<float>, grid, dim3(NT), sh, s, g, to_dev(plan),
                       (const float2 *)tw, (const float2 *)in, (float2 *)out, B, (int)backward);
  NPME_HIP_CHECK(hipGetLastError());
}

void fft_pass_r2c_rows(bool fp64, int N, int Lin, long nRows, long inRowStride,
                       long outRowStride, const FFTPlan1D &plan, const void *tw, const void *in,
                       void *out, hipStream_t s)
{
  const long nPairs = nRows / 2;
  const int B = batch_for(N, fp64);
  const dim3 grid((unsigned)((nPairs + B - 1) / B));
  const size_t sh = lds_bytes(N, B, fp64);
  if (fp64)
    hipLaunchKernelGGL(k_fft_r2c<double>, grid, dim3(NT), sh, s, N, Lin, nPairs, inRowStride,
                       outRowStride, to_dev(plan), (const double2 *)tw, (const double *)in,
                       (double2 *)out, B);
  else
    hipLaunchKernelGGL(k_fft_r2c<float>, grid, dim3(NT), sh, s, N, Lin, nPairs, inRowStride,
                       outRowStride, to_dev(plan), (const float2 *)tw, (const float *)in,
                       (float2 *)out, B);
  NPME_HIP_CHECK(hipGetLastError());
}

void fft_pass_c2r_rows(bool fp64, int N, int Lout, long nRows, long inRowStride,
                       long outRowStride, const FFTPlan1D &plan, const void *tw, const void *in,
                       void *out, hipStream_t s)
{
  const long nPairs = nRows / 2;
  const int B = batch_for(N, fp64);
  const dim3 grid((unsigned)((nPairs + B - 1) / B));
  const size_t sh = lds_bytes(N, B, fp64);
  if (fp64)
    hipLaunchKernelGGL(k_fft_c2r<double>, grid, dim3(NT), sh, s, N, Lout, nPairs, inRowStride,
                       outRowStride, to_dev(plan), (const double2 *)tw, (const double2 *)in,
                       (double *)out, B);
  else
    hipLaunchKernelGGL(k_fft_c2r<float>, grid, dim3(NT), sh, s, N, Lout, nPairs, inRowStride,
                       outRowStride, to_dev(plan), (const float2 *)tw, (const float2 *)in,
                       (float *)out, B);
  NPME_HIP_CHECK(hipGetLastError());
}

void fft_pass_convolve(bool fp64, bool complexG, int N, int Lin, long nCols, long pos_stride,
                       int N2, int G2, long C3, const FFTPlan1D &plan, const void *tw,
                       const void *G, void *data, hipStream_t s)
{
  // C3: number of axis-3 entries per axis-2 row in the data; for complex
  // kernels the axis-3 index is folded (|k3|) against N3 = C3.
  const int B = batch_for(N, fp64);
  const dim3 grid((unsigned)((nCols + B - 1) / B));
  const size_t sh = lds_bytes(N, B, fp64);
  const int N3fold = complexG ? (int)C3 : 0;
  if (fp64) {
    if (complexG)
      hipLaunchKernelGGL((k_fft_conv<double, true>), grid, dim3(NT), sh, s, N, Lin, nCols,
                         pos_stride, N2, G2, C3, N3fold, to_dev(plan), (const double2 *)tw, G,
                         (double2 *)data, B);
    else
      hipLaunchKernelGGL((k_fft_conv<double, false>), grid, dim3(NT), sh, s, N, Lin, nCols,
                         pos_stride, N2, G2, C3, N3fold, to_dev(plan), (const double2 *)tw, G,
                         (double2 *)data, B);
  } else {
    if (complexG)
      hipLaunchKernelGGL((k_fft_conv<float, true>), grid, dim3(NT), sh, s, N, Lin, nCols,
                         pos_stride, N2, G2, C3, N3fold, to_dev(plan), (const float2 *)tw, G,
                         (float2 *)data, B);
    else
      hipLaunchKernelGGL((k_fft_conv<float, false>), grid, dim3(NT), sh, s, N, Lin, nCols,
                         pos_stride, N2, G2, C3, N3fold, to_dev(plan), (const float2 *)tw, G,
                         (float2 *)data, B);
  }
  NPME_HIP_CHECK(hipGetLastError());
}

void fft_pass_even(int N, long nCols, long nInner, long outer, long inner, long pos,
                   const FFTPlan1D &plan, const void *tw, void *data, hipStream_t s)
{
  const int B = batch_for(N, true);
  const dim3 grid((unsigned)((nCols + B - 1) / B));
  const size_t sh = lds_bytes(N, B, true);
  hipLaunchKernelGGL(k_fft_even, grid, dim3(NT), sh, s, N, nCols, nInner, outer, inner, pos,
                     to_dev(plan), (const double2 *)tw, (double2 *)data, B);
  NPME_HIP_CHECK(hipGetLastError());
}

}  // namespace npme
