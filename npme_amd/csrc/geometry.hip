// Set-up of one geometry on the device: embedded cell keys, the stable sort
// of the charges by cell, the per-cell ranges, cluster-relative positions,
// spline windows and the per-tile charge lists of spreading and
// interpolation.  Every index is computed with the same FP64 operations as
// the host set-up (no contraction into FMAs), and the sorts are stable, so
// the permutation and every list are bit-identical to the host's; the
// per-charge cell index follows src/PartitionBox.cpp:311-351 and
// src/PartitionEmbeddedBox.cpp:266-455, the spline window
// src/RecSumSupportFunctions.h:123-149.
#include <cstring>
#include <deque>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include "npme_internal.h"

namespace npme {

// error bits written by the set-up kernels
enum { GEO_ERR_CELL = 1, GEO_ERR_WINDOW = 2, GEO_ERR_FIXED = 4 };

__device__ __forceinline__ bool d_coord2cell(double x, double cs, unsigned long ncell,
                                             unsigned long &v)
{
#pragma clang fp contract(off)
  const double tol = 1.0e-14;
  if (x < -tol) return false;
  unsigned long t = (unsigned long)(x / cs);
  if (t > ncell) return false;
  if (t == ncell) {
    if (fabs(x - (double)t * cs) < tol) t--;
    else return false;
  }
  v = t;
  return true;
}

__global__ __launch_bounds__(256) void k_cell_keys(size_t n, const double *__restrict__ coord,
                                                   GeoParams g, unsigned long *__restrict__ key,
                                                   int *__restrict__ idx,
                                                   int4 *__restrict__ gcell, int *__restrict__ err)
{
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned long a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
  bool ok = true;
  for (int d = 0; d < 3; d++) {
    const double x1 = coord[3 * i + d] + (-g.mn[d]);
    ok = ok && d_coord2cell(x1, g.csA, g.nA[d], a[d]);
    const double x2 = x1 - (double)a[d] * g.csA;
    ok = ok && d_coord2cell(x2, g.csB, g.nP, b[d]);
  }
  if (!ok) {
    err[0] = GEO_ERR_CELL;
    a[0] = a[1] = a[2] = b[0] = b[1] = b[2] = 0;
  }
  const unsigned long Nc = g.nP * g.nP * g.nP;
  const unsigned long tA = (a[0] * g.nA[1] + a[1]) * g.nA[2] + a[2];
  key[i] = tA * Nc + (b[0] * g.nP + b[1]) * g.nP + b[2];
  idx[i] = (int)i;
  gcell[i] = make_int4((int)(a[0] * g.nP + b[0]), (int)(a[1] * g.nP + b[1]),
                       (int)(a[2] * g.nP + b[2]), 0);
}

template <typename K>
__device__ __forceinline__ size_t lower_bound_d(const K *a, size_t n, K v)
{
  size_t lo = 0, hi = n;
  while (lo < hi) {
    const size_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// (start, count) of every cell of the dense cell grid; an empty cell starts
// where its key would be
__global__ __launch_bounds__(256) void k_cell_ranges(long ncell, GeoParams g,
                                                     const unsigned long *__restrict__ skey,
                                                     size_t n, int *__restrict__ cellSE)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const long g2 = c % g.nb[2], g1 = (c / g.nb[2]) % g.nb[1], g0 = c / ((long)g.nb[1] * g.nb[2]);
  const unsigned long nP = g.nP, Nc = nP * nP * nP;
  const unsigned long a0 = g0 / nP, a1 = g1 / nP, a2 = g2 / nP;
  const unsigned long kc =
      ((a0 * g.nA[1] + a1) * g.nA[2] + a2) * Nc + ((g0 % nP) * nP + (g1 % nP)) * nP + (g2 % nP);
  const size_t lo = lower_bound_d(skey, n, kc);
  const size_t hi = lower_bound_d(skey, n, kc + 1);
  cellSE[2 * c] = (int)lo;
  cellSE[2 * c + 1] = (int)(hi - lo);
}

// cell of each sorted charge
__global__ __launch_bounds__(256) void k_gather_cells(size_t n, const int *__restrict__ perm,
                                                      const int4 *__restrict__ gcell,
                                                      int4 *__restrict__ cellOf)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) cellOf[s] = gcell[perm[s]];
}

// per sorted charge: cluster-relative position (FP64 or int32 fixed
// point), spline window start u0 = m0 + N/4 and fractional argument, and
// the number of 8^3 spreading tiles its window touches
template <bool FP64>
__global__ __launch_bounds__(256) void k_charge_setup(size_t n, const int *__restrict__ perm,
                                                      const double *__restrict__ coord,
                                                      const int4 *__restrict__ cellOf, GeoParams g,
                                                      void *__restrict__ src,
                                                      float4 *__restrict__ srcRes,
                                                      int4 *__restrict__ ms,
                                                      void *__restrict__ frac,
                                                      int *__restrict__ ntile,
                                                      int *__restrict__ err)
{
#pragma clang fp contract(off)
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int i = perm[s];
  const int4 gc = cellOf[s];
  const int gcv[3] = {gc.x, gc.y, gc.z};
  double loc[3], fr[3];
  int u0[3];
  int cnt = 1;
  for (int d = 0; d < 3; d++) {
    const double x = coord[3 * (size_t)i + d];
    // the list walk reads positions in one frame (the cells' low corner),
    // the cell-row sum relative to the charge's cluster corner
    loc[d] = g.globalFrame ? x + (-g.mn[d]) : (x + (-g.mn[d])) - (double)(gcv[d] / (int)g.nP) * g.csA;
    const double w = g.NoverL[d] * (x - g.R[d]);
    const long m0 = (w < 0) ? (long)(-w) + 1 : (long)(-w);
    if (m0 < -g.N[d] / 4 || m0 + g.Bn - 1 > g.N[d] / 4 - 1) {
      err[0] = GEO_ERR_WINDOW;
      u0[d] = 0;
      fr[d] = 0;
    } else {
      u0[d] = (int)(m0 + g.N[d] / 4);
      fr[d] = w + (double)m0;
    }
    cnt *= (u0[d] + (int)g.Bn - 1) / g.ts[d] - u0[d] / g.ts[d] + 1;
  }
  ms[s] = make_int4(u0[0], u0[1], u0[2], 0);
  ntile[s] = cnt;
  if (FP64) {
    ((double4 *)src)[s] = make_double4(loc[0], loc[1], loc[2], 0.0);
    ((double4 *)frac)[s] = make_double4(fr[0], fr[1], fr[2], 0.0);
  } else {
    int v[3];
    float res[3];
    for (int d = 0; d < 3; d++) {
      const double u = loc[d] * g.fxScale;
      const long r = g.globalFrame ? (fabs(u) < 4.0e18 ? llround(u) : 0) : lround(u);
      if (g.globalFrame ? !(fabs(u) < 4.0e18) : (r < -(1L << 20) || r > (1L << g.fxShift) + (1L << 20))) {
        err[0] = GEO_ERR_FIXED;
        v[d] = 0;
      } else {
        // the list walk's frame wraps modulo 2^32
        v[d] = (int)(unsigned)(unsigned long)r;
      }
      res[d] = (float)(u - (double)r);
    }
    ((int4 *)src)[s] = make_int4(v[0], v[1], v[2], 0);
    srcRes[s] = make_float4(res[0], res[1], res[2], 0.0f);
    ((float4 *)frac)[s] = make_float4((float)fr[0], (float)fr[1], (float)fr[2], 0.0f);
  }
}

// (tile, charge) entries of the spreading lists, in charge order
__global__ __launch_bounds__(256) void k_tile_emit(size_t n, const int4 *__restrict__ ms,
                                                   const int *__restrict__ off, GeoParams g,
                                                   unsigned *__restrict__ tkey,
                                                   int *__restrict__ tval)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int4 u = ms[s];
  const int lo0 = u.x / g.ts[0], hi0 = (u.x + (int)g.Bn - 1) / g.ts[0];
  const int lo1 = u.y / g.ts[1], hi1 = (u.y + (int)g.Bn - 1) / g.ts[1];
  const int lo2 = u.z / g.ts[2], hi2 = (u.z + (int)g.Bn - 1) / g.ts[2];
  int o = off[s];
  for (int a = lo0; a <= hi0; a++)
    for (int b = lo1; b <= hi1; b++)
      for (int c = lo2; c <= hi2; c++) {
        tkey[o] = (unsigned)(((long)a * g.nt[1] + b) * g.nt[2] + c);
        tval[o] = (int)s;
        o++;
      }
}

// interpolation tile of each charge (the tile holding its window start)
__global__ __launch_bounds__(256) void k_itile_key(size_t n, const int4 *__restrict__ ms,
                                                   GeoParams g, unsigned *__restrict__ key,
                                                   int *__restrict__ val)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int4 u = ms[s];
  key[s] = (unsigned)(((long)(u.x / g.TI) * g.nti[1] + u.y / g.TI) * g.nti[2] + u.z / g.TI);
  val[s] = (int)s;
}

// start[t] = number of entries with key < t, t = 0..nt
__global__ __launch_bounds__(256) void k_starts(long nt, const unsigned *__restrict__ skey,
                                                size_t m, int *__restrict__ start)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nt) return;
  start[t] = (int)lower_bound_d(skey, m, (unsigned)t);
}

// neighbour-list slices of kNlSlice targets: a slice's per-lane length
// comes from its longest list; the scan input is 64 * that length (the last
// slot stays 0 for the total)
__global__ __launch_bounds__(256) void k_slice_len(size_t n, const int *__restrict__ count,
                                                   int *__restrict__ sliceLen,
                                                   long *__restrict__ sliceWords)
{
  const size_t sl = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nSlices = (n + kNlSlice - 1) / kNlSlice;
  if (sl > nSlices) return;
  if (sl == nSlices) {
    sliceWords[sl] = 0;
    return;
  }
  int m = 0;
  const size_t b = kNlSlice * sl, e = (b + kNlSlice < n) ? b + kNlSlice : n;
  for (size_t s = b; s < e; s++) m = max(m, count[s]);
  const int ml = nlist_lane_len(m);
  sliceLen[sl] = ml;
  sliceWords[sl] = 64L * ml;
}

__global__ __launch_bounds__(256) void k_sort_coords(size_t n, const int *__restrict__ perm,
                                                     const double *__restrict__ coord,
                                                     double4 *__restrict__ xs)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const size_t i = (size_t)perm[s];
  xs[s] = make_double4(coord[3 * i], coord[3 * i + 1], coord[3 * i + 2], 0.0);
}

// per cell: count * (charges in the stencil around it), as double
__global__ __launch_bounds__(256) void k_stencil_pairs(long ncell, int nb0, int nb1, int nb2,
                                                       int nN, const int2 *__restrict__ cellSE,
                                                       double *__restrict__ out)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int n0 = cellSE[c].y;
  double sum = 0;
  if (n0 > 0) {
    const int g2 = (int)(c % nb2), g1 = (int)((c / nb2) % nb1), g0 = (int)(c / ((long)nb1 * nb2));
    for (int dx = -nN; dx <= nN; dx++)
      for (int dy = -nN; dy <= nN; dy++)
        for (int dz = -nN; dz <= nN; dz++) {
          const int a = g0 + dx, b = g1 + dy, e = g2 + dz;
          if (a < 0 || b < 0 || e < 0 || a >= nb0 || b >= nb1 || e >= nb2) continue;
          sum += (double)cellSE[((long)a * nb1 + b) * nb2 + e].y;
        }
  }
  out[c] = (double)n0 * sum;
}

static dim3 grid_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

static int key_bits(unsigned long maxKey)
{
  int b = 1;
  while (b < 64 && (maxKey >> b) != 0) b++;
  return b;
}

// Temporary device storage of a phase, kept until the phase's final
// stream synchronisation: the pool is shared with other streams, so memory
// returns to it only when the work that uses it has finished.
using Keep = std::deque<DevBuf>;
static void keep_or_sync(DevBuf &tmp, Keep *keep, hipStream_t s)
{
  if (keep) {
    keep->emplace_back();
    keep->back().swap(tmp);
  } else {
    NPME_HIP_CHECK(hipStreamSynchronize(s));  // tmp is released on return
  }
}

// stable sort of (key, value) pairs; temporary storage from the pool
template <typename K>
static void sort_pairs(const K *kin, K *kout, const int *vin, int *vout, size_t m, int bits,
                       hipStream_t s, Keep *keep = nullptr)
{
  size_t bytes = 0;
  NPME_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, m, 0, bits, s));
  DevBuf tmp(bytes);
  NPME_HIP_CHECK(rocprim::radix_sort_pairs(tmp.p, bytes, kin, kout, vin, vout, m, 0, bits, s));
  keep_or_sync(tmp, keep, s);
}

template <typename T, typename S>
static void exclusive_scan(const T *in, S *out, size_t m, hipStream_t s, Keep *keep = nullptr)
{
  size_t bytes = 0;
  NPME_HIP_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, (S)0, m, rocprim::plus<S>(), s));
  DevBuf tmp(bytes);
  NPME_HIP_CHECK(rocprim::exclusive_scan(tmp.p, bytes, in, out, (S)0, m, rocprim::plus<S>(), s));
  keep_or_sync(tmp, keep, s);
}

static void throw_geo_err(int e)
{
  if (e & GEO_ERR_CELL) throw Error{NPME_ERR_ARG, "coordinate outside the cell box"};
  if (e & GEO_ERR_WINDOW) throw Error{NPME_ERR_GRID, "spline window outside the compact grid"};
  if (e & GEO_ERR_FIXED) throw Error{NPME_ERR_ARG, "position outside its cluster"};
}

static void check_err(const int *d_err, hipStream_t s)
{
  int e = 0;
  NPME_HIP_CHECK(hipMemcpyAsync(&e, d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  if (e & GEO_ERR_CELL) throw Error{NPME_ERR_ARG, "coordinate outside the cell box"};
  if (e & GEO_ERR_WINDOW) throw Error{NPME_ERR_GRID, "spline window outside the compact grid"};
  if (e & GEO_ERR_FIXED) throw Error{NPME_ERR_ARG, "position outside its cluster"};
}

void geometry_cells_device(const GeoParams &g, size_t n, const double *d_coord, int *d_perm,
                           int4 *d_cellOf, int *d_cellSE, hipStream_t s)
{
  DevBuf key(n * sizeof(unsigned long)), skey(n * sizeof(unsigned long)), idx(n * sizeof(int)),
      err(sizeof(int)), gcell(n * sizeof(int4));
  NPME_HIP_CHECK(hipMemsetAsync(err.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_cell_keys, grid_for(n), dim3(256), 0, s, n, d_coord, g,
                     (unsigned long *)key.p, (int *)idx.p, (int4 *)gcell.p, (int *)err.p);
  NPME_HIP_CHECK(hipGetLastError());
  // a charge outside the box gets key 0 and cell 0, so the sort and the
  // ranges stay in bounds; the error is raised after the phase
  Keep keep;
  const unsigned long Nc = g.nP * g.nP * g.nP;
  const unsigned long maxKey = g.nA[0] * g.nA[1] * g.nA[2] * Nc;
  sort_pairs((const unsigned long *)key.p, (unsigned long *)skey.p, (const int *)idx.p, d_perm,
             n, key_bits(maxKey), s, &keep);
  const long ncell = (long)g.nb[0] * g.nb[1] * g.nb[2];
  hipLaunchKernelGGL(k_cell_ranges, grid_for(ncell), dim3(256), 0, s, ncell, g,
                     (const unsigned long *)skey.p, n, d_cellSE);
  NPME_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_gather_cells, grid_for(n), dim3(256), 0, s, n, (const int *)d_perm,
                     (const int4 *)gcell.p, d_cellOf);
  NPME_HIP_CHECK(hipGetLastError());
  check_err((const int *)err.p, s);  // synchronises the stream
}

void geometry_charges_device(const GeoParams &g, bool fp64, size_t n, const double *d_coord,
                             const int *d_perm, const int4 *d_cellOf, void *d_src, void *d_srcRes,
                             int4 *d_ms, void *d_frac, DevBuf &tileStart, DevBuf &tileList,
                             DevBuf &itileStart, DevBuf &itileList, size_t &nTileEntries,
                             hipStream_t s)
{
  DevBuf ntile(n * sizeof(int)), off((n + 1) * sizeof(int)), err(sizeof(int));
  NPME_HIP_CHECK(hipMemsetAsync(err.p, 0, sizeof(int), s));
  if (fp64)
    hipLaunchKernelGGL(k_charge_setup<true>, grid_for(n), dim3(256), 0, s, n, d_perm, d_coord,
                       d_cellOf, g, d_src, (float4 *)d_srcRes, d_ms, d_frac, (int *)ntile.p,
                       (int *)err.p);
  else
    hipLaunchKernelGGL(k_charge_setup<false>, grid_for(n), dim3(256), 0, s, n, d_perm, d_coord,
                       d_cellOf, g, d_src, (float4 *)d_srcRes, d_ms, d_frac, (int *)ntile.p,
                       (int *)err.p);
  NPME_HIP_CHECK(hipGetLastError());
  Keep keep;

  // spreading lists: emit (tile, charge) in charge order, stable sort by
  // tile.  The error flag comes back with the list size, before any kernel
  // uses the windows.
  exclusive_scan((const int *)ntile.p, (int *)off.p, n, s, &keep);
  int last[3] = {0, 0, 0};
  NPME_HIP_CHECK(hipMemcpyAsync(&last[0], (int *)off.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipMemcpyAsync(&last[1], (int *)ntile.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipMemcpyAsync(&last[2], err.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  throw_geo_err(last[2]);
  const size_t m = (size_t)last[0] + (size_t)last[1];
  nTileEntries = m;
  {
    DevBuf tk(m * sizeof(unsigned)), tv(m * sizeof(int)), sk(m * sizeof(unsigned));
    tileList.reset(m * sizeof(int));
    hipLaunchKernelGGL(k_tile_emit, grid_for(n), dim3(256), 0, s, n, d_ms, (const int *)off.p, g,
                       (unsigned *)tk.p, (int *)tv.p);
    NPME_HIP_CHECK(hipGetLastError());
    const long ntiles = (long)g.nt[0] * g.nt[1] * g.nt[2];
    sort_pairs((const unsigned *)tk.p, (unsigned *)sk.p, (const int *)tv.p, (int *)tileList.p, m,
               key_bits((unsigned long)ntiles), s, &keep);
    tileStart.reset((ntiles + 1) * sizeof(int));
    hipLaunchKernelGGL(k_starts, grid_for(ntiles + 1), dim3(256), 0, s, ntiles,
                       (const unsigned *)sk.p, m, (int *)tileStart.p);
    NPME_HIP_CHECK(hipGetLastError());
    for (DevBuf *b : {&tk, &tv, &sk}) {
      keep.emplace_back();
      keep.back().swap(*b);
    }
  }
  // interpolation lists: one tile per charge
  {
    DevBuf ik(n * sizeof(unsigned)), iv(n * sizeof(int)), sk(n * sizeof(unsigned));
    itileList.reset(n * sizeof(int));
    hipLaunchKernelGGL(k_itile_key, grid_for(n), dim3(256), 0, s, n, d_ms, g, (unsigned *)ik.p,
                       (int *)iv.p);
    NPME_HIP_CHECK(hipGetLastError());
    const long nit = (long)g.nti[0] * g.nti[1] * g.nti[2];
    sort_pairs((const unsigned *)ik.p, (unsigned *)sk.p, (const int *)iv.p, (int *)itileList.p, n,
               key_bits((unsigned long)nit), s, &keep);
    itileStart.reset((nit + 1) * sizeof(int));
    hipLaunchKernelGGL(k_starts, grid_for(nit + 1), dim3(256), 0, s, nit, (const unsigned *)sk.p,
                       n, (int *)itileStart.p);
    NPME_HIP_CHECK(hipGetLastError());
    NPME_HIP_CHECK(hipStreamSynchronize(s));  // before the scratch returns to the pool
  }
}

__global__ __launch_bounds__(256) void k_widen(long m, const int *__restrict__ in,
                                               long *__restrict__ out)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = in[i];
}

__global__ __launch_bounds__(256) void k_narrow(long m, const long *__restrict__ in,
                                                int *__restrict__ out)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = (int)in[i];
}

// Cell-row segments of the direct sum without a neighbour list (k_direct):
// per block of 8 sorted targets, the contiguous source ranges of the
// cluster-rows of cells its cutoff sphere can reach, with the same FP64
// decisions (no FMA contraction) as the host build, so the segments are
// identical; COUNT writes the number per block, FILL the segments
// (sB, sE, ax, ay << 16 | az) at segStart[b], and the largest cluster
// offset from a target's cluster goes to maxOff (atomic max).
template <bool FILL>
__global__ __launch_bounds__(256) void k_cell_segments(long nBlocks, size_t n, GeoParams g,
                                                       double Rm,
                                                       const double4 *__restrict__ xs,
                                                       const int4 *__restrict__ cellOf,
                                                       const int2 *__restrict__ cellSE,
                                                       int *__restrict__ count,
                                                       const int *__restrict__ segStart,
                                                       int4 *__restrict__ segs,
                                                       int *__restrict__ maxOff)
{
#pragma clang fp contract(off)
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nBlocks) return;
  const int nP = (int)g.nP;
  const double csB = g.csB;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  int alo[3] = {1 << 30, 1 << 30, 1 << 30}, ahi[3] = {-1, -1, -1};
  const size_t s1 = min(n, (size_t)(8 * b + 8));
  for (size_t s = 8 * b; s < s1; s++) {
    const double4 p = xs[s];
    const int4 c = cellOf[s];
    const double px[3] = {p.x, p.y, p.z};
    const int cc[3] = {c.x, c.y, c.z};
    for (int d = 0; d < 3; d++) {
      const double x = px[d] + (-g.mn[d]);
      lo[d] = fmin(lo[d], x);
      hi[d] = fmax(hi[d], x);
      alo[d] = min(alo[d], cc[d] / nP);
      ahi[d] = max(ahi[d], cc[d] / nP);
    }
  }
  int off = 1, k = 0;
  const int start = FILL ? segStart[b] : 0;
  const int gx0 = max(0, (int)floor((lo[0] - Rm) / csB));
  const int gx1 = min(g.nb[0] - 1, (int)floor((hi[0] + Rm) / csB));
  const int gy0 = max(0, (int)floor((lo[1] - Rm) / csB));
  const int gy1 = min(g.nb[1] - 1, (int)floor((hi[1] + Rm) / csB));
  for (int gx = gx0; gx <= gx1; gx++) {
    const double gapx = fmax(0.0, fmax(gx * csB - hi[0], lo[0] - (gx + 1) * csB));
    if (gapx >= Rm) continue;
    for (int gy = gy0; gy <= gy1; gy++) {
      const double gapy = fmax(0.0, fmax(gy * csB - hi[1], lo[1] - (gy + 1) * csB));
      const double rem = Rm * Rm - gapx * gapx - gapy * gapy;
      if (rem <= 0) continue;
      const double sr = sqrt(rem);
      const int zlo = max(0, (int)floor((lo[2] - sr) / csB));
      const int zhi = min(g.nb[2] - 1, (int)floor((hi[2] + sr) / csB));
      const long row = ((long)gx * g.nb[1] + gy) * g.nb[2];
      for (int az = zlo / nP; az <= zhi / nP; az++) {
        const int z0 = max(zlo, az * nP), z1 = min(zhi, az * nP + nP - 1);
        const int sB = cellSE[row + z0].x;
        const int sE = cellSE[row + z1].x + cellSE[row + z1].y;
        if (sE > sB) {
          if (FILL) segs[start + k] = make_int4(sB, sE, gx / nP, ((gy / nP) << 16) | az);
          k++;
          const int a[3] = {gx / nP, gy / nP, az};
          for (int d = 0; d < 3; d++) off = max(off, max(abs(a[d] - alo[d]), abs(a[d] - ahi[d])));
        }
      }
    }
  }
  if (!FILL) {
    count[b] = k;
    atomicMax(maxOff, off);
  }
}

long cell_segments_device(const GeoParams &g, size_t n, double Rdir, const double4 *d_xs,
                          const int4 *d_cellOf, const int *d_cellSE, int *d_segStart,
                          DevBuf &segs, int &maxOff, hipStream_t s)
{
  const long nBlocks = ((long)n + 7) / 8;
  const double Rm = Rdir * (1.0 + 1.0e-9) + 1.0e-12;
  DevBuf cnt((nBlocks + 1) * sizeof(int)), off(sizeof(int));
  const int one = 1;
  NPME_HIP_CHECK(hipMemcpyAsync(off.p, &one, sizeof(int), hipMemcpyHostToDevice, s));
  NPME_HIP_CHECK(hipMemsetAsync((int *)cnt.p + nBlocks, 0, sizeof(int), s));
  const dim3 grid((unsigned)((nBlocks + 255) / 256));
  hipLaunchKernelGGL(k_cell_segments<false>, grid, dim3(256), 0, s, nBlocks, n, g, Rm, d_xs, d_cellOf,
                     (const int2 *)d_cellSE, (int *)cnt.p, (const int *)nullptr, (int4 *)nullptr,
                     (int *)off.p);
  NPME_HIP_CHECK(hipGetLastError());
  // segment counts in long, so an overflow of the int offsets is seen
  DevBuf wide((nBlocks + 1) * sizeof(long)), start((nBlocks + 1) * sizeof(long));
  hipLaunchKernelGGL(k_widen, grid_for(nBlocks + 1), dim3(256), 0, s, nBlocks + 1,
                     (const int *)cnt.p, (long *)wide.p);
  NPME_HIP_CHECK(hipGetLastError());
  const long total = scan_words_device((const long *)wide.p, (long *)start.p, nBlocks + 1, s);
  NPME_HIP_CHECK(hipMemcpyAsync(&maxOff, off.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  if (total > 0x7fffffffL) throw Error{NPME_ERR_NOMEM, "too many direct-sum segments"};
  hipLaunchKernelGGL(k_narrow, grid_for(nBlocks + 1), dim3(256), 0, s, nBlocks + 1,
                     (const long *)start.p, d_segStart);
  segs.reset(std::max<size_t>((size_t)total, 1) * sizeof(int4));
  hipLaunchKernelGGL(k_cell_segments<true>, grid, dim3(256), 0, s, nBlocks, n, g, Rm, d_xs, d_cellOf,
                     (const int2 *)d_cellSE, (int *)nullptr, (const int *)d_segStart, (int4 *)segs.p,
                     (int *)nullptr);
  NPME_HIP_CHECK(hipGetLastError());
  NPME_HIP_CHECK(hipStreamSynchronize(s));  // before the scratch returns to the pool
  return total;
}

long scan_words_device(const long *d_in, long *d_out, size_t m, hipStream_t s)
{
  exclusive_scan(d_in, d_out, m, s);
  long total = 0;
  NPME_HIP_CHECK(hipMemcpyAsync(&total, d_out + (m - 1), sizeof(long), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  return total;
}

size_t nlist_slices_device(size_t n, const int *d_count, int *d_sliceLen, long *d_sliceStart,
                           hipStream_t s)
{
  const size_t nSlices = (n + kNlSlice - 1) / kNlSlice;
  DevBuf words((nSlices + 1) * sizeof(long));
  hipLaunchKernelGGL(k_slice_len, grid_for(nSlices + 1), dim3(256), 0, s, n, d_count, d_sliceLen,
                     (long *)words.p);
  NPME_HIP_CHECK(hipGetLastError());
  exclusive_scan((const long *)words.p, d_sliceStart, nSlices + 1, s);
  long total = 0;
  NPME_HIP_CHECK(hipMemcpyAsync(&total, d_sliceStart + nSlices, sizeof(long), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  return (size_t)total;
}

// per-axis minimum and maximum of the coordinates: one partial per
// workgroup, then one workgroup over the partials.  Exact; a NaN
// coordinate makes its axis' bounds NaN.
__device__ __forceinline__ double nan_min(double a, double b) { return (a != a || a < b) ? a : b; }
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a || b < a) ? a : b; }

constexpr int BOUNDS_WG = 1024;
__global__ __launch_bounds__(256) void k_coord_bounds(size_t n, const double *__restrict__ coord,
                                                      double *__restrict__ part)
{
  __shared__ double sh[6][256];
  double v[6] = {HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    for (int d = 0; d < 3; d++) {
      const double x = coord[3 * i + d];
      v[d] = nan_min(x, v[d]);
      v[3 + d] = nan_max(x, v[3 + d]);
    }
  for (int d = 0; d < 6; d++) sh[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; d++) {
        sh[d][threadIdx.x] = nan_min(sh[d][threadIdx.x + w], sh[d][threadIdx.x]);
        sh[3 + d][threadIdx.x] = nan_max(sh[3 + d][threadIdx.x + w], sh[3 + d][threadIdx.x]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_bounds_final(int np, const double *__restrict__ part,
                                                      double *__restrict__ out)
{
  __shared__ double sh[6][256];
  double v[6] = {HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int b = threadIdx.x; b < np; b += blockDim.x)
    for (int d = 0; d < 3; d++) {
      v[d] = nan_min(part[6 * b + d], v[d]);
      v[3 + d] = nan_max(part[6 * b + 3 + d], v[3 + d]);
    }
  for (int d = 0; d < 6; d++) sh[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; d++) {
        sh[d][threadIdx.x] = nan_min(sh[d][threadIdx.x + w], sh[d][threadIdx.x]);
        sh[3 + d][threadIdx.x] = nan_max(sh[3 + d][threadIdx.x + w], sh[3 + d][threadIdx.x]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = sh[threadIdx.x][0];
}

size_t coord_bounds_scratch(size_t n)
{
  return (6 * std::min<size_t>(BOUNDS_WG, (n + 255) / 256) + 6) * sizeof(double);
}

void coord_bounds_launch(size_t n, const double *d_coord, double *d_scratch, hipStream_t s)
{
  const int nb = (int)std::min<size_t>(BOUNDS_WG, (n + 255) / 256);
  hipLaunchKernelGGL(k_coord_bounds, dim3(nb), dim3(256), 0, s, n, d_coord, d_scratch);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, s, nb, (const double *)d_scratch,
                     d_scratch + 6 * nb);
  NPME_HIP_CHECK(hipGetLastError());
}

void coord_bounds_device(size_t n, const double *d_coord, double mn[3], double mx[3], hipStream_t s)
{
  const int nb = (int)std::min<size_t>(BOUNDS_WG, (n + 255) / 256);
  DevBuf scratch(coord_bounds_scratch(n));
  coord_bounds_launch(n, d_coord, (double *)scratch.p, s);
  const double *out = (const double *)scratch.p + 6 * nb;
  double v[6];
  NPME_HIP_CHECK(hipMemcpyAsync(v, out, sizeof(v), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  for (int d = 0; d < 3; d++) {
    mn[d] = v[d];
    mx[d] = v[3 + d];
  }
}

void sorted_coords_device(size_t n, const int *d_perm, const double *d_coord, double4 *d_xs,
                          hipStream_t s)
{
  hipLaunchKernelGGL(k_sort_coords, grid_for(n), dim3(256), 0, s, n, d_perm, d_coord, d_xs);
  NPME_HIP_CHECK(hipGetLastError());
}

double stencil_candidates_device(const int nb[3], int nNeigh, const int *d_cellSE, hipStream_t s)
{
  const long ncell = (long)nb[0] * nb[1] * nb[2];
  DevBuf per(ncell * sizeof(double)), tot(sizeof(double));
  hipLaunchKernelGGL(k_stencil_pairs, grid_for(ncell), dim3(256), 0, s, ncell, nb[0], nb[1], nb[2],
                     nNeigh, (const int2 *)d_cellSE, (double *)per.p);
  NPME_HIP_CHECK(hipGetLastError());
  size_t bytes = 0;
  NPME_HIP_CHECK(rocprim::reduce(nullptr, bytes, (const double *)per.p, (double *)tot.p, 0.0,
                                 (size_t)ncell, rocprim::plus<double>(), s));
  DevBuf tmp(bytes);
  NPME_HIP_CHECK(rocprim::reduce(tmp.p, bytes, (const double *)per.p, (double *)tot.p, 0.0,
                                 (size_t)ncell, rocprim::plus<double>(), s));
  double v = 0;
  NPME_HIP_CHECK(hipMemcpyAsync(&v, tot.p, sizeof(double), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  return v;
}

}  // namespace npme

namespace npme {
__global__ __launch_bounds__(256) void k_sentinel_near(size_t n, const int4 *__restrict__ src, int4 v,
                                                      int thr, int *__restrict__ hit)
{
  const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int4 p = src[s];
  const int dx = (int)((unsigned)p.x - (unsigned)v.x), dy = (int)((unsigned)p.y - (unsigned)v.y),
            dz = (int)((unsigned)p.z - (unsigned)v.z);
  if (abs(dx) < thr && abs(dy) < thr && abs(dz) < thr) hit[0] = 1;
}

bool sentinel_near_any(size_t n, const void *d_src, int4 v, int thr, hipStream_t s)
{
  DevBuf hit(sizeof(int));
  int h = 0;
  NPME_HIP_CHECK(hipMemsetAsync(hit.p, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_sentinel_near, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                     (const int4 *)d_src, v, thr, (int *)hit.p);
  NPME_HIP_CHECK(hipGetLastError());
  NPME_HIP_CHECK(hipMemcpyAsync(&h, hit.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  return h != 0;
}
}  // namespace npme
