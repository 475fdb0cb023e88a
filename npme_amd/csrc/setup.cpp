// Per-geometry set-up and the CalcV driver.
//
// Set-up (host FP64, integer indices bit-identical to the reference):
//   * keywords -> kernel, DM fit, FFT grid, B-spline order
//   * embedded cell partition (cell = Rdir/nNeigh inside clusters of
//     nCellClust1D^3 cells) and the stable sort of the charges by cell,
//     src/PartitionEmbeddedBox.cpp:266-455, src/PartitionBox.cpp:311-351
//   * PME-corrected box and spline windows mStart per charge,
//     src/RecSumSupportFunctions.cpp:276-296, .h:123-149
//   * per-tile charge lists for the atomic-free spreading
//   * kernel grid G = F(|k|) lam^2 / (N1 N2 N3): real-space compact grid of
//     f_LR in FP64 on the device, then an even-extension DFT per axis
//     (src/RecSumGrid.cpp:45-651), lam^2 from src/RecSumV.cpp:565-581
// Device layout: the compact grid is stored in m-order, u = m + N/4
// (a fixed cyclic shift of the reference's DFT order p = m mod N/2), so a
// spline window is always a contiguous box and the zero padding of the
// full grid is the tail [N/2, N) of every axis.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <sstream>

#include "npme_internal.h"

namespace npme {

static void kw_error(const std::string &m) { throw Error{NPME_ERR_KEYWORD, m}; }

static double ms_since(std::chrono::steady_clock::time_point &t)
{
  const auto now = std::chrono::steady_clock::now();
  const double ms = std::chrono::duration<double, std::milli>(now - t).count();
  t = now;
  return ms;
}

template <typename T> static T *dalloc(Handle &h, size_t count)
{
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  void *p = dev_alloc(bytes);
  if (!p) {
    std::ostringstream os;
    os << "device allocation of " << bytes << " bytes failed";
    throw Error{NPME_ERR_NOMEM, os.str()};
  }
  h.deviceBytes += dev_block_bytes(bytes);
  return (T *)p;
}

static void drop_graph(Handle &h)
{
  if (h.calcvExec) (void)hipGraphExecDestroy(h.calcvExec);
  h.calcvExec = nullptr;
  h.graphQ = h.graphV = nullptr;
  h.graphCalls = 0;
}

void handle_free_device(Handle &h, bool keepG)
{
  drop_graph(h);
  void *ptrs[] = {h.d_itileStart, h.d_itileList, h.d_perm, h.d_src, h.d_srcRes, h.d_blockBox, h.d_segStart, h.d_nlist,
                  h.d_sliceStart, h.d_sliceLen, h.d_qi, h.d_cell, h.d_cellStart, h.d_ms, h.d_frac,
                  h.d_tileStart, h.d_tileList, h.d_grid, h.d_buf1, h.d_buf2,
                  h.d_tables, h.d_Vdir, h.d_V, h.d_qin, h.d_coordS, h.d_xs,
                  h.d_qs, h.d_qsi, h.d_utab, h.d_cellFlags};
  // queued work of this handle may still read its buffers
  if (h.stream) (void)hipStreamSynchronize(h.stream);
  if (h.stream2) (void)hipStreamSynchronize(h.stream2);
  for (void *p : ptrs) dev_free(p);
  if (!keepG) {
    dev_free(h.d_G);
    h.d_G = nullptr;
    h.gValid = false;
    h.gBytes = 0;
  }
  h.d_perm = nullptr;
  h.d_src = h.d_srcRes = h.d_qi = h.d_blockBox = nullptr;
  h.d_qs = h.d_qsi = nullptr;
  h.d_segStart = nullptr;
  h.d_itileStart = h.d_itileList = nullptr;
  h.d_nlist = nullptr;
  h.d_sliceStart = nullptr;
  h.d_sliceLen = nullptr;
  h.useNlist = false;
  h.nlistEntries = 0;
  h.d_cell = nullptr;
  h.d_cellStart = nullptr;
  h.d_ms = nullptr;
  h.d_frac = nullptr;
  h.d_tileStart = h.d_tileList = nullptr;
  h.d_grid = h.d_buf1 = h.d_buf2 = nullptr;
  h.d_tw[0] = h.d_tw[1] = h.d_tw[2] = nullptr;
  h.d_tables = nullptr;
  h.d_Vdir = h.d_V = h.d_qin = h.d_coordS = nullptr;
  h.d_xs = nullptr;
  h.d_utab = nullptr;
  h.d_cellFlags = nullptr;
  h.coordStaged = h.stagedBoundsReady = false;
  h.deviceBytes = h.d_G ? h.gBytes : 0;
}

static void ensure_streams(Handle &h)
{
  NPME_HIP_CHECK(hipGetDevice(&h.device));
  if (!h.stream) NPME_HIP_CHECK(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking));
  if (!h.stream2) NPME_HIP_CHECK(hipStreamCreateWithFlags(&h.stream2, hipStreamNonBlocking));
  if (!h.evFork) NPME_HIP_CHECK(hipEventCreateWithFlags(&h.evFork, hipEventDisableTiming));
  if (!h.evJoin) NPME_HIP_CHECK(hipEventCreateWithFlags(&h.evJoin, hipEventDisableTiming));
}

void handle_stage_coords(Handle &h, size_t n, const double *coord)
{
  if (n < 2 || !coord) throw Error{NPME_ERR_ARG, "staged positions: nCharge must be at least 2"};
  ensure_streams(h);
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  handle_free_device(h, true);
  h.ready = false;
  h.d_coordS = dalloc<double>(h, 3 * n);
  NPME_HIP_CHECK(hipMemcpyAsync(h.d_coordS, coord, 3 * n * sizeof(double), hipMemcpyHostToDevice,
                                h.stream));
  h.stageBuf.reset(coord_bounds_scratch(n));
  coord_bounds_launch(n, h.d_coordS, (double *)h.stageBuf.p, h.stream);
  h.stagedN = n;
  h.coordStaged = true;
  h.stagedBoundsReady = false;
}

void handle_staged_bounds(Handle &h, double mn[3], double mx[3])
{
  if (!h.coordStaged) throw Error{NPME_ERR_ARG, "no staged positions"};
  if (!h.stagedBoundsReady) {
    double v[6];
    const double *out = (const double *)((const char *)h.stageBuf.p + h.stageBuf.bytes) - 6;
    NPME_HIP_CHECK(hipMemcpyAsync(v, out, sizeof(v), hipMemcpyDeviceToHost, h.stream));
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
    for (int d = 0; d < 3; d++) {
      h.stagedMn[d] = v[d];
      h.stagedMx[d] = v[3 + d];
    }
    h.stagedBoundsReady = true;
  }
  for (int d = 0; d < 3; d++) {
    mn[d] = h.stagedMn[d];
    mx[d] = h.stagedMx[d];
  }
}

// ---------------------------------------------------------------- cells
// cell index (size_t)(x / cs) with the 1e-14 upper-edge correction
// (src/PartitionBox.cpp:311-351)
static unsigned long coord2cell(double x, double cs, unsigned long ncell)
{
  const double tol = 1.0e-14;
  if (x < -tol) throw Error{NPME_ERR_ARG, "coordinate below the cell box"};
  unsigned long v = (unsigned long)(x / cs);
  if (v > ncell) throw Error{NPME_ERR_ARG, "coordinate above the cell box"};
  if (v == ncell) {
    if (std::fabs(x - v * cs) < tol) v--;
    else throw Error{NPME_ERR_ARG, "coordinate above the cell box"};
  }
  return v;
}

std::vector<double> twiddles(int N)
{
  const double PI = 3.14159265358979323846;
  std::vector<double> t(2 * (size_t)N);
  for (int k = 0; k < N; k++) {
    const double a = 2.0 * PI * (double)k / (double)N;
    t[2 * k] = std::cos(a);
    t[2 * k + 1] = -std::sin(a);
  }
  return t;
}

// top-hat smoothing of the Fourier extension (src/RecSumGrid.cpp:1124-1231)
static double hramp(double x, double a)
{
  const double erf_max = 6.0;
  if (x < -1.0) return 0.0;
  if (x > 1.0) return 1.0;
  const double z = erf_max / a;
  const double xMax = z / std::sqrt(1.0 + z * z);
  if (x >= xMax) return 1.0;
  if (x <= -xMax) return 0.0;
  return 0.5 * (1.0 + std::erf(a * x / std::sqrt(1.0 - x * x)));
}

static double tsmooth(double x, double X, double Xmax, double a)
{
  const double eps = 1.0e-12 * Xmax;
  if (x < -Xmax - eps || x > Xmax + eps)
    throw Error{NPME_ERR_GRID, "top-hat argument outside the Fourier extension"};
  const double del = (Xmax - X) * 0.5;
  if (x < -X) return hramp((x + X + del) / del, a);
  if (x <= X) return 1.0;
  return hramp((-x + X + del) / del, a);
}

// per-axis tables of the full (translated) kernel grid
// (NPME_RecSumGrid_FullTrans_Func, src/RecSumGrid.cpp:663-794): sample n at
// s L/N with s = n, or n - N for n >= N/2 (src/RecSumGrid.cpp:1024-1042);
// its second image and the top-hat weights as NPME_CalcOverlapSmoothFunc
// (src/RecSumGrid.cpp:1193-1231); both image coordinates shifted by D after
// the weights are taken (src/RecSumGrid.cpp:77-100).  Per axis
// [x1 + D | x2 + D | T1 | T2].
std::vector<double> fulltrans_tables(const long N[3], const double ta[3], const double del[3],
                                     const double X[3], const double D[3])
{
  std::vector<double> tab;
  for (int d = 0; d < 3; d++) {
    const long n = N[d];
    const double Xmax = X[d] + del[d], Xt = 2 * Xmax - X[d], dx = 2 * Xmax / n;
    std::vector<double> x1(n), x2(n), T1(n), T2(n);
    for (long i = 0; i < n; i++) {
      const double x = (double)(i >= n / 2 ? i - n : i) * dx;
      if (x > X[d]) {
        T1[i] = tsmooth(x, X[d], Xt, ta[d]);
        x2[i] = x - 2 * Xmax;
        T2[i] = tsmooth(x2[i], X[d], Xt, ta[d]);
      } else if (x > -X[d]) {
        T1[i] = 1.0;
        x2[i] = 1.0;  // weight 0: never evaluated
        T2[i] = 0.0;
      } else {
        T1[i] = tsmooth(x, X[d], Xt, ta[d]);
        x2[i] = x + 2 * Xmax;
        T2[i] = tsmooth(x2[i], X[d], Xt, ta[d]);
      }
      x1[i] = x + D[d];
      x2[i] += D[d];
    }
    tab.insert(tab.end(), x1.begin(), x1.end());
    tab.insert(tab.end(), x2.begin(), x2.end());
    tab.insert(tab.end(), T1.begin(), T1.end());
    tab.insert(tab.end(), T2.begin(), T2.end());
  }
  return tab;
}

// min |r + R0| over the box of edges l centred at the origin
// (NPME_CalcMinDistance_R0_r, src/SupportFunctions.cpp:687-718)
double min_distance_r0(const double R0[3], const double l[3])
{
  double s = 0;
  for (int d = 0; d < 3; d++) {
    const double m = std::fabs(R0[d]);
    if (!(m < 0.5 * l[d])) s += (0.5 * l[d] - m) * (0.5 * l[d] - m);
  }
  return std::sqrt(s);
}

double2 *fft3d_forward_fp64(const long N[3], const FFTPlan1D plan[3], const void *const tw64[3],
                            double2 *a, double2 *b, hipStream_t s)
{
  const long N0 = N[0], N1 = N[1], N2 = N[2];
  // rows along axis 3, columns along axis 2 (stride N2), columns along axis 1
  const PassGeom g3{(int)N2, (int)N2, (int)N2, N0 * N1, N0 * N1, 0, N2, 1, 0, N2, 1};
  const PassGeom g2{(int)N1, (int)N1, (int)N1, N0 * N2, N2, N1 * N2, 1, N2, N1 * N2, 1, N2};
  const PassGeom g1{(int)N0, (int)N0, (int)N0, N1 * N2, N1 * N2, 0, 1, N1 * N2, 0, 1, N1 * N2};
  fft_pass_complex(true, g3, plan[2], tw64[2], a, b, false, s);
  fft_pass_complex(true, g2, plan[1], tw64[1], b, a, false, s);
  fft_pass_complex(true, g1, plan[0], tw64[0], a, b, false, s);
  return b;
}

// ---------------------------------------------------------------- neighbour list
// Built when the keyword neighborList is not 'off', the cluster offsets fit
// the entry code (nNeigh <= nCellClust1D) and the list fits a quarter of the
// free device memory ('on' lifts the memory limit); otherwise the direct
// sum walks the cell rows of each block (k_direct).
// Expected list size from the cell occupancy alone: the candidate pairs of
// the cell stencil times the fraction of the stencil volume inside the
// cutoff sphere.  When even half of it exceeds the budget (clustered charges
// with thousands per cell), the exact count, a scan of every candidate
// pair, is skipped.
static bool nlist_clearly_too_big(Handle &h, const std::string &mode)
{
  if (mode == "on") return false;
  size_t freeB = 0, totalB = 0;
  NPME_HIP_CHECK(hipMemGetInfo(&freeB, &totalB));
  // no list exceeds kNlSlice ceil(n / kNlSlice) targets of n - 1 entries
  // each, plus fewer than 4 kNlTeam of padding: when even that fits, the
  // estimate is not needed
  const double n = (double)h.n;
  if ((n + kNlSlice) * (n + 4.0 * kNlTeam) * sizeof(unsigned) <= (double)(freeB / 4)) return false;
  const double cand = stencil_candidates_device(h.nb, (int)h.nNeigh, h.d_cellStart, h.stream);
  const double side = (2.0 * h.nNeigh + 1.0) * h.csB, R = h.ker.Rdir;
  const double est = cand * (4.18879020478639098 * R * R * R) / (side * side * side);
  return 0.5 * est * sizeof(unsigned) > (double)(freeB / 4);
}

static void build_neighbor_list(Handle &h, const std::string &mode)
{
  if (nlist_clearly_too_big(h, mode)) return;
  const DirectParams dp = direct_params(h);
  const int nP = (int)h.nCellClust1D;
  std::vector<int> cnt(h.n);
  NlistWork work;
  {
    DevBuf d_count(h.n * sizeof(int));
    launch_nlist(false, dp, h.n, h.d_cell, h.d_cellStart, h.d_xs, nP,
                 (int *)d_count.p, nullptr, nullptr, work, h.stream);
    NPME_HIP_CHECK(hipMemcpyAsync(cnt.data(), d_count.p, h.n * sizeof(int), hipMemcpyDeviceToHost, h.stream));
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  }
  const size_t nSlices = (h.n + kNlSlice - 1) / kNlSlice;
  std::vector<long> sliceStart(nSlices + 1, 0);
  std::vector<int> sliceLen(nSlices, 0);
  for (size_t sl = 0; sl < nSlices; sl++) {
    int m = 0;
    for (size_t s = kNlSlice * sl; s < std::min(h.n, kNlSlice * sl + kNlSlice); s++) m = std::max(m, cnt[s]);
    sliceLen[sl] = nlist_lane_len(m);
    sliceStart[sl + 1] = sliceStart[sl] + 64L * sliceLen[sl];
  }
  const size_t entries = (size_t)sliceStart[nSlices];
  size_t freeB = 0, totalB = 0;
  NPME_HIP_CHECK(hipMemGetInfo(&freeB, &totalB));
  if (mode != "on" && entries * sizeof(unsigned) > freeB / 4) return;
  h.d_nlist = dalloc<unsigned>(h, entries);
  h.d_sliceStart = dalloc<long>(h, nSlices + 1);
  h.d_sliceLen = dalloc<int>(h, nSlices);
  // copies on the handle's stream, never the null stream: a null-stream
  // operation can invalidate a graph capture running on another thread
  NPME_HIP_CHECK(hipMemcpyAsync(h.d_sliceStart, sliceStart.data(), (nSlices + 1) * sizeof(long),
                                hipMemcpyHostToDevice, h.stream));
  NPME_HIP_CHECK(hipMemcpyAsync(h.d_sliceLen, sliceLen.data(), nSlices * sizeof(int),
                                hipMemcpyHostToDevice, h.stream));
  NPME_HIP_CHECK(hipMemsetD32Async((hipDeviceptr_t)h.d_nlist, nlist_pad(h.n, h.fp64 ? 5 : 4), entries, h.stream));
  launch_nlist(true, dp, h.n, h.d_cell, h.d_cellStart, h.d_xs, nP, nullptr, h.d_sliceStart,
               h.d_nlist, work, h.stream);
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  h.nlistEntries = entries;
  h.useNlist = true;
}

// ---------------------------------------------------------------- per-charge set-up
static void *adopt(Handle &h, DevBuf &b)
{
  h.deviceBytes += dev_block_bytes(b.bytes);
  return b.release();
}

// the list walk's fixed point: 2^30 units per cluster edge.  Positions are
// wrapped modulo 2^32, and a listed pair is closer than Rdir <= one cluster
// edge (nNeigh <= nCellClust1D) < 2^31 units, so its int32 difference is exact
constexpr int kListShift = 30;

// the zero-charge source n that pads the neighbour list.  FP64: outside the
// cells' box by two cluster edges on every axis, so it is at least that far
// from every charge.  Mixed (wrapped frame, where no point is far from all
// charges): charge 0's position moved by large odd constants, moved again
// while any charge sits within 2^20 units (1e-3 cluster edges) of it on all
// three axes, so every term it adds is finite (and zero).
static void place_sentinel(Handle &h)
{
  const size_t n = h.n;
  if (h.fp64) {
    const double4 v = make_double4(-2.0 * h.csA, -2.0 * h.csA, -2.0 * h.csA, 0.0);
    NPME_HIP_CHECK(hipMemcpyAsync(static_cast<double4 *>(h.d_src) + n, &v, sizeof(v),
                                  hipMemcpyHostToDevice, h.stream));
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
    return;
  }
  int4 c0;
  NPME_HIP_CHECK(hipMemcpyAsync(&c0, h.d_src, sizeof(c0), hipMemcpyDeviceToHost, h.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  if (!h.useNlist) {
    const int4 z = make_int4(0, 0, 0, 0);
    NPME_HIP_CHECK(hipMemcpyAsync(static_cast<int4 *>(h.d_src) + n, &z, sizeof(z), hipMemcpyHostToDevice,
                                  h.stream));
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));  // before z goes out of scope
    return;
  }
  unsigned x = (unsigned)c0.x, y = (unsigned)c0.y, z = (unsigned)c0.z;
  for (int attempt = 0;; attempt++) {
    x += 0x9E3779B9u;
    y += 0x7F4A7C15u;
    z += 0x85EBCA6Bu;
    const int4 v = make_int4((int)x, (int)y, (int)z, 0);  // w: FP32 charge 0
    if (!sentinel_near_any(n, h.d_src, v, 1 << 20, h.stream)) {
      NPME_HIP_CHECK(hipMemcpyAsync(static_cast<int4 *>(h.d_src) + n, &v, sizeof(v),
                                    hipMemcpyHostToDevice, h.stream));
      NPME_HIP_CHECK(hipStreamSynchronize(h.stream));  // before v goes out of scope
      return;
    }
    if (attempt == 63) throw Error{NPME_ERR_UNSUPPORTED, "no place for the list's padding source"};
  }
}

static int fixed_shift(int maxOff)
{
  // positions lie in [0, 2^S] units of their cluster and offsets reach
  // maxOff clusters, so every pair difference the kernels form stays
  // within (maxOff + 1) 2^S <= 2^30
  int S = 30;
  while (((long)maxOff + 1) << S > (1L << 30)) S--;
  if (S < 16) throw Error{NPME_ERR_UNSUPPORTED, "direct-sum cluster offsets too large"};
  return S;
}

// Device set-up (geometry.hip); returns false when the neighbour list does
// not fit the memory budget (the host set-up then builds the cell-row lists)
// the positions are already on the device (h.d_coordS)
static bool particles_device(Handle &h, GeoParams &g, const std::string &nlmode, bool nlPossible)
{
  const size_t n = h.n;
  hipStream_t s = h.stream;
  const bool f64 = h.fp64;
  const size_t rs = f64 ? 8 : 4;
  const long ncell = (long)h.nb[0] * h.nb[1] * h.nb[2];
  auto t0 = std::chrono::steady_clock::now();
  h.d_perm = dalloc<int>(h, n);
  h.d_cell = dalloc<int4>(h, n);
  h.d_cellStart = dalloc<int>(h, 2 * (size_t)ncell);
  geometry_cells_device(g, n, h.d_coordS, h.d_perm, h.d_cell, h.d_cellStart, s);
  h.d_xs = dalloc<double4>(h, n);
  sorted_coords_device(n, h.d_perm, h.d_coordS, h.d_xs, s);
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  h.setupMs[2] = ms_since(t0);

  // direct-sum structure: the neighbour list (count, slice lengths and
  // offsets, then fill) when it is possible and fits, else the cell-row
  // segments of k_direct
  bool list = nlPossible && !nlist_clearly_too_big(h, nlmode);
  if (list) {
    const DirectParams dp = direct_params(h);
    const size_t nSlices = (n + kNlSlice - 1) / kNlSlice;
    size_t entries = 0;
    h.d_sliceStart = dalloc<long>(h, nSlices + 1);
    h.d_sliceLen = dalloc<int>(h, nSlices);
    NlistWork work;
    {
      DevBuf cnt(n * sizeof(int));
      launch_nlist(false, dp, n, h.d_cell, h.d_cellStart, h.d_xs, (int)h.nCellClust1D, (int *)cnt.p,
                   nullptr, nullptr, work, s);
      entries = nlist_slices_device(n, (const int *)cnt.p, h.d_sliceLen, h.d_sliceStart, s);
    }
    size_t freeB = 0, totalB = 0;
    NPME_HIP_CHECK(hipMemGetInfo(&freeB, &totalB));
    if (nlmode != "on" && entries * sizeof(unsigned) > freeB / 4) {
      list = false;
    } else {
      h.d_nlist = dalloc<unsigned>(h, entries);
      NPME_HIP_CHECK(hipMemsetD32Async((hipDeviceptr_t)h.d_nlist, nlist_pad(n, f64 ? 5 : 4), std::max<size_t>(entries, 1), s));
      launch_nlist(true, dp, n, h.d_cell, h.d_cellStart, h.d_xs, (int)h.nCellClust1D, nullptr,
                   h.d_sliceStart, h.d_nlist, work, s);
      h.nlistEntries = entries;
      h.useNlist = true;
      NPME_HIP_CHECK(hipStreamSynchronize(s));
    }
  }
  int maxOff = 1;
  h.nSegments = 0;
  if (!list) {
    const long nBlocks = ((long)n + 7) / 8;
    h.d_segStart = dalloc<int>(h, nBlocks + 1);
    DevBuf segs;
    h.nSegments = (size_t)cell_segments_device(g, n, h.ker.Rdir, h.d_xs, h.d_cell, h.d_cellStart,
                                               h.d_segStart, segs, maxOff, s);
    h.d_blockBox = (int4 *)adopt(h, segs);
  }
  h.setupMs[3] = ms_since(t0);

  // positions, spline windows and tile lists; the fixed-point scale
  // depends on how far the direct sum's sources reach (list: the adjacent
  // clusters)
  h.fxShift = list ? kListShift : fixed_shift(maxOff);
  g.fxShift = h.fxShift;
  g.fxScale = std::ldexp(1.0, h.fxShift) / h.csA;
  g.globalFrame = list ? 1 : 0;
  // [n + 1]: entry n is the neighbour list's zero-charge sentinel
  h.d_src = dalloc<char>(h, 4 * rs * (n + 1));
  if (!f64) {
    h.d_srcRes = dalloc<float4>(h, n + 1);
    NPME_HIP_CHECK(hipMemsetAsync(static_cast<float4 *>(h.d_srcRes) + n, 0, sizeof(float4), s));
  }
  h.d_frac = dalloc<char>(h, 4 * rs * n);
  h.d_ms = dalloc<int4>(h, n);
  DevBuf ts, tl, its, itl;
  size_t m = 0;
  geometry_charges_device(g, f64, n, h.d_coordS, h.d_perm, h.d_cell, h.d_src, h.d_srcRes, h.d_ms,
                          h.d_frac, ts, tl, its, itl, m, s);
  h.d_tileStart = (int *)adopt(h, ts);
  h.d_tileList = (int *)adopt(h, tl);
  h.d_itileStart = (int *)adopt(h, its);
  h.d_itileList = (int *)adopt(h, itl);
  const int vw = h.complexK ? 8 : 4;
  if (h.complexK) {
    h.d_qi = dalloc<char>(h, rs * (n + 1));
    NPME_HIP_CHECK(hipMemsetAsync((char *)h.d_qi + rs * n, 0, rs, s));
    h.d_qsi = dalloc<char>(h, rs * n);
  }
  h.d_qs = dalloc<char>(h, rs * n);
  h.d_Vdir = dalloc<double>(h, vw * n);
  h.d_V = dalloc<double>(h, vw * n);
  h.d_qin = dalloc<double>(h, (h.complexK ? 2 : 1) * n);
  NPME_HIP_CHECK(hipStreamSynchronize(s));
  place_sentinel(h);
  h.setupMs[4] = ms_since(t0);
  return true;
}

// Host set-up: the same arrays built on the host (FP64, stable sort), plus
// the per-block source segments of the cell-row direct sum
static void particles_host(Handle &h, GeoParams &g, const std::string &nlmode, bool nlPossible,
                           const double *coord)
{
  const size_t n = h.n;
  const unsigned long nP = g.nP;
  const double csA = g.csA, csB = g.csB;
  const double *mn = g.mn;
  const unsigned long *nA = g.nA;
  const double Rdir = h.ker.Rdir;
  const unsigned long Nc = nP * nP * nP;
  std::vector<unsigned long> key(n);
  std::vector<int> gcell(3 * n);
  for (size_t i = 0; i < n; i++) {
    unsigned long a[3], b[3];
    for (int d = 0; d < 3; d++) {
      const double x1 = coord[3 * i + d] + (-mn[d]);
      a[d] = coord2cell(x1, csA, nA[d]);
      const double x2 = x1 - a[d] * csA;
      b[d] = coord2cell(x2, csB, nP);
      gcell[3 * i + d] = (int)(a[d] * nP + b[d]);
    }
    const unsigned long tA = (a[0] * nA[1] + a[1]) * nA[2] + a[2];
    key[i] = tA * Nc + (b[0] * nP + b[1]) * nP + b[2];
  }
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return key[x] < key[y]; });

  const long ncell = (long)h.nb[0] * h.nb[1] * h.nb[2];
  // (start, count) of every cell of the dense cell grid in the sorted order;
  // an empty cell starts where its key would be, so a run of consecutive
  // cells of one cluster row is the range [start(first), start(last)+count)
  std::vector<unsigned long> skey(n);
  for (size_t s = 0; s < n; s++) skey[s] = key[perm[s]];
  std::vector<int> cellSE(2 * (size_t)ncell, 0);
  for (int g0 = 0; g0 < h.nb[0]; g0++)
    for (int g1 = 0; g1 < h.nb[1]; g1++)
      for (int g2 = 0; g2 < h.nb[2]; g2++) {
        const unsigned long a0 = g0 / nP, a1 = g1 / nP, a2 = g2 / nP;
        const unsigned long kc = ((a0 * nA[1] + a1) * nA[2] + a2) * Nc +
                                 ((g0 % nP) * nP + (g1 % nP)) * nP + (g2 % nP);
        auto lo = std::lower_bound(skey.begin(), skey.end(), kc);
        auto hi = std::upper_bound(lo, skey.end(), kc);
        const long c = ((long)g0 * h.nb[1] + g1) * h.nb[2] + g2;
        cellSE[2 * c] = (int)(lo - skey.begin());
        cellSE[2 * c + 1] = (int)(hi - lo);
      }
  std::vector<int4> cellOf(n);
  std::vector<double> loc(4 * n);
  for (size_t s = 0; s < n; s++) {
    const int i = perm[s];
    int gc[3];
    for (int d = 0; d < 3; d++) {
      gc[d] = gcell[3 * i + d];
      // position relative to the low corner of the charge's cluster, as the
      // reference forms it (x2 = x1 - n cellSizeA)
      loc[4 * s + d] = (coord[3 * i + d] + (-mn[d])) - (double)(gc[d] / (int)nP) * csA;
    }
    loc[4 * s + 3] = 0;
    cellOf[s] = make_int4(gc[0], gc[1], gc[2], 0);
  }
  // source segments of every block of 8 consecutive sorted targets: rows of
  // cells within Rdir (slightly inflated, the kernel applies the exact
  // cutoff) of the box spanned by the block's positions, split per cluster
  const long nBlocks = ((long)n + 7) / 8;
  std::vector<int> segStart(nBlocks + 1, 0);
  int maxOff = 1;  // the neighbour list only reaches the adjacent clusters
  std::vector<int4> segs;
  segs.reserve((size_t)nBlocks * 48);
  {
    const double Rm = Rdir * (1.0 + 1.0e-9) + 1.0e-12;
    for (long b = 0; b < nBlocks; b++) {
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      int alo[3] = {1 << 30, 1 << 30, 1 << 30}, ahi[3] = {-1, -1, -1};
      for (size_t s = 8 * b; s < std::min(n, (size_t)(8 * b + 8)); s++) {
        const int i = perm[s];
        for (int d = 0; d < 3; d++) {
          const double x = coord[3 * i + d] + (-mn[d]);
          lo[d] = std::min(lo[d], x);
          hi[d] = std::max(hi[d], x);
          alo[d] = std::min(alo[d], gcell[3 * i + d] / (int)nP);
          ahi[d] = std::max(ahi[d], gcell[3 * i + d] / (int)nP);
        }
      }
      // largest cluster offset between a target of the block and a segment
      auto note_off = [&](int d, int a) {
        maxOff = std::max(maxOff, std::max(std::abs(a - alo[d]), std::abs(a - ahi[d])));
      };
      const int gx0 = std::max(0, (int)std::floor((lo[0] - Rm) / csB));
      const int gx1 = std::min(h.nb[0] - 1, (int)std::floor((hi[0] + Rm) / csB));
      const int gy0 = std::max(0, (int)std::floor((lo[1] - Rm) / csB));
      const int gy1 = std::min(h.nb[1] - 1, (int)std::floor((hi[1] + Rm) / csB));
      for (int gx = gx0; gx <= gx1; gx++) {
        const double gapx = std::max(0.0, std::max(gx * csB - hi[0], lo[0] - (gx + 1) * csB));
        if (gapx >= Rm) continue;
        for (int gy = gy0; gy <= gy1; gy++) {
          const double gapy = std::max(0.0, std::max(gy * csB - hi[1], lo[1] - (gy + 1) * csB));
          const double rem = Rm * Rm - gapx * gapx - gapy * gapy;
          if (rem <= 0) continue;
          const double sr = std::sqrt(rem);
          const int zlo = std::max(0, (int)std::floor((lo[2] - sr) / csB));
          const int zhi = std::min(h.nb[2] - 1, (int)std::floor((hi[2] + sr) / csB));
          const long row = ((long)gx * h.nb[1] + gy) * h.nb[2];
          for (int az = zlo / (int)nP; az <= zhi / (int)nP; az++) {
            const int z0 = std::max(zlo, az * (int)nP), z1 = std::min(zhi, az * (int)nP + (int)nP - 1);
            const int sB = cellSE[2 * (row + z0)];
            const int sE = cellSE[2 * (row + z1)] + cellSE[2 * (row + z1) + 1];
            if (sE > sB) {
              segs.push_back(make_int4(sB, sE, gx / (int)nP, ((gy / (int)nP) << 16) | az));
              note_off(0, gx / (int)nP);
              note_off(1, gy / (int)nP);
              note_off(2, az);
            }
          }
        }
      }
      if (segs.size() > 0x7fffffffUL) throw Error{NPME_ERR_NOMEM, "too many direct-sum segments"};
      segStart[b + 1] = (int)segs.size();
    }
  }

  // ---- spline windows (u = m + N/4) and per-tile lists
  std::vector<int4> ms(n);
  std::vector<double> frac(4 * n);
  for (size_t s = 0; s < n; s++) {
    const int i = perm[s];
    int u0[3];
    for (int d = 0; d < 3; d++) {
      const double w = h.N[d] / h.L[d] * (coord[3 * i + d] - h.R[d]);
      const long m0 = (w < 0) ? (long)(-w) + 1 : (long)(-w);
      if (m0 < -h.N[d] / 4 || m0 + h.Bn - 1 > h.N[d] / 4 - 1)
        throw Error{NPME_ERR_GRID, "spline window outside the compact grid"};
      u0[d] = (int)(m0 + h.N[d] / 4);
      frac[4 * s + d] = w + (double)m0;
    }
    frac[4 * s + 3] = 0;
    ms[s] = make_int4(u0[0], u0[1], u0[2], 0);
  }
  const long ntiles = (long)h.nt[0] * h.nt[1] * h.nt[2];
  std::vector<int> tcount(ntiles + 1, 0);
  auto tile_range = [&](int u0, int d, int &lo, int &hi) {
    lo = u0 / h.T[d];
    hi = (u0 + (int)h.Bn - 1) / h.T[d];
  };
  for (size_t s = 0; s < n; s++) {
    int lo[3], hi[3];
    tile_range(ms[s].x, 0, lo[0], hi[0]);
    tile_range(ms[s].y, 1, lo[1], hi[1]);
    tile_range(ms[s].z, 2, lo[2], hi[2]);
    for (int a = lo[0]; a <= hi[0]; a++)
      for (int b = lo[1]; b <= hi[1]; b++)
        for (int c = lo[2]; c <= hi[2]; c++) tcount[((long)a * h.nt[1] + b) * h.nt[2] + c + 1]++;
  }
  for (long t = 0; t < ntiles; t++) tcount[t + 1] += tcount[t];
  std::vector<int> tlist(tcount[ntiles]);
  {
    std::vector<int> fill(tcount.begin(), tcount.end() - 1);
    for (size_t s = 0; s < n; s++) {
      int lo[3], hi[3];
      tile_range(ms[s].x, 0, lo[0], hi[0]);
      tile_range(ms[s].y, 1, lo[1], hi[1]);
      tile_range(ms[s].z, 2, lo[2], hi[2]);
      for (int a = lo[0]; a <= hi[0]; a++)
        for (int b = lo[1]; b <= hi[1]; b++)
          for (int c = lo[2]; c <= hi[2]; c++) tlist[fill[((long)a * h.nt[1] + b) * h.nt[2] + c]++] = (int)s;
    }
  }
  const long nitiles = (long)h.nti[0] * h.nti[1] * h.nti[2];
  std::vector<int> icount(nitiles + 1, 0), ilist(n);
  auto itile = [&](const int4 &u) {
    return ((long)(u.x / h.TI) * h.nti[1] + u.y / h.TI) * h.nti[2] + u.z / h.TI;
  };
  for (size_t s = 0; s < n; s++) icount[itile(ms[s]) + 1]++;
  for (long t = 0; t < nitiles; t++) icount[t + 1] += icount[t];
  {
    std::vector<int> fill(icount.begin(), icount.end() - 1);
    for (size_t s = 0; s < n; s++) ilist[fill[itile(ms[s])]++] = (int)s;
  }

  // ---- uploads
  const bool f64 = h.fp64;
  const size_t rs = f64 ? 8 : 4;
  auto up = [&](void *dst, const void *src, size_t bytes) {
    NPME_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h.stream));
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));  // the host arrays are temporaries
  };
  h.d_perm = dalloc<int>(h, n);
  up(h.d_perm, perm.data(), n * sizeof(int));
  h.d_frac = dalloc<char>(h, 4 * rs * n);
  if (f64) {
    up(h.d_frac, frac.data(), 4 * n * sizeof(double));
  } else {
    std::vector<float> ff(frac.begin(), frac.end());
    up(h.d_frac, ff.data(), 4 * n * sizeof(float));
  }
  h.d_segStart = dalloc<int>(h, segStart.size());
  up(h.d_segStart, segStart.data(), segStart.size() * sizeof(int));
  h.d_blockBox = dalloc<int4>(h, segs.size());
  up(h.d_blockBox, segs.data(), segs.size() * sizeof(int4));
  h.nSegments = segs.size();
  if (h.complexK) {
    h.d_qi = dalloc<char>(h, rs * (n + 1));  // entry n: the list's zero-charge sentinel
    NPME_HIP_CHECK(hipMemsetAsync((char *)h.d_qi + rs * n, 0, rs, h.stream));
    h.d_qsi = dalloc<char>(h, rs * n);
  }
  h.d_qs = dalloc<char>(h, rs * n);
  h.d_cell = dalloc<int4>(h, n);
  up(h.d_cell, cellOf.data(), n * sizeof(int4));
  h.d_cellStart = dalloc<int>(h, 2 * (size_t)ncell);
  up(h.d_cellStart, cellSE.data(), 2 * (size_t)ncell * sizeof(int));
  h.d_ms = dalloc<int4>(h, n);
  up(h.d_ms, ms.data(), n * sizeof(int4));
  h.d_tileStart = dalloc<int>(h, ntiles + 1);
  up(h.d_tileStart, tcount.data(), (ntiles + 1) * sizeof(int));
  h.d_itileStart = dalloc<int>(h, nitiles + 1);
  up(h.d_itileStart, icount.data(), (nitiles + 1) * sizeof(int));
  h.d_itileList = dalloc<int>(h, n);
  up(h.d_itileList, ilist.data(), n * sizeof(int));
  h.d_tileList = dalloc<int>(h, tlist.size());
  up(h.d_tileList, tlist.data(), tlist.size() * sizeof(int));
  const int vw = h.complexK ? 8 : 4;
  h.d_Vdir = dalloc<double>(h, vw * n);
  h.d_V = dalloc<double>(h, vw * n);
  h.d_qin = dalloc<double>(h, (h.complexK ? 2 : 1) * n);
  h.d_coordS = dalloc<double>(h, 3 * n);
  up(h.d_coordS, coord, 3 * n * sizeof(double));
  h.d_xs = dalloc<double4>(h, n);
  sorted_coords_device(n, h.d_perm, h.d_coordS, h.d_xs, h.stream);

  // ---- direct sum: the neighbour list when it fits, else the cell rows;
  // the fixed-point scale depends on which of the two reads the positions
  if (nlPossible) build_neighbor_list(h, nlmode);
  h.fxShift = h.useNlist ? kListShift : fixed_shift(maxOff);
  if (h.useNlist) {
    // the list walk reads positions in one frame: relative to the low
    // corner of the cells (the cell-row sum keeps them per cluster)
    for (size_t s = 0; s < n; s++)
      for (int d = 0; d < 3; d++) loc[4 * s + d] = coord[3 * (size_t)perm[s] + d] + (-mn[d]);
  }
  // [n + 1]: entry n is the neighbour list's zero-charge sentinel
  h.d_src = dalloc<char>(h, 4 * rs * (n + 1));
  if (f64) {
    up(h.d_src, loc.data(), 4 * n * sizeof(double));
  } else {
    // int32 fixed point (2^S units per cluster edge) in the x, y, z words,
    // relative to the cluster corner for the cell rows and to the cells'
    // low corner, wrapped modulo 2^32, for the list; the w word receives
    // the charge, and the sub-unit residual of each coordinate is kept
    std::vector<int32_t> fx(4 * n, 0);
    std::vector<float> fr(4 * n, 0.0f);
    const double scale = std::ldexp(1.0, h.fxShift) / csA;
    for (size_t s = 0; s < n; s++)
      for (int d = 0; d < 3; d++) {
        const double u = loc[4 * s + d] * scale;
        const long long v = std::llround(u);
        if (h.useNlist ? !(std::fabs(u) < 4.0e18)
                       : (v < -(1L << 20) || v > (1L << h.fxShift) + (1L << 20)))
          throw Error{NPME_ERR_ARG, "position outside its cluster"};
        fx[4 * s + d] = (int32_t)(uint32_t)(unsigned long long)v;
        fr[4 * s + d] = (float)(u - (double)v);
      }
    up(h.d_src, fx.data(), 4 * n * sizeof(int32_t));
    h.d_srcRes = dalloc<float4>(h, n + 1);
    NPME_HIP_CHECK(hipMemsetAsync(static_cast<float4 *>(h.d_srcRes) + n, 0, sizeof(float4), h.stream));
    up(h.d_srcRes, fr.data(), 4 * n * sizeof(float));
  }
  place_sentinel(h);
}

// FFT buffers, twiddles and the kernel multiplier G = F lam^2 / (N1 N2 N3)
// (FP64 build, stored in the working precision)
static void grid_setup(Handle &h)
{
  const bool f64 = h.fp64;
  const size_t rs = f64 ? 8 : 4;
  const long M1 = h.M[0], M2 = h.M[1], M3 = h.M[2];
  const size_t cs = 2 * rs;  // complex element size
  h.d_grid = dalloc<char>(h, (h.complexK ? cs : rs) * (size_t)(M1 * M2 * M3));
  h.d_buf1 = dalloc<char>(h, cs * (size_t)(M1 * M2 * h.Kp));
  h.d_buf2 = dalloc<char>(h, cs * (size_t)(M1 * h.N[1] * h.Kp));
  // host-built tables (twiddles of the CalcV passes in the working
  // precision, FP64 twiddles and the extension table of the kernel grid,
  // lambda^2) go to the device as one block in one copy
  std::vector<char> blob;
  auto put = [&blob](const void *src, size_t bytes) {
    const size_t off = (blob.size() + 255) & ~(size_t)255;
    blob.resize(off + bytes);
    std::memcpy(blob.data() + off, src, bytes);
    return off;
  };
  size_t offTw[3], offTw64[3];
  for (int d = 0; d < 3; d++) {
    h.plan[d] = make_plan((int)h.N[d]);
    const std::vector<double> t = twiddles((int)h.N[d]);
    offTw64[d] = put(t.data(), t.size() * sizeof(double));
    if (f64) {
      offTw[d] = offTw64[d];
    } else {
      const std::vector<float> f(t.begin(), t.end());
      offTw[d] = put(f.data(), f.size() * sizeof(float));
    }
  }

  // G of the previous set-up is reused when everything it depends on is
  // unchanged (a ResetCoord that keeps the box size)
  GridKey key;
  for (int d = 0; d < 3; d++) {
    key.N[d] = h.N[d];
    key.X[d] = h.X[d];
    key.del[d] = h.del[d];
    key.ta[d] = h.ta[d];
  }
  key.Bn = h.Bn;
  key.fp64 = f64;
  key.complexK = h.complexK;
  key.kind = h.ker.kind;
  key.Nder = h.ker.Nder;
  key.alpha = h.ker.alpha;
  key.k0r = h.ker.k0r;
  key.k0i = h.ker.k0i;
  key.Rdir = h.ker.Rdir;
  key.beta = h.ker.beta;
  key.trans = h.trans;
  for (int d = 0; d < 3; d++) key.D[d] = h.trans ? h.transD[d] : 0.0;
  // a user kernel's callbacks may change between set-ups: its G is never reused
  h.gReused = h.gValid && h.d_G && key == h.gKey && !kind_is_user(h.ker.kind);
  if (h.gReused) {
    h.d_tables = dalloc<char>(h, blob.size());
    NPME_HIP_CHECK(hipMemcpyAsync(h.d_tables, blob.data(), blob.size(), hipMemcpyHostToDevice, h.stream));
    for (int d = 0; d < 3; d++) h.d_tw[d] = static_cast<char *>(h.d_tables) + offTw[d];
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));  // before blob is released
    return;
  }
  dev_free(h.d_G);
  h.d_G = nullptr;
  h.deviceBytes -= h.gBytes;
  h.gValid = false;
  h.gBytes = 0;

  if (h.trans) {
    // box pair: the full kernel at r + D on the full grid, its 3-D DFT, and
    // G over the full N1 x N2 x G3 grid (complex)
    if (min_distance_r0(h.transD, h.L) < 1.0e-6) {
      std::ostringstream os;
      os << "box-box separation (" << h.transD[0] << ", " << h.transD[1] << ", " << h.transD[2]
         << ") should be increased: the kernel grid box of edges (" << h.L[0] << ", " << h.L[1] << ", "
         << h.L[2] << ") reaches the singularity";
      throw Error{NPME_ERR_ARG, os.str()};
    }
    const std::vector<double> tab = fulltrans_tables(h.N, h.ta, h.del, h.X, h.transD);
    const size_t offTab = put(tab.data(), tab.size() * sizeof(double));
    std::vector<double> lam;
    for (int d = 0; d < 3; d++) {
      std::vector<double> l(h.N[d]);
      lambda2(h.N[d], h.Bn, l.data());
      lam.insert(lam.end(), l.begin(), l.end());
    }
    const size_t offLam = put(lam.data(), lam.size() * sizeof(double));
    h.d_tables = dalloc<char>(h, blob.size());
    NPME_HIP_CHECK(hipMemcpyAsync(h.d_tables, blob.data(), blob.size(), hipMemcpyHostToDevice, h.stream));
    char *tb = static_cast<char *>(h.d_tables);
    const void *tw64[3];
    for (int d = 0; d < 3; d++) {
      h.d_tw[d] = tb + offTw[d];
      tw64[d] = tb + offTw64[d];
    }
    const size_t tot = (size_t)h.N[0] * h.N[1] * h.N[2];
    DevBuf dA(tot * sizeof(double2)), dB(tot * sizeof(double2));
    std::vector<double> Fhost;
    if (kind_is_user(h.ker.kind)) {
      user_fgrid_host_n(h.user, 0, h.N, tab, true, Fhost);
      NPME_HIP_CHECK(hipMemcpyAsync(dA.p, Fhost.data(), Fhost.size() * sizeof(double),
                                    hipMemcpyHostToDevice, h.stream));
    } else {
      launch_fgrid_trans(h.ker, h.N, (const double *)(tb + offTab), (double2 *)dA.p, h.stream);
    }
    double2 *F = fft3d_forward_fp64(h.N, h.plan, tw64, (double2 *)dA.p, (double2 *)dB.p, h.stream);
    const long G3 = h.complexK ? h.N[2] : h.N[2] / 2 + 1;
    const size_t totG = (size_t)h.N[0] * h.N[1] * ((G3 + 7) / 8) * 8;
    const size_t before = h.deviceBytes;
    h.d_G = dalloc<char>(h, cs * totG);
    h.gBytes = h.deviceBytes - before;
    launch_make_G_full(f64, h.N, G3, F, (const double *)(tb + offLam), h.d_G, h.stream);
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));  // before the temporaries return to the pool
    h.gKey = key;
    h.gValid = true;
    return;
  }

  // kernel grid: real-space compact grid of f_LR with the smooth Fourier
  // extension, then an even-extension DFT along each axis, all FP64
  const long Mf[3] = {h.N[0] / 2 + 1, h.N[1] / 2 + 1, h.N[2] / 2 + 1};
  const long totF = Mf[0] * Mf[1] * Mf[2];
  std::vector<double> tab;
  for (int d = 0; d < 3; d++) {
    const double Xmax = h.X[d] + h.del[d], Xt = 2 * Xmax - h.X[d];
    const double dx = 2 * Xmax / h.N[d];
    std::vector<double> x1(Mf[d]), x2(Mf[d]), T1(Mf[d]), T2(Mf[d]);
    for (long i = 0; i < Mf[d]; i++) {
      x1[i] = i * dx;
      const double eps = 1.0e-12 * Xmax;
      if (x1[i] > Xmax + eps) throw Error{NPME_ERR_GRID, "kernel grid beyond the extension"};
      if (x1[i] > h.X[d]) {
        T1[i] = tsmooth(x1[i], h.X[d], Xt, h.ta[d]);
        x2[i] = x1[i] - 2 * Xmax;
        T2[i] = tsmooth(x2[i], h.X[d], Xt, h.ta[d]);
      } else {
        T1[i] = 1.0;
        x2[i] = 1.0;
        T2[i] = 0.0;
      }
    }
    tab.insert(tab.end(), x1.begin(), x1.end());
    tab.insert(tab.end(), x2.begin(), x2.end());
    tab.insert(tab.end(), T1.begin(), T1.end());
    tab.insert(tab.end(), T2.begin(), T2.end());
  }
  // F: double per point for the real kernels, double2 for Helmholtz
  const size_t offTab = put(tab.data(), tab.size() * sizeof(double));
  std::vector<double> lam;
  for (int d = 0; d < 3; d++) {
    std::vector<double> l(h.N[d]);
    lambda2(h.N[d], h.Bn, l.data());
    lam.insert(lam.end(), l.begin(), l.begin() + Mf[d]);
  }
  const size_t offLam = put(lam.data(), lam.size() * sizeof(double));
  h.d_tables = dalloc<char>(h, blob.size());
  NPME_HIP_CHECK(hipMemcpyAsync(h.d_tables, blob.data(), blob.size(), hipMemcpyHostToDevice, h.stream));
  char *tb = static_cast<char *>(h.d_tables);
  const void *tw64[3];
  for (int d = 0; d < 3; d++) {
    h.d_tw[d] = tb + offTw[d];
    tw64[d] = tb + offTw64[d];
  }
  DevBuf d_F((h.complexK ? 2 : 1) * totF * sizeof(double));
  std::vector<double> Fhost;
  if (kind_is_user(h.ker.kind)) {
    // user kernels: f_LR sampled through the host callback at the same
    // points and weights (src/RecSumGrid.cpp:45-162)
    user_fgrid_host(h.user, Mf, tab, Fhost);
    NPME_HIP_CHECK(hipMemcpyAsync(d_F.p, Fhost.data(), Fhost.size() * sizeof(double),
                                  hipMemcpyHostToDevice, h.stream));
  } else {
    launch_fgrid_real_space(h.ker, Mf, (const double *)(tb + offTab), (double *)d_F.p, h.stream);
  }
  // even-extension DFT along axis 3, 2, 1 of F[Mf0][Mf1][Mf2]
  double *F = (double *)d_F.p;
  auto pass = h.complexK ? fft_pass_even
                         : [](int N, long nCols, long nInner, long outer, long inner, long pos,
                              const FFTPlan1D &plan, const void *tw, void *data, hipStream_t st) {
                             fft_pass_even_real(N, nCols, nInner, outer, inner, pos, plan, tw,
                                                (double *)data, st);
                           };
  pass((int)h.N[2], Mf[0] * Mf[1], Mf[0] * Mf[1], 0, Mf[2], 1, h.plan[2], tw64[2], F, h.stream);
  pass((int)h.N[1], Mf[0] * Mf[2], Mf[2], Mf[1] * Mf[2], 1, Mf[2], h.plan[1], tw64[1], F, h.stream);
  pass((int)h.N[0], Mf[1] * Mf[2], Mf[1] * Mf[2], 0, 1, Mf[1] * Mf[2], h.plan[0], tw64[0], F, h.stream);
  // G in blocks of 8 along axis 3 (the last block padded)
  const size_t totG = (size_t)Mf[0] * Mf[1] * ((Mf[2] + 7) / 8) * 8;
  const size_t before = h.deviceBytes;
  h.d_G = dalloc<char>(h, (h.complexK ? cs : rs) * totG);
  h.gBytes = h.deviceBytes - before;
  launch_make_G(f64, h.complexK, Mf, h.K3, F, (const double *)(tb + offLam), h.d_G, h.stream);
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));  // before the temporaries return to the pool
  h.gKey = key;
  h.gValid = true;
}

// ---------------------------------------------------------------- set-up
void handle_setup(Handle &h, const Keywords &kw, size_t n, const double *coord)
{
  auto t0 = std::chrono::steady_clock::now();
  for (double &v : h.setupMs) v = 0;
  h.ready = false;
  h.kw = kw;
  h.n = n;
  if (n < 2) throw Error{NPME_ERR_ARG, "nCharge must be at least 2"};
  std::string funcType, split, prec;
  double Rdir;
  int Nder = 0;
  // user-defined kernels (GenFunc with plugins) carry their own split: the
  // keywords give only the grid and the cutoff
  if (!h.userK) {
    if (!kw.get("funcType", funcType)) kw_error("missing keyword 'funcType'");
    if (!kw.get("EwaldSplit", split)) kw_error("missing keyword 'EwaldSplit'");
    if (split != "DerivMatch" && split != "LaplaceOrig")
      kw_error("EwaldSplit = " + split + " is undefined ('DerivMatch' or 'LaplaceOrig')");
    if (split == "LaplaceOrig" && funcType != "Laplace")
      kw_error("keywords for EwaldSplit and funcType are incompatible");
    if (!kw.get("nDeriv", Nder)) kw_error("missing keyword 'nDeriv'");
  }
  if (!kw.get("Rdir", Rdir)) kw_error("missing keyword 'Rdir'");
  if (!(Rdir > 0)) kw_error("Rdir must be positive");
  if (!kw.get("BnOrder", h.Bn)) kw_error("missing keyword 'BnOrder'");
  const char *nk[3] = {"N1", "N2", "N3"}, *bk[3] = {"n1", "n2", "n3"};
  const char *dk[3] = {"del1", "del2", "del3"}, *ak[3] = {"a1", "a2", "a3"};
  for (int d = 0; d < 3; d++) {
    if (!kw.get(nk[d], h.N[d])) kw_error(std::string("missing keyword '") + nk[d] + "'");
    if (!kw.get(bk[d], h.nblk[d])) kw_error(std::string("missing keyword '") + bk[d] + "'");
    if (!kw.get(dk[d], h.del[d])) kw_error(std::string("missing keyword '") + dk[d] + "'");
    if (!kw.get(ak[d], h.ta[d])) kw_error(std::string("missing keyword '") + ak[d] + "'");
  }
  if (!kw.get("nNeigh", h.nNeigh)) kw_error("missing keyword 'nNeigh'");
  if (!kw.get("nCellClust1D", h.nCellClust1D)) kw_error("missing keyword 'nCellClust1D'");
  // useCompactF 0 asks for the kernel grid F on the full N1 x N2 x N3 grid
  // (src/RecSumGrid.cpp:293-379, src/RecSumV.cpp:583-637).  The built-in
  // kernels are radial and the grid samples are symmetric, so the full F is
  // the even extension of the compact one and the convolution is the same:
  // both values run the compact F.
  long useCompactF = 1;
  if (kw.get("useCompactF", useCompactF) && useCompactF != 0 && useCompactF != 1)
    kw_error("useCompactF must be 0 or 1");
  h.fp64 = kw.get("precision", prec) && prec == "fp64";
  {
    std::string fk = "auto";
    kw.get("fftKernels", fk);
    if (fk != "auto" && fk != "generic") kw_error("fftKernels must be 'auto' or 'generic'");
    h.fftPow8 = fk == "auto";
  }
  if (kw.get("precision", prec) && prec != "fp64" && prec != "mixed")
    kw_error("precision must be 'mixed' or 'fp64'");

  int kind;
  double alpha = -1.0, k0r = 0, k0i = 0;
  if (h.userK) kind = h.user.complex ? KIND_USER_CX : KIND_USER;
  else if (funcType == "Laplace") kind = split == "LaplaceOrig" ? KIND_LAPLACE_ERF : KIND_LAPLACE;
  else if (funcType == "Ralpha") {
    kind = KIND_RALPHA;
    if (!kw.get("alpha", alpha)) kw_error("funcType = Ralpha needs 'alpha'");
  } else if (funcType == "Helmholtz") {
    kind = KIND_HELMHOLTZ;
    if (!kw.get("k0_r", k0r) || !kw.get("k0_i", k0i)) kw_error("funcType = Helmholtz needs k0_r, k0_i");
    if (k0i < -1.0e-12) kw_error("k0_i is negative");
  } else {
    kw_error("funcType = " + funcType + " is undefined");
    return;
  }
  h.complexK = kind_is_complex(kind);
  std::string err;
  h.utS = 0;
  h.utErr[0] = h.utErr[1] = 0;
  std::vector<double> utab;
  if (kind_is_user(kind)) {
    // NPME_InterfaceReal_GenFunc::SetUp (src/Interface.cpp:1606-1671):
    // fself = funcLR(0, 0, 0); the short-range part is tabulated here and
    // the long-range part sampled on the kernel grid in grid_setup
    h.ker = DMKernel{};
    h.ker.kind = kind;
    h.ker.Rdir = Rdir;
    if (!h.trans) {  // a box pair uses only func, on the kernel grid
      user_check_radial(h.user, Rdir);
      // the reference calls funcSR on every point pair of the interacting
      // cell pairs (src/PotentialGenFunc.cpp:523-645), up to the farthest
      // points of cells nNeigh apart on each axis: sqrt(3) (nNeigh + 1)
      // cell edges
      h.utR = std::sqrt(3.0) * (double)(h.nNeigh + 1) * (Rdir / (double)h.nNeigh) * (1.0 + 1.0e-6);
      user_build_table(h.user, h.utR, utab, h.utS, h.utErr);
      user_fself(h.user, h.fself);
    }
    h.coordHost.assign(coord, coord + 3 * n);
  } else if (kind == KIND_LAPLACE_ERF) {
    // NPME_InterfaceReal_Laplace_Original::SetUp (src/Interface.cpp:2152-2230)
    double tol;
    if (!kw.get("tol", tol)) kw_error("missing keyword 'tol'");
    if (!(Rdir > 0) || !(tol > 0) || !(tol < 1)) kw_error("EwaldSplit LaplaceOrig needs Rdir > 0, 0 < tol < 1");
    erf_kernel_init(h.ker, Nder, Rdir, rdir2beta(Rdir, tol));
  } else if (!dm_kernel_init(h.ker, kind, alpha, k0r, k0i, Nder, Rdir, err)) {
    kw_error(err);
  }
  if (!kind_is_user(kind)) {
    double f1[2];
    dm_eval_lr(h.ker, 0.0, h.fself, f1);
    h.coordHost.clear();
  }
  if (h.trans) h.fself[0] = h.fself[1] = 0.0;  // a box pair has no self term

  // NPME_CheckFFTParm (src/RecSumSupportFunctions.cpp:37-135)
  for (int d = 0; d < 3; d++) {
    const long N = h.N[d], b = h.nblk[d];
    std::ostringstream os;
    if (N % 4 != 0) os << "N" << d + 1 << " = " << N << " is not a multiple of 4";
    else if (h.Bn >= N / 4) os << "BnOrder = " << h.Bn << " >= N" << d + 1 << "/4";
    else if (b > N / 4) os << "block size n" << d + 1 << " = " << b << " > N/4";
    else if (b < 1 || N % (2 * b) != 0) os << "N" << d + 1 << " must be a multiple of 2 n" << d + 1;
    if (!os.str().empty()) throw Error{NPME_ERR_GRID, os.str()};
    if (make_plan((int)N).nst < 0) throw Error{NPME_ERR_GRID, "FFT size has a prime factor > 7"};
    h.M[d] = N / 2;
  }
  if (h.Bn < 2 || h.Bn > 16) throw Error{NPME_ERR_UNSUPPORTED, "BnOrder must be in [2, 16]"};
  // an odd order's interpolation factor |b(m)|^2 vanishes at m = N/2 (the
  // reference documents BnOrder as an even integer)
  if (h.Bn % 2 != 0) throw Error{NPME_ERR_UNSUPPORTED, "BnOrder must be even"};
  // the mixed form holds a 1e-6 relative error against the FP64 path up to
  // order 12 (measured 2.5e-7 for orders 4-12 on a 64^3 grid, 1.06e-6 at
  // order 14)
  if (!h.fp64 && h.Bn > kMixedMaxBn)
    throw Error{NPME_ERR_UNSUPPORTED, "precision mixed is provided up to BnOrder " +
                                          std::to_string(kMixedMaxBn) +
                                          "; use precision fp64 for higher orders"};
  // one FFT column must fit a workgroup's LDS in FP64 (kernel-grid set-up)
  for (int d = 0; d < 3; d++)
    if (h.N[d] + 1 > 16 * 256)
      throw Error{NPME_ERR_UNSUPPORTED, "FFT length above 4095 is not provided"};
  h.K3 = h.complexK ? h.N[2] : h.N[2] / 2 + 1;
  // rows of the FFT buffers padded to a multiple of 8 complex values, so a
  // workgroup's 8 adjacent columns are one aligned 64-byte segment per row
  h.Kp = (h.K3 + 7) / 8 * 8;

  // ---- streams; the device memory of an earlier set-up is released
  // (positions staged for this set-up are kept)
  ensure_streams(h);
  std::string smode = "device", nlmode = "auto";
  kw.get("setup", smode);
  kw.get("neighborList", nlmode);
  if (smode != "device" && smode != "host") kw_error("setup must be 'device' or 'host'");
  if (nlmode != "auto" && nlmode != "on" && nlmode != "off")
    kw_error("neighborList must be 'auto', 'on' or 'off'");
  if (h.trans) nlmode = "off";  // no direct sum between the boxes
  // LaplaceOrig and user kernels sum the reference's cell pairs, not a
  // cutoff list
  if (kind_ref_pairs(kind)) nlmode = "off";
  const bool nlPossible = nlmode != "off" && h.nNeigh <= h.nCellClust1D && n < (1u << 26);
  const bool devSetup = smode == "device";
  const bool staged = devSetup && h.coordStaged && h.stagedN == n;
  if (!staged) {
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
    handle_free_device(h, true);
  }
  h.setupMs[1] = ms_since(t0);

  // ---- bounds of the charges: on the device for the device set-up (the
  // positions go there first), else on the host; both exact
  double mn[3], mx[3];
  if (staged) {
    handle_staged_bounds(h, mn, mx);
    h.coordStaged = h.stagedBoundsReady = false;
  } else if (devSetup) {
    h.d_coordS = dalloc<double>(h, 3 * n);
    NPME_HIP_CHECK(hipMemcpyAsync(h.d_coordS, coord, 3 * n * sizeof(double), hipMemcpyHostToDevice,
                                  h.stream));
    coord_bounds_device(n, h.d_coordS, mn, mx, h.stream);
  } else {
    bool finite = true;
    for (int d = 0; d < 3; d++) mn[d] = mx[d] = coord[d];
    for (size_t i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) {
        const double x = coord[3 * i + d];
        finite = finite && std::isfinite(x);
        mn[d] = std::min(mn[d], x);
        mx[d] = std::max(mx[d], x);
      }
    if (!finite) mn[0] = NAN;
  }
  for (int d = 0; d < 3; d++)
    if (!std::isfinite(mn[d]) || !std::isfinite(mx[d]))
      throw Error{NPME_ERR_ARG, "coordinates must be finite"};

  // ---- geometry of the reciprocal sum: box size and centre as
  // src/SupportFunctions.cpp NPME_FindBoxDimCenter (extremes seeded with
  // -/+1e12)
  box_from_bounds(mn, mx, h.X0, h.R0);
  pme_corr_box(h.X0, h.R0, h.N, h.del, h.Bn, h.X, h.R);
  for (int d = 0; d < 3; d++) h.L[d] = 2 * (h.X[d] + h.del[d]);

  // ---- embedded cell partition: cells of Rdir/nNeigh inside clusters of
  // nCellClust1D^3 cells, over the bounding box of the charges
  GeoParams g{};
  const double csB = Rdir / h.nNeigh;
  const unsigned long nP = (unsigned long)h.nCellClust1D;
  const double csA = nP * csB;
  for (int d = 0; d < 3; d++) {
    g.nA[d] = (unsigned long)((mx[d] - mn[d]) / csA) + 1;
    h.nb[d] = (int)(g.nA[d] * nP);
    h.boxMin[d] = g.mn[d] = mn[d];
    g.nb[d] = h.nb[d];
    g.NoverL[d] = h.N[d] / h.L[d];
    g.R[d] = h.R[d];
    g.N[d] = h.N[d];
  }
  h.csB = g.csB = csB;
  h.csA = g.csA = csA;
  g.nP = nP;
  g.Bn = h.Bn;
  // spreading tiles: 16 x 8 x 8 points for the real kernels (matrix-core
  // spreading, FP32 and FP64), else 8^3; interpolation tiles of TI^3 points
  // whose block plus halo fits 64 KB of LDS
  h.T[0] = !h.complexK ? 16 : 8;
  h.T[1] = h.T[2] = 8;
  for (int d = 0; d < 3; d++) {
    g.ts[d] = h.T[d];
    h.nt[d] = g.nt[d] = (int)((h.M[d] + h.T[d] - 1) / h.T[d]);
  }
  {
    // largest tile edge of 16, 12, 8, 6, 4, 2, 1 whose theta block plus halo
    // fits 64 KB (FP64 real at order 6: 12, a 17^3 block of 39 KB, which
    // reads the grid 2.8 times over against 4.3 times for edge 8); every
    // accepted order fits at some edge >= 1 (complex FP64 at order 16:
    // edge 1, 16^3 points of 16 bytes)
    const long es = (h.fp64 ? 8 : 4) * (h.complexK ? 2 : 1);
    auto bytes = [&](int TI) {
      const long E = TI + h.Bn - 1;
      return E * E * E * es;
    };
    const int edges[] = {16, 12, 8, 6, 4, 2, 1};
    for (int e : edges) {
      h.TI = e;
      if (bytes(e) <= 65536) break;
    }
    int maxLds = 0;
    NPME_HIP_CHECK(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, h.device));
    if (bytes(h.TI) > std::min<long>(65536, maxLds)) {
      std::ostringstream os;
      os << "interpolation tile of " << bytes(h.TI) << " bytes exceeds the LDS of a workgroup ("
         << std::min<long>(65536, maxLds) << " bytes)";
      throw Error{NPME_ERR_UNSUPPORTED, os.str()};
    }
  }
  g.TI = h.TI;
  for (int d = 0; d < 3; d++) h.nti[d] = g.nti[d] = (int)((h.M[d] + h.TI - 1) / h.TI);

  // ---- per-charge state: on the device when the neighbour list is used
  h.setupMs[0] = ms_since(t0);
  bool done = false;
  if (devSetup) done = particles_device(h, g, nlmode, nlPossible);
  if (!done) {
    handle_free_device(h);
    particles_host(h, g, nlmode, nlPossible, coord);
  }
  h.deviceSetup = done;
  ms_since(t0);
  if (!utab.empty()) {
    // the short-range table in FP64: the reference-pair direct sum
    // evaluates it in FP64 in both forms
    h.d_utab = dalloc<double>(h, utab.size());
    NPME_HIP_CHECK(hipMemcpyAsync(h.d_utab, utab.data(), utab.size() * sizeof(double),
                                  hipMemcpyHostToDevice, h.stream));
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));  // before utab is released
  }
  if (kind_ref_pairs(h.ker.kind)) {
    const size_t ncell = (size_t)h.nb[0] * h.nb[1] * h.nb[2];
    h.d_cellFlags = dalloc<unsigned>(h, ncell * refpair_words((int)h.nNeigh));
    launch_cellpair_contact(direct_params(h), h.d_cellStart, h.d_xs, h.d_cellFlags, h.stream);
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  }
  grid_setup(h);
  if (h.trans) {
    // the interpolation adds the (absent) direct sum: zero, once
    const size_t vb = h.n * 4 * (h.complexK ? 2 : 1) * sizeof(double);
    NPME_HIP_CHECK(hipMemsetAsync(h.d_Vdir, 0, vb, h.stream));
    NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  }
  h.setupMs[5] = ms_since(t0);
  h.ready = true;
}

// a failed SetUp / ResetCoord leaves no usable device state
static void require_ready(const Handle &h)
{
  if (!h.ready) throw Error{NPME_ERR_ARG, "no valid set-up: the last SetUp or ResetCoord failed"};
}

void handle_destroy_streams(Handle &h)
{
  drop_graph(h);
  for (hipEvent_t e : h.ev) (void)hipEventDestroy(e);
  h.ev.clear();
  if (h.evFork) (void)hipEventDestroy(h.evFork);
  if (h.evJoin) (void)hipEventDestroy(h.evJoin);
  h.evFork = h.evJoin = nullptr;
  if (h.stream2) (void)hipStreamDestroy(h.stream2);
  if (h.stream) (void)hipStreamDestroy(h.stream);
  h.stream = h.stream2 = nullptr;
}

// ---------------------------------------------------------------- CalcV
static void mark(Handle &h, int i)
{
  if (!h.profiling) return;
  while ((int)h.ev.size() <= i) {
    hipEvent_t e;
    NPME_HIP_CHECK(hipEventCreate(&e));
    h.ev.push_back(e);
  }
  NPME_HIP_CHECK(hipEventRecord(h.ev[i], h.stream));
}

DirectParams direct_params(const Handle &h)
{
  DirectParams dp;
  dp.kind = h.ker.kind;
  dp.Nder = h.ker.Nder;
  dp.Rdir = h.ker.Rdir;
  dp.alpha = h.ker.alpha;
  dp.beta = h.ker.beta;
  dp.k0r = h.ker.k0r;
  dp.k0i = h.ker.k0i;
  std::memcpy(dp.a, h.ker.a, sizeof(dp.a));
  std::memcpy(dp.b, h.ker.b, sizeof(dp.b));
  dp.nNeigh = (int)h.nNeigh;
  dp.cs = h.csB;
  dp.csA = h.csA;
  // the cell of a charge is found from differences of coordinates of this
  // magnitude, rounded, with an edge tolerance of 1e-14
  double mag = 0;
  for (int d = 0; d < 3; d++) {
    dp.nb[d] = h.nb[d];
    dp.mn[d] = h.boxMin[d];
    mag = std::max(mag, std::fabs(h.boxMin[d]) + h.nb[d] * h.csB);
  }
  dp.cellTol = 1.0e-9 * h.csB + 1.0e-12 * mag + 1.0e-13;
  dp.fxShift = h.fxShift;
  dp.fxInv = std::ldexp(h.csA, -h.fxShift);
  dp.utab = h.d_utab;
  dp.utS = h.utS;
  dp.utR = h.utR;
  dp.nlEshift = h.fp64 ? 5 : 4;
  return dp;
}

double handle_count_pairs(Handle &h)
{
  require_ready(h);
  unsigned long long tot = 0;
  DevBuf d_tot(sizeof(unsigned long long));
  NPME_HIP_CHECK(hipMemsetAsync(d_tot.p, 0, sizeof(unsigned long long), h.stream));
  launch_count_pairs(direct_params(h), h.n, h.d_cell, h.d_cellStart, h.d_xs,
                     (unsigned long long *)d_tot.p, h.stream);
  NPME_HIP_CHECK(hipMemcpyAsync(&tot, d_tot.p, sizeof(tot), hipMemcpyDeviceToHost, h.stream));
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
  return (double)tot;
}

// The launches of one CalcV.  The direct sum (gather of the packed sources,
// pair sum into Vdir) runs on stream2 while the reciprocal sum runs on
// stream: stream2 waits for a fork event recorded on stream, and stream
// waits for a join event recorded on stream2 before the interpolation,
// which reads Vdir.  Both events are recorded before their waits are
// queued, so a wait never depends on work queued after it.  Without split
// (per-stage timing) everything runs on stream in order, with the stage
// events.
static void calcv_launches(Handle &h, const double *d_q, double *d_Vout, bool split)
{
  const bool f64 = h.fp64, cx = h.complexK;
  hipStream_t s = h.stream;
  const long M1 = h.M[0], M2 = h.M[1], M3 = h.M[2], N1 = h.N[0], N2 = h.N[1], N3 = h.N[2];
  const long K3 = h.K3, Kp = h.Kp;
  hipStream_t s2 = split ? h.stream2 : s;
  const DirectParams dp = direct_params(h);
  const bool doDirect = !h.trans;  // a box pair has no direct sum
  const bool refPairs = kind_ref_pairs(h.ker.kind);
  // the charges in sorted order: one gather on stream writes the direct
  // sum's packed sources and the spreading copy (a box pair has no direct
  // sum, so only the spreading copy)
  auto gather = [&] {
    if (doDirect)
      launch_gather_both(f64, cx, h.n, h.d_perm, d_q, h.d_src, h.d_qi, h.d_qs, h.d_qsi, s);
    else
      launch_gather_q(f64, cx, h.n, h.d_perm, d_q, h.d_qs, h.d_qsi, s);
  };
  auto direct = [&](hipStream_t ds) {
    if (refPairs)
      launch_direct_refpairs(dp, h.n, h.d_cell, h.d_cellStart, h.d_xs, h.d_perm, d_q, h.d_cellFlags,
                             h.d_Vdir, ds);
    else if (h.useNlist)
      launch_direct_nl(f64, dp, h.n, h.d_src, h.d_srcRes, h.d_qi, h.d_sliceStart, h.d_sliceLen,
                       h.d_nlist, h.csA, h.d_Vdir, ds);
    else
      launch_direct(f64, dp, h.n, h.d_src, h.d_srcRes, h.d_qi, h.d_cell, h.d_segStart, h.d_blockBox,
                    (int)h.nCellClust1D, h.csA, h.d_Vdir, ds);
  };
  int st = 0;
  mark(h, st++);
  gather();
  if (split && doDirect) {
    NPME_HIP_CHECK(hipEventRecord(h.evFork, s));  // after the gather
    NPME_HIP_CHECK(hipStreamWaitEvent(s2, h.evFork, 0));
    direct(s2);
    NPME_HIP_CHECK(hipEventRecord(h.evJoin, s2));
  }
  mark(h, st++);  // stage "gather"
  if (!split && doDirect) {
    // the direct sum in order on stream, timed as its own stage
    direct(s);
    mark(h, st++);
  }
  launch_spread(f64, cx, h.Bn, h.M, h.T, h.nt, h.d_tileStart, h.d_tileList, h.d_qs, h.d_qsi,
                h.d_ms, h.d_frac, h.d_grid, s);
  mark(h, st++);  // stage "spread"
  if (!cx) {
    fft_pass_r2c_rows(f64, (int)N3, (int)M3, M1 * M2, M3, Kp, h.plan[2], h.d_tw[2], h.d_grid,
                      h.d_buf1, s);
    mark(h, st++);
    PassGeom b{(int)N2, (int)M2, (int)N2, M1 * Kp, Kp, M2 * Kp, 1, Kp, N2 * Kp, 1, Kp};
    b.validInner = K3;
    fft_pass_complex(f64, b, h.plan[1], h.d_tw[1], h.d_buf1, h.d_buf2, false, s, h.fftPow8);
    mark(h, st++);
    fft_pass_convolve(f64, false, (int)N1, (int)M1, N2 * Kp, N2 * Kp, (int)N2, (int)(N2 / 2 + 1),
                      Kp, K3, h.plan[0], h.d_tw[0], h.d_G, h.d_buf2, s, h.fftPow8, h.trans,
                      h.transFlags);
    mark(h, st++);
    PassGeom d{(int)N2, (int)N2, (int)M2, M1 * Kp, Kp, N2 * Kp, 1, Kp, M2 * Kp, 1, Kp};
    d.validInner = K3;
    fft_pass_complex(f64, d, h.plan[1], h.d_tw[1], h.d_buf2, h.d_buf1, true, s, h.fftPow8);
    mark(h, st++);
    fft_pass_c2r_rows(f64, (int)N3, (int)M3, M1 * M2, Kp, M3, h.plan[2], h.d_tw[2], h.d_buf1,
                      h.d_grid, s);
    mark(h, st++);
  } else {
    PassGeom a{(int)N3, (int)M3, (int)N3, M1 * M2, M1 * M2, 0, M3, 1, 0, Kp, 1};
    fft_pass_complex(f64, a, h.plan[2], h.d_tw[2], h.d_grid, h.d_buf1, false, s, h.fftPow8);
    mark(h, st++);
    PassGeom b{(int)N2, (int)M2, (int)N2, M1 * Kp, Kp, M2 * Kp, 1, Kp, N2 * Kp, 1, Kp};
    b.validInner = N3;
    fft_pass_complex(f64, b, h.plan[1], h.d_tw[1], h.d_buf1, h.d_buf2, false, s, h.fftPow8);
    mark(h, st++);
    fft_pass_convolve(f64, true, (int)N1, (int)M1, N2 * Kp, N2 * Kp, (int)N2, (int)(N2 / 2 + 1),
                      Kp, N3, h.plan[0], h.d_tw[0], h.d_G, h.d_buf2, s, h.fftPow8, h.trans,
                      h.transFlags);
    mark(h, st++);
    PassGeom d{(int)N2, (int)N2, (int)M2, M1 * Kp, Kp, N2 * Kp, 1, Kp, M2 * Kp, 1, Kp};
    d.validInner = N3;
    fft_pass_complex(f64, d, h.plan[1], h.d_tw[1], h.d_buf2, h.d_buf1, true, s, h.fftPow8);
    mark(h, st++);
    PassGeom e{(int)N3, (int)N3, (int)M3, M1 * M2, M1 * M2, 0, Kp, 1, 0, M3, 1};
    fft_pass_complex(f64, e, h.plan[2], h.d_tw[2], h.d_buf1, h.d_grid, true, s, h.fftPow8);
    mark(h, st++);
  }
  const double C[3] = {N1 / h.L[0], N2 / h.L[1], N3 / h.L[2]};
  if (split && doDirect) NPME_HIP_CHECK(hipStreamWaitEvent(s, h.evJoin, 0));  // Vdir is complete
  launch_interp(f64, cx, h.Bn, h.M, C, h.TI, h.nti, h.d_itileStart, h.d_itileList, h.d_ms,
                h.d_frac, h.d_grid, h.d_Vdir, h.d_perm, d_q, h.fself, d_Vout, s);
  mark(h, st++);
  if (!split) {
    NPME_HIP_CHECK(hipStreamSynchronize(s));
    h.nstage = st - 1;
    for (int i = 0; i + 1 < st; i++) {
      float ms = 0;
      NPME_HIP_CHECK(hipEventElapsedTime(&ms, h.ev[i], h.ev[i + 1]));
      h.stage_ms[i] = ms;
    }
  }
}

// the CalcV launches captured from stream as one graph
static hipGraphExec_t capture_calcv(Handle &h, const double *d_q, double *d_Vout)
{
  NPME_HIP_CHECK(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
  hipGraph_t g = nullptr;
  try {
    calcv_launches(h, d_q, d_Vout, true);
  } catch (...) {
    if (hipStreamEndCapture(h.stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    throw;
  }
  NPME_HIP_CHECK(hipStreamEndCapture(h.stream, &g));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  NPME_HIP_CHECK(e);
  return ex;
}

// the CalcV of m set-up handles (charges in each handle's d_qin, V to its
// d_V) captured as one graph and launched once on stream sb: each handle's
// stream waits for an event recorded on sb, the handles' launches are
// recorded, and sb waits for each handle's last launch; the handles' CalcVs
// then run concurrently on the device
void handle_calcv_batch(Handle *const *hs, int m, hipStream_t sb, hipEvent_t evStart,
                        const hipEvent_t *evDone)
{
  for (int k = 0; k < m; k++) require_ready(*hs[k]);
  NPME_HIP_CHECK(hipStreamBeginCapture(sb, hipStreamCaptureModeThreadLocal));
  hipGraph_t g = nullptr;
  try {
    NPME_HIP_CHECK(hipEventRecord(evStart, sb));
    for (int k = 0; k < m; k++) {
      Handle &h = *hs[k];
      NPME_HIP_CHECK(hipStreamWaitEvent(h.stream, evStart, 0));
      calcv_launches(h, h.d_qin, h.d_V, true);
      NPME_HIP_CHECK(hipEventRecord(evDone[k], h.stream));
      NPME_HIP_CHECK(hipStreamWaitEvent(sb, evDone[k], 0));
    }
  } catch (...) {
    if (hipStreamEndCapture(sb, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
    throw;
  }
  const hipError_t ec = hipStreamEndCapture(sb, &g);
  if (ec != hipSuccess || !g) {
    // an invalidated capture (HIP invalidates a capture when a
    // synchronising operation touches its streams, possibly from another
    // thread): the same launches, issued directly
    (void)hipGetLastError();
    for (int k = 0; k < m; k++) {
      Handle &h = *hs[k];
      calcv_launches(h, h.d_qin, h.d_V, true);
    }
    for (int k = 0; k < m; k++) NPME_HIP_CHECK(hipStreamSynchronize(hs[k]->stream));
    return;
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  NPME_HIP_CHECK(e);
  const hipError_t el = hipGraphLaunch(ex, sb);
  const hipError_t es = el == hipSuccess ? hipStreamSynchronize(sb) : el;
  (void)hipGraphExecDestroy(ex);
  NPME_HIP_CHECK(es);
}

void handle_set_trans_rev(Handle &h, int rev)
{
  // G holds DFT(g) of the translated kernel; reversing axis d uses G at
  // -k_d.  The R2C data of real charges stores k3 >= 0 only: G(k1, k2, -k3)
  // = conj(G(-k1, -k2, k3)) for the real g of a real kernel.
  int f = rev & 7;
  if (!h.complexK && (rev & 4)) f = ((~rev) & 3) | 8;
  if (f != h.transFlags) drop_graph(h);  // the flags are arguments of a captured launch
  h.transRev = rev & 7;
  h.transFlags = f;
}

void handle_calcv(Handle &h, const double *d_q, double *d_Vout)
{
  require_ready(h);
  if (h.profiling) {
    calcv_launches(h, d_q, d_Vout, false);
    return;
  }
  // repeated calls with the same pointers on one set-up replay a graph
  if (d_q == h.graphQ && d_Vout == h.graphV) {
    h.graphCalls++;
  } else {
    drop_graph(h);
    h.graphQ = d_q;
    h.graphV = d_Vout;
    h.graphCalls = 1;
  }
  if (h.useGraph && h.graphCalls >= 2) {
    if (!h.calcvExec) h.calcvExec = capture_calcv(h, d_q, d_Vout);
    NPME_HIP_CHECK(hipGraphLaunch(h.calcvExec, h.stream));
  } else {
    calcv_launches(h, d_q, d_Vout, true);
  }
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
}


void handle_exact(Handle &h, const double *d_q, size_t nt, const long *d_idx, double *d_Vout)
{
  require_ready(h);
  launch_exact(h.ker.kind, h.ker.alpha, h.ker.k0r, h.ker.k0i, h.n, h.d_coordS, h.d_perm, d_q, nt,
               d_idx, d_Vout, h.stream);
  NPME_HIP_CHECK(hipStreamSynchronize(h.stream));
}

}  // namespace npme
