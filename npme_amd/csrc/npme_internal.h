// Internal declarations shared by the host set-up code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <exception>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/npme_amd.h"

namespace npme {

// ------------------------------------------------------------- errors
void set_error(const std::string &msg);
std::string last_error();  // the calling thread's last error message
struct Status {
  int code;
  std::string msg;
};

#define NPME_HIP_CHECK(expr)                                                  \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw npme::DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Error {
  int code;
  std::string msg;
};
struct DeviceError {
  std::string msg;
  explicit DeviceError(std::string m) : msg(std::move(m)) {}
};

// runs f, mapping exceptions to a status code and the thread's last error
template <typename F> inline int guarded(F &&f)
{
  try {
    f();
    return NPME_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const DeviceError &e) {
    set_error(e.msg);
    return NPME_ERR_DEVICE;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return NPME_ERR_NOMEM;
  } catch (const std::exception &e) {
    set_error(e.what());
    return NPME_ERR_ARG;
  }
}

// ------------------------------------------------------------- kernels
// KIND_LAPLACE_ERF: 1/r with the Ewald split erf(beta r)/r + erfc(beta r)/r
// (EwaldSplit = LaplaceOrig) instead of the derivative match
// KIND_USER / KIND_USER_CX: user-defined real / complex kernels
// (NPME_InterfaceReal_GenFunc / NPME_InterfaceComplex_GenFunc with
// NPME_KfuncReal / NPME_KfuncComplex plugins): the short-range part is
// evaluated on the device from a host-built table
enum KernelKind {
  KIND_LAPLACE = 0,
  KIND_RALPHA = 1,
  KIND_HELMHOLTZ = 2,
  KIND_LAPLACE_ERF = 3,
  KIND_USER = 4,
  KIND_USER_CX = 5
};
__host__ __device__ constexpr bool kind_is_complex(int kind)
{
  return kind == KIND_HELMHOLTZ || kind == KIND_USER_CX;
}
__host__ __device__ constexpr bool kind_is_user(int kind)
{
  return kind == KIND_USER || kind == KIND_USER_CX;
}

// short-range table of a user kernel: f(r) and g(r) = f'(r)/r (complex:
// real and imaginary parts) on r = Rdir 2^-(k+1) (1 + x), x in [0, 1), per
// octave k < kUserOct in S equal pieces, each a polynomial of degree
// kUserDeg in t in [-1, 1]; coefficients per piece in the order
// f, g (complex: Re f, Im f, Re g, Im g), lowest degree first
constexpr int kUserOct = 32;
constexpr int kUserDeg = 7;

// highest B-spline order of the mixed-precision form
constexpr long kMixedMaxBn = 12;

// Derivative-match split of a radial kernel (host, FP64).
struct DMKernel {
  int kind = KIND_LAPLACE;
  double alpha = -1.0;
  double k0r = 0, k0i = 0;
  int Nder = 0;
  double Rdir = 0;
  double a[2 * 31] = {0};  // (re, im) pairs
  double b[2 * 31] = {0};
  double beta = 0;         // KIND_LAPLACE_ERF
};
bool dm_kernel_init(DMKernel &k, int kind, double alpha, double k0r, double k0i, int Nder,
                    double Rdir, std::string &err);
// the Ewald split of 1/r (EwaldSplit = LaplaceOrig) with parameter beta
void erf_kernel_init(DMKernel &k, int Nder, double Rdir, double beta);
// beta with erfc(beta Rdir)/Rdir ~ tol (src/KernelFunctionLaplace.cpp:41-72)
double rdir2beta(double Rdir, double tol);
// f_LR at r2 (FP64): f0[2] (re, im)
void dm_eval_lr(const DMKernel &k, double r2, double f0[2], double f1[2]);

// ------------------------------------------------------------- host math
double bspline(int n, double w);
void bspline_shifts(int n, double t, double *out);  // out[j] = M_n(t + j)
long optimal_grid_size(long n);
void fft_size_block(long Nideal, long nideal, long &N, long &nblk);
void box_dim_center(size_t n, const double *coord, double X0[3], double R0[3]);
void pme_corr_box(const double X0[3], const double R0[3], const long N[3], const double del[3],
                  long Bn, double X[3], double R[3]);
// what the RSEM grid choice reads of the charges: box size and centre
// (box_dim_center) and the norm of the charges, sqrt(sum |q|^2)
struct ChargeStats {
  double X0[3], R0[3];
  double Q0;
};
void charge_stats(size_t n, const double *coord, const double *q, bool complexQ, ChargeStats &st);
// box size and centre from exact per-axis bounds (box_dim_center's seeds)
void box_from_bounds(const double mn[3], const double mx[3], double X0[3], double R0[3]);
double sum_squares(size_t m, const double *v);
bool rsem_static(const ChargeStats &st, long Bn, double alpha, double tol, double Rdir,
                 int nDerivIn, long N[3], long nblk[3], int &Nder);
bool rsem_static_pred_error(size_t n, const double *q, const double *coord, long Bn, double alpha,
                            int Nder, double Rdir, const long N[3], double &eps, bool &reliable);
bool rsem_helmholtz(const ChargeStats &st, long Bn, double k0r, double k0i, double tol,
                    double Rdir, int nDerivIn, long N[3], long nblk[3], int &Nder,
                    std::string &err);
void lambda2(long N, long Bn, double *lam2);

// ------------------------------------------------------------- keywords
struct Keywords {
  std::vector<std::pair<std::string, std::string>> kv;  // insertion order kept
  bool get(const std::string &k, std::string &v) const;
  bool get(const std::string &k, double &v) const;
  bool get(const std::string &k, long &v) const;
  bool get(const std::string &k, int &v) const;
  void set(const std::string &k, const std::string &v);
  void set(const std::string &k, double v);
  void set(const std::string &k, long v);
};

// ------------------------------------------------------------- device memory
// caching allocator (devmem.cpp): blocks return to per-device free lists;
// a block may be freed only when no queued work still uses it
void *dev_alloc(size_t bytes);  // nullptr when the device is out of memory
void dev_free(void *p);
size_t dev_block_bytes(size_t bytes);
void dev_trim();
void dev_stats(size_t *inUse, size_t *cached);

// the calling thread's k-th next dev_alloc fails (k >= 0; -1 disables):
// a hook for the tests of the error paths
void dev_fail_after(long k);

// RAII block from the pool.  Destroyed while an exception unwinds (a set-up
// that failed part way), the block may still be in use by kernels queued
// before the failure: the device is synchronised before the block goes
// back to the pool, where another handle or server slot could receive it.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() {}
  explicit DevBuf(size_t b) { reset(b); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf()
  {
    if (p && std::uncaught_exceptions() > 0) (void)hipDeviceSynchronize();
    dev_free(p);
  }
  void swap(DevBuf &o)
  {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
  }
  void reset(size_t b)
  {
    dev_free(p);
    p = dev_alloc(b);
    bytes = b;
    if (!p) throw Error{NPME_ERR_NOMEM, "device allocation of " + std::to_string(b) + " bytes failed"};
  }
  void *release()
  {
    void *q = p;
    p = nullptr;
    return q;
  }
};

// ------------------------------------------------------------- FFT plan
struct FFTPlan1D {
  int N = 0;
  int nst = 0;
  int radix[24] = {0};
};
FFTPlan1D make_plan(int N);

// ------------------------------------------------------------- device state
struct Handle;

// one CalcV pipeline instance (device buffers for one geometry)
// everything the kernel multiplier G depends on: the grid, the box extent
// (not its centre), the smooth extension and the kernel
struct GridKey {
  long N[3] = {0, 0, 0};
  double X[3] = {0, 0, 0}, del[3] = {0, 0, 0}, ta[3] = {0, 0, 0};
  long Bn = 0;
  bool fp64 = false, complexK = false;
  int kind = -1, Nder = 0;
  double alpha = 0, k0r = 0, k0i = 0, Rdir = 0, beta = 0;
  bool trans = false;           // translated full kernel (box pair)
  double D[3] = {0, 0, 0};      // its translation
  bool operator==(const GridKey &o) const
  {
    for (int d = 0; d < 3; d++)
      if (N[d] != o.N[d] || X[d] != o.X[d] || del[d] != o.del[d] || ta[d] != o.ta[d] ||
          D[d] != o.D[d])
        return false;
    return Bn == o.Bn && fp64 == o.fp64 && complexK == o.complexK && kind == o.kind &&
           Nder == o.Nder && alpha == o.alpha && k0r == o.k0r && k0i == o.k0i && Rdir == o.Rdir &&
           beta == o.beta && trans == o.trans;
  }
};

// host-side state of a user-defined kernel set (borrowed callbacks)
struct UserKernels {
  bool complex = false;
  npme_kfunc_real r[3] = {};      // func, funcLR, funcSR
  npme_kfunc_complex c[3] = {};
};

struct Handle {
  std::mutex mtx;
  bool ready = false;  // the last set-up completed (device state valid)
  // G of the previous set-up, kept by a new set-up of the same handle and
  // reused when its key matches (ResetCoord of a box of the same size)
  GridKey gKey;
  bool gValid = false, gReused = false;
  size_t gBytes = 0;
  // positions copied to the device ahead of a set-up (handle_stage_coords):
  // the set-up of the same n uses them and their bounds
  bool coordStaged = false, stagedBoundsReady = false;
  size_t stagedN = 0;
  double stagedMn[3] = {0, 0, 0}, stagedMx[3] = {0, 0, 0};
  DevBuf stageBuf;
  // configuration
  Keywords kw;
  DMKernel ker;
  // user-defined kernels: the callbacks, the short-range table (device, in
  // the working precision) and its error, and the positions on the host for
  // the exact sum, which evaluates the user's full kernel on the host
  bool userK = false;
  UserKernels user;
  void *d_utab = nullptr;
  int utS = 0;
  double utErr[2] = {0, 0};
  double utR = 0;  // the table's radius: the longest pair of interacting cells
  std::vector<double> coordHost;
  bool complexK = false;
  bool fp64 = false;
  long Bn = 6;
  long N[3] = {0, 0, 0}, nblk[3] = {0, 0, 0}, M[3] = {0, 0, 0}, K3 = 0;
  long Kp = 0;  // row length of the FFT buffers: K3 (or N3) rounded up to 8
  double del[3] = {0, 0, 0}, ta[3] = {4, 4, 4};
  long nNeigh = 2, nCellClust1D = 3;
  size_t n = 0;
  double fself[2] = {0, 0};
  // geometry
  double X0[3], R0[3], X[3], R[3], L[3];
  // cell grid (global cellB grid)
  int nb[3] = {0, 0, 0};
  double csB = 0, csA = 0;
  double boxMin[3] = {0, 0, 0};
  // spreading tiles
  int T[3] = {8, 8, 8};
  int nt[3] = {0, 0, 0};
  // device buffers
  hipStream_t stream = nullptr;
  // the direct sum runs on stream2, concurrently with the reciprocal sum on
  // stream; the host joins them (no cross-stream device waits)
  hipStream_t stream2 = nullptr;
  int device = 0;
  int *d_perm = nullptr;         // [n]
  void *d_src = nullptr;         // [n] double4 (local xyz, q) or mixed: int32 fixed-point xyz + FP32 q
  int fxShift = 29;              // mixed: one cluster edge = 2^fxShift fixed-point units
  void *d_srcRes = nullptr;      // mixed: [n] float4 sub-unit residuals of the positions
  void *d_blockBox = nullptr;    // int4 source segments of the 8-target blocks
  int *d_segStart = nullptr;     // [ceil(n/8) + 1] first segment of each block
  size_t nSegments = 0;
  // sliced-ELLPACK neighbour list (used when it fits the memory budget)
  bool useNlist = false;
  unsigned *d_nlist = nullptr;
  long *d_sliceStart = nullptr;
  int *d_sliceLen = nullptr;
  size_t nlistEntries = 0;
  void *d_qi = nullptr;          // [n] imag part (float/double) for complex kernels
  void *d_qs = nullptr;          // [n] sorted charges (re) for spreading, working precision
  void *d_qsi = nullptr;         // [n] sorted charges (im), complex kernels
  int4 *d_cell = nullptr;        // [n] cell coords (x,y,z), pad
  int *d_cellStart = nullptr;    // [ncell + 1]
  int4 *d_ms = nullptr;          // [n] stencil start u0 per dim (u = m + N/4)
  void *d_frac = nullptr;        // [n] float4/double4 fractional spline arguments
  int *d_tileStart = nullptr;    // [ntiles + 1]
  int *d_tileList = nullptr;     // [nlist]
  int TI = 16;                   // interpolation tile edge
  int nti[3] = {0, 0, 0};
  int *d_itileStart = nullptr;   // [ntiles + 1] charges per interpolation tile
  int *d_itileList = nullptr;    // [n]
  void *d_grid = nullptr;        // Q / theta: M reals (real) or M complex
  void *d_buf1 = nullptr;        // M1 x M2 x Kp complex (rows of K3 or N3 values)
  void *d_buf2 = nullptr;        // M1 x N2 x Kp complex
  void *d_G = nullptr;           // k-space multiplier (N1/2+1)(N2/2+1)(K3 or N3/2+1)
  void *d_tw[3] = {nullptr, nullptr, nullptr};  // twiddle tables per axis length (in d_tables)
  void *d_tables = nullptr;  // the host-built tables of a set-up, one block
  double *d_Vdir = nullptr;      // [n][4] (x2 for complex) direct sum, sorted order
  double *d_V = nullptr;         // [n][4] (x2) output staging, input order
  double *d_qin = nullptr;       // [n] (x2) input charges staging
  double *d_coordS = nullptr;    // [n][3] FP64 input coordinates, input order (exact sums)
  double4 *d_xs = nullptr;       // [n] FP64 coordinates in sorted order (pair search)
  unsigned *d_cellFlags = nullptr;  // LaplaceOrig: interacting (cell, stencil offset) bits
  FFTPlan1D plan[3];
  size_t deviceBytes = 0;
  bool deviceSetup = false;      // per-charge set-up ran on the device
  bool fftPow8 = true;           // compile-time FFT kernels for N = 64, 512
  // wall times of the last set-up (ms): host geometry, previous state freed,
  // cells and sort, neighbour list, per-charge arrays and tile lists, FFT
  // buffers and kernel grid
  double setupMs[6] = {0, 0, 0, 0, 0, 0};
  // CalcV as a graph: the second CalcV call with the same charge and V
  // pointers on one set-up captures the launches (the direct-sum branch and
  // the reciprocal branch joined by an event) and later calls replay it.
  // Any set-up or release of the device state drops the graph.
  bool useGraph = true;
  hipEvent_t evFork = nullptr, evJoin = nullptr;  // fork/join of the two streams
  hipGraphExec_t calcvExec = nullptr;
  const void *graphQ = nullptr, *graphV = nullptr;  // pointers of the last CalcV
  int graphCalls = 0;  // CalcV calls in a row with those pointers
  // box pair (two point sets, translated kernel): the charges are the
  // targets followed by the sources, each relative to the centre of its own
  // bounding box; the kernel grid holds the full kernel at r + transD with
  // transD = source centre - target centre, G over the full grid; no direct
  // sum and no self term.  transFlags: the k-reversal / conjugation bits of
  // the convolution pass for the reversal transRev (bits x, y, z) of CalcV.
  bool trans = false;
  double transD[3] = {0, 0, 0};
  size_t nTarget = 0;
  int transRev = 0, transFlags = 0;
  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> ev;
  double stage_ms[16] = {0};
  int nstage = 0;
};

void handle_setup(Handle &h, const Keywords &kw, size_t n, const double *coord);
// box pairs: the reversal (bits x, y, z) used by the next CalcV
void handle_set_trans_rev(Handle &h, int rev);
// exp(-2 pi i t / N), t = 0 .. N-1, as (re, im) pairs in FP64
std::vector<double> twiddles(int N);
// NPME_AddMissingKeywords: defaults, validation, RSEM grid (throws Error)
void add_missing_keywords(Keywords &kw, size_t nCharge, bool isChargeReal, const double *coord,
                          const double *chargeReal, const double *chargeComplex,
                          const ChargeStats *stats = nullptr);

// geometry of the cell partition and the grids, shared by the host and the
// device set-up (geometry.hip)
struct GeoParams {
  double mn[3];            // low corner of the cell box
  double csA, csB;         // cluster and cell edges
  unsigned long nA[3];     // clusters per axis
  unsigned long nP;        // cells per cluster edge
  int nb[3];               // cells per axis
  double NoverL[3], R[3];  // spline argument w = N/L (x - R)
  long N[3];
  long Bn;
  double fxScale;          // mixed form: 2^fxShift / csA
  int fxShift;
  int ts[3];               // spreading tile edge per axis (16 x 8 x 8 or 8^3)
  int nt[3];               // spreading tiles per axis
  int TI, nti[3];          // interpolation tiles
  int globalFrame;         // positions relative to mn (neighbour-list walk), not to the cluster
};
// whether any of the n int4 fixed-point positions lies within thr units of
// v on all three axes (wrapped int32 differences)
bool sentinel_near_any(size_t n, const void *d_src, int4 v, int thr, hipStream_t s);
void geometry_cells_device(const GeoParams &g, size_t n, const double *d_coord, int *d_perm,
                           int4 *d_cellOf, int *d_cellSE, hipStream_t s);
void geometry_charges_device(const GeoParams &g, bool fp64, size_t n, const double *d_coord,
                             const int *d_perm, const int4 *d_cellOf, void *d_src, void *d_srcRes,
                             int4 *d_ms,
                             void *d_frac, DevBuf &tileStart, DevBuf &tileList, DevBuf &itileStart,
                             DevBuf &itileList, size_t &nTileEntries, hipStream_t s);
// cell-row segments of the direct sum without a neighbour list, on the
// device (segStart: nBlocks + 1 ints, nBlocks = ceil(n / 8)); returns their
// number and the largest cluster offset (at least 1)
long cell_segments_device(const GeoParams &g, size_t n, double Rdir, const double4 *d_xs,
                          const int4 *d_cellOf, const int *d_cellSE, int *d_segStart,
                          DevBuf &segs, int &maxOff, hipStream_t s);
size_t nlist_slices_device(size_t n, const int *d_count, int *d_sliceLen, long *d_sliceStart,
                           hipStream_t s);
// coordinates in sorted order: xs[s] = coord[perm[s]]
// exclusive prefix sums of m values into out[0..m-1]; returns out[m-1]
// (the total when in[m-1] is 0)
long scan_words_device(const long *d_in, long *d_out, size_t m, hipStream_t s);
// exact per-axis bounds of the n positions on the device (NaN if any is NaN)
void coord_bounds_device(size_t n, const double *d_coord, double mn[3], double mx[3], hipStream_t s);
// the same queued without waiting: scratch of coord_bounds_scratch(n)
// bytes; the six results (minima, maxima) are its last six doubles
size_t coord_bounds_scratch(size_t n);
void coord_bounds_launch(size_t n, const double *d_coord, double *d_scratch, hipStream_t s);
void sorted_coords_device(size_t n, const int *d_perm, const double *d_coord, double4 *d_xs,
                          hipStream_t s);
// candidate pairs of the cell stencil: sum over cells of count * (counts of
// the (2 nNeigh + 1)^3 cells around it)
double stencil_candidates_device(const int nb[3], int nNeigh, const int *d_cellSE, hipStream_t s);
// releases the device buffers of a set-up; keepG keeps G and its key for
// the next set-up of the same handle
void handle_free_device(Handle &h, bool keepG = false);
// copies n positions to the device and queues their bounds, without
// waiting, for the next handle_setup of the same positions (the caller
// computes what else it needs meanwhile); releases the previous set-up
void handle_stage_coords(Handle &h, size_t n, const double *coord);
// waits for the staged positions' bounds
void handle_staged_bounds(Handle &h, double mn[3], double mx[3]);
// destroys the handle's streams and timing events (after handle_free_device)
void handle_destroy_streams(Handle &h);
void handle_calcv(Handle &h, const double *d_q, double *d_Vout);
// CalcV of m handles (charges in d_qin, V to d_V of each) as one captured
// graph launched on sb (events: one start, one per handle)
void handle_calcv_batch(Handle *const *hs, int m, hipStream_t sb, hipEvent_t evStart,
                        const hipEvent_t *evDone);

void handle_exact(Handle &h, const double *d_q, size_t nt, const long *d_idx, double *d_Vout);


// ------------------------------------------------------------- launches
// FFT passes (fft.hip)
struct PassGeom {
  int N;           // FFT length
  int Lin, Lout;   // valid input / output positions
  long nCols;      // number of columns
  long nInner;     // columns per outer group
  long inOuter, inInner, inPos;     // input strides (elements)
  long outOuter, outInner, outPos;  // output strides
  long validInner = -1;  // columns with inner index >= this are padding (-1: none)
};
// pow8: the compile-time kernels may be used for N = 64 and 512 (radix-8 plans)
void fft_pass_complex(bool fp64, const PassGeom &g, const FFTPlan1D &plan, const void *tw,
                      const void *in, void *out, bool backward, hipStream_t s, bool pow8 = true);
void fft_pass_r2c_rows(bool fp64, int N, int Lin, long nRows, long inRowStride,
                       long outRowStride, const FFTPlan1D &plan, const void *tw, const void *in,
                       void *out, hipStream_t s);
void fft_pass_c2r_rows(bool fp64, int N, int Lout, long nRows, long inRowStride,
                       long outRowStride, const FFTPlan1D &plan, const void *tw, const void *in,
                       void *out, hipStream_t s);
// middle pass: forward FFT along the axis, multiply by G, backward FFT.
// fullG: G is complex over the full N1 x N2 x Kv grid of a translated
// kernel, gflags: k -> -k on axis 1, 2, 3 (bits 1, 2, 4), conjugate (bit 8)
void fft_pass_convolve(bool fp64, bool complexG, int N, int Lin, long nCols, long pos_stride,
                       int N2, int G2, long K3, long Kv, const FFTPlan1D &plan, const void *tw,
                       const void *G, void *data, hipStream_t s, bool pow8 = true,
                       bool fullG = false, int gflags = 0);
// even-extension DFT of the compact kernel grid along one axis (set-up,
// FP64): complex lines, and real lines ([.] doubles) two per transform
void fft_pass_even(int N, long nCols, long nInner, long outer, long inner, long pos,
                   const FFTPlan1D &plan, const void *tw, void *data, hipStream_t s);
void fft_pass_even_real(int N, long nCols, long nInner, long outer, long inner, long pos,
                        const FFTPlan1D &plan, const void *tw, double *data, hipStream_t s);

// particle kernels (particles.hip)
struct DirectParams {
  int kind;
  int Nder;
  double Rdir;
  double alpha;
  double beta;  // KIND_LAPLACE_ERF
  double k0r, k0i;
  double a[2 * 31], b[2 * 31];
  int nNeigh;
  double cs;
  int nb[3];
  double mn[3];   // low corner of the cell grid
  double csA;     // cluster edge (nCellClust1D cells)
  double cellTol; // bound on a charge's distance outside its cell's box
  int fxShift;    // mixed form: one cluster edge = 2^fxShift fixed-point units
  double fxInv;   // cluster edge / 2^fxShift
  const void *utab = nullptr;  // user kernels: short-range table, S pieces per octave
  int utS = 0;
  double utR = 0;              // its radius (r / utR in (0, 1])
  // neighbour-list entries are source byte offsets: log2 of the source
  // element size (5 for FP64 double4, 4 for the mixed form's 16 bytes)
  int nlEshift = 4;
};
double handle_count_pairs(Handle &h);
// LaplaceOrig's direct sum over the reference's interacting cell pairs (no
// cutoff): bit words per cell, the flags, and the FP64 sum into Vdir
int refpair_words(int nNeigh);
void launch_cellpair_contact(const DirectParams &pp, const int *cellStart, const double4 *xs,
                             unsigned *flags, hipStream_t s);
void launch_direct_refpairs(const DirectParams &pp, size_t n, const int4 *cell, const int *cellStart,
                            const double4 *xs, const int *perm, const double *q,
                            const unsigned *flags, double *Vdir, hipStream_t s);
// kernels whose direct sum takes the reference's pair set (no cutoff):
// LaplaceOrig's erfc and the user-defined short-range parts
inline bool kind_ref_pairs(int kind) { return kind == KIND_LAPLACE_ERF || kind == KIND_USER || kind == KIND_USER_CX; }
// user-defined kernels, host side (genfunc.cpp)
void user_fself(const UserKernels &u, double fself[2]);
void user_check_radial(const UserKernels &u, double Rdir);
void user_build_table(const UserKernels &u, double Rdir, std::vector<double> &coef, int &S,
                      double err[2]);
void user_fgrid_host(const UserKernels &u, const long Mf[3], const std::vector<double> &tab,
                     std::vector<double> &F);
void user_exact_host(const UserKernels &u, size_t n, const double *coord, const double *q,
                     size_t nt, const long *idx, double *V);
// the same count with the partials in a static device array (noalloc.hip)
double handle_count_pairs_noalloc(Handle &h);
DirectParams direct_params(const Handle &h);
void launch_gather_q(bool fp64, bool complexK, size_t n, const int *perm, const double *q_in,
                     void *qre, void *qim, hipStream_t s);
// the charges in sorted order for both sums in one pass: the direct sum's
// packed source charges (src[s].w, qi) and the spreading copy (qre, qim)
void launch_gather_both(bool fp64, bool complexK, size_t n, const int *perm, const double *q_in,
                        void *src, void *qi, void *qre, void *qim, hipStream_t s);
void launch_direct(bool fp64, const DirectParams &p, size_t n, const void *src, const void *res,
                   const void *qi, const int4 *cell, const int *segStart, const void *segs, int nP, double csA,
                   double *Vdir, hipStream_t s);
// padding entry of the neighbour list: the sentinel source n, as a byte offset
unsigned nlist_pad(size_t sentinel, int eshift);
// neighbour-list slices: a team of kNlTeam lanes walks one target's list
// (lane j of the team takes entries j, j + kNlTeam, ...), so a slice (one
// wavefront) holds kNlSlice targets.  A slice's length is its longest list
// rounded up to 4 kNlTeam; sliceLen holds that length divided by kNlTeam,
// the entries each lane walks (a multiple of 4), and a slice takes
// 64 sliceLen words
constexpr int kNlTeamLog = 2;
constexpr int kNlTeam = 1 << kNlTeamLog;
constexpr int kNlSlice = 64 / kNlTeam;
// per-lane length of a slice whose longest list has m entries
__host__ __device__ constexpr int nlist_lane_len(int m)
{
  return ((m + 4 * kNlTeam - 1) / (4 * kNlTeam)) * 4;
}
// scratch of one list build, from its count pass to its fill pass: the
// hit masks of the wavefront-per-cell walk and their per-cell offsets
struct NlistWork {
  DevBuf maskStart, masks;
  bool masked = false;
};
// count pass (fill false: count[s] per target) or fill pass of the
// neighbour list; both passes of one build take the same work
void launch_nlist(bool fill, const DirectParams &p, size_t n, const int4 *cell, const int *cellStart,
                  const double4 *xs, int nP, int *count, const long *sliceStart, unsigned *list,
                  NlistWork &work, hipStream_t s);
void launch_direct_nl(bool fp64, const DirectParams &p, size_t n, const void *src, const void *res,
                      const void *qi, const long *sliceStart, const int *sliceLen, const unsigned *list,
                      double csA, double *Vdir, hipStream_t s);
void launch_count_pairs(const DirectParams &p, size_t n, const int4 *cell, const int *cellStart,
                        const double4 *xs, unsigned long long *total, hipStream_t s);
void launch_spread(bool fp64, bool complexK, long Bn, const long M[3], const int T[3],
                   const int nt[3], const int *tileStart, const int *tileList, const void *qre,
                   const void *qim, const int4 *ms, const void *frac, void *grid, hipStream_t s);
void launch_interp(bool fp64, bool complexK, long Bn, const long M[3], const double C[3], int T,
                   const int nt[3], const int *tileStart, const int *tileList, const int4 *ms,
                   const void *frac, const void *grid, const double *Vdir, const int *perm,
                   const double *q_in, const double fself[2], double *Vout, hipStream_t s);
void launch_exact(int kind, double alpha, double k0r, double k0i, size_t n,
                  const double *coordS, const int *perm, const double *q_in, size_t nt,
                  const long *idx, double *Vout, hipStream_t s);
// F on the compact real-space grid: double per point for the real kernels,
// double2 for Helmholtz
void launch_fgrid_real_space(const DMKernel &k, const long Mf[3], const double *d_tab,
                             void *F, hipStream_t s);
// translated full kernel grid (box pairs): f(r) = func(|r + D|) sampled on
// the full grid from per-axis tables [x1 + D | x2 + D | T1 | T2], (re, im)
// per point; G = F lam^2 / (N1 N2 N3) over N1 x N2 x G3 (complex)
void launch_fgrid_trans(const DMKernel &km, const long N[3], const double *d_tab, double2 *F,
                        hipStream_t s);
void launch_make_G_full(bool fp64, const long N[3], long G3, const double2 *F, const double *lam,
                        void *G, hipStream_t s);
// host tables of the translated grid (src/RecSumGrid.cpp:663-794,
// :1024-1042, :1193-1231) and the separation check
// (src/SupportFunctions.cpp:687-718)
std::vector<double> fulltrans_tables(const long N[3], const double ta[3], const double del[3],
                                     const double X[3], const double D[3]);
double min_distance_r0(const double R0[3], const double l[3]);
// f_0 of a user kernel (which: 0 func, 1 funcLR) on a grid given by
// per-axis tables [x1 | x2 | T1 | T2] of n[d] values: F[n0][n1][n2] doubles
// (real) or (re, im) pairs (complex, or cxOut)
void user_fgrid_host_n(const UserKernels &u, int which, const long n[3], const std::vector<double> &tab,
                       bool cxOut, std::vector<double> &F);
// forward 3-D FFT in FP64 of a complex [N0][N1][N2] grid (twiddles tw64 of
// each axis); a holds the input, b is scratch; returns the buffer holding
// the result
double2 *fft3d_forward_fp64(const long N[3], const FFTPlan1D plan[3], const void *const tw64[3],
                            double2 *a, double2 *b, hipStream_t s);
void launch_make_G(bool fp64, bool complexK, const long Mf[3], long K3, const double *F,
                   const double *lam, void *G, hipStream_t s);

}  // namespace npme

// the opaque types of the C ABI
struct npme_keywords {
  npme::Keywords kw;
};
struct npme_handle {
  npme::Handle h;
};
