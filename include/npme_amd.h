/*
 * npme_amd.h -- C ABI of the MI355X NPME potential evaluation.
 *
 * Drop-in boundary for the NPME library calls of ElkingD/npme
 * (src/Interface.h).  Plain pointers and sizes only.  Complex data is
 * interleaved (re, im) doubles, the memory layout of C99 `_Complex double`
 * used by the reference.  V arrays are [nCharge][4] = (V, dV/dx, dV/dy,
 * dV/dz) per charge, complex ones [nCharge][4][2].
 *
 * Every entry point returns an int status (NPME_OK = 0) and never exits the
 * process; npme_last_error() returns the message of the last failure on the
 * calling thread.  One handle owns one device-resident state (grids, cell
 * lists, twiddles); calls on one handle are serialised by the library,
 * distinct handles may be used concurrently.
 */
#ifndef NPME_AMD_H
#define NPME_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define NPME_OK              0
#define NPME_ERR_KEYWORD     1   /* missing / contradictory keyword        */
#define NPME_ERR_GRID        2   /* FFT grid or B-spline constraint        */
#define NPME_ERR_NOMEM       3   /* device memory / admission refused      */
#define NPME_ERR_DEVICE      4   /* HIP runtime error                      */
#define NPME_ERR_ARG         5   /* bad argument (size mismatch, NULL ...) */
#define NPME_ERR_UNSUPPORTED 6   /* combination not provided               */
#define NPME_ERR_IO          7   /* file cannot be opened, read or parsed  */

typedef struct npme_keywords npme_keywords;
typedef struct npme_handle   npme_handle;

/* ---- keyword set: NPME_KeywordInput (src/ReadPrint.h:28-102) ---------- */
npme_keywords *npme_keywords_create(void);
void npme_keywords_destroy(npme_keywords *kw);
/* AddKeyword (src/ReadPrint.h:64-98); replaces an existing value */
int  npme_keywords_set(npme_keywords *kw, const char *key, const char *value);
/* GetValue: copies the value, returns 1 if the key exists, 0 if not */
int  npme_keywords_get(const npme_keywords *kw, const char *key, char *buf,
                       size_t buflen);
/* NPME_KeywordInput(file): "@Keyword" section, "key value [comment]" lines
 * (src/ReadPrint.cpp:38-66) */
int  npme_keywords_read_file(npme_keywords *kw, const char *path);
/* number of keywords and the i-th key (for printing / copying) */
int  npme_keywords_count(const npme_keywords *kw);
int  npme_keywords_key(const npme_keywords *kw, int i, char *buf, size_t buflen);

/* ---- NPME_AddMissingKeywords (src/Interface.h:37-51,
 *      src/Interface.cpp:53-517): defaults, validation and the RSEM grid
 *      choice.  isChargeReal selects chargeReal[n] or chargeComplex[n][2];
 *      the other pointer may be NULL. ------------------------------------ */
int npme_add_missing_keywords(npme_keywords *kw, size_t nCharge,
                              int isChargeReal, const double *coord,
                              const double *chargeReal,
                              const double *chargeComplex);

/* ---- set-up: NPME_Interface_SetUpRealKernel / SetUpComplexKernel
 *      (src/Interface.h:800-829) selecting NPME_InterfaceReal_Laplace_DM,
 *      NPME_InterfaceReal_GenFunc (r^alpha) or
 *      NPME_InterfaceComplex_Helmholtz_DM from the keywords (funcType,
 *      EwaldSplit), then ::SetUp (src/Interface.cpp:1606-1671, :1847-1956,
 *      :2661-2791).  The keyword set must already hold every value
 *      (npme_add_missing_keywords).  Extension keyword: precision = mixed
 *      (default; FP32 spreading and FFT, FP32 direct-sum pair terms summed
 *      in FP32 partials of 8 that are carried into FP64) or fp64. ------- */
int npme_create(npme_handle **h, const npme_keywords *kw, size_t nCharge,
                const double *coord);

/* ---- user-defined kernels: NPME_KfuncReal / NPME_KfuncComplex
 * (src/KernelFunction.h:25-116) as C callbacks.  ctx is passed back
 * unchanged.  The in-place conventions are the reference's:
 *   value(ctx, n, x_f0, y, z): x_f0[i] holds x on entry, f(x, y, z) on
 *     return (complex: x_f0_r holds x, then Re f; f0_i receives Im f);
 *   grad(ctx, n, f0, x_fX, y_fY, z_fZ): f0 receives f; x_fX, y_fY, z_fZ
 *     hold x, y, z on entry and df/dx, df/dy, df/dz on return (complex: the
 *     real parts in place, the imaginary parts in fX_i, fY_i, fZ_i).
 * The library may call a callback from several host threads at once (the
 * reference calls Calc from OpenMP threads).  The callbacks are borrowed:
 * they must stay valid until the handle is destroyed. */
typedef struct {
  void *ctx;
  void (*value)(void *ctx, size_t n, double *x_f0, const double *y, const double *z);
  void (*grad)(void *ctx, size_t n, double *f0, double *x_fX, double *y_fY, double *z_fZ);
} npme_kfunc_real;
typedef struct {
  void *ctx;
  void (*value)(void *ctx, size_t n, double *x_f0_r, double *f0_i, const double *y,
                const double *z);
  void (*grad)(void *ctx, size_t n, double *f0_r, double *f0_i, double *x_fX_r, double *fX_i,
               double *y_fY_r, double *fY_i, double *z_fZ_r, double *fZ_i);
} npme_kfunc_complex;

/* NPME_InterfaceReal_GenFunc::SetUp (src/Interface.h:291-296,
 * src/Interface.cpp:1606-1671) and NPME_InterfaceComplex_GenFunc::SetUp
 * (src/Interface.h:563-650) with user kernels: func is the full kernel
 * (CalcV_exact, evaluated on the host as the reference does), funcLR the
 * smooth long-range part, sampled on the kernel grid on the host at every
 * set-up (and fself = funcLR(0, 0, 0)), funcSR the short-range part, which
 * must be radial: it is tabulated on the host as piecewise polynomials in r
 * over [0, R_t], R_t = sqrt(3) (nNeigh + 1) Rdir / nNeigh, the longest pair
 * of the interacting cell pairs on which the reference calls it
 * (src/PotentialGenFunc.cpp:523-645), and evaluated from that table by the
 * device direct sum over those pairs in FP64 (npme_get_table_error reports
 * the table's error).  The keywords must hold
 * the grid and split (Rdir, BnOrder, N1-3, n1-3, del1-3, a1-3, nNeigh,
 * nCellClust1D); funcType, EwaldSplit and nDeriv are not used. */
int npme_create_genfunc_real(npme_handle **h, const npme_keywords *kw, size_t nCharge,
                             const double *coord, const npme_kfunc_real *func,
                             const npme_kfunc_real *funcLR, const npme_kfunc_real *funcSR);
int npme_create_genfunc_complex(npme_handle **h, const npme_keywords *kw, size_t nCharge,
                                const double *coord, const npme_kfunc_complex *func,
                                const npme_kfunc_complex *funcLR,
                                const npme_kfunc_complex *funcSR);
/* the derivative-match split of a built-in kernel, as SetParm of
 * NPME_Kfunc_{Laplace,Ralpha,Helmholtz}_{LR,SR}_DM computes it
 * (src/FunctionDerivMatch.cpp:42-450): inside Rdir the long-range part is
 * sum_k a_k r^2k and its (1/r) d/dr is sum_k b_k r^2k, k = 0..Nder; a and b
 * are [Nder + 1][2] (re, im).  funcType "Laplace", "Ralpha" (alpha) or
 * "Helmholtz" (k0_r, k0_i). */
int npme_dm_coefficients(const char *funcType, double alpha, double k0_r, double k0_i,
                         int Nder, double Rdir, double *a, double *b);
/* NPME_EwaldSplitOrig_Rdir2Beta (src/KernelFunctionLaplace.cpp:41-72): the
 * beta of the erf/erfc split of 1/r for a cutoff Rdir and tolerance tol */
int npme_ewald_rdir2beta(double Rdir, double tol, double *beta);
/* largest error of the short-range table of a user kernel against funcSR at
 * check points between the interpolation nodes, relative to the largest
 * |f| (err[0]) and |f'(r)/r| (err[1]) over the same octave of r; both 0 for
 * the built-in kernels.  No reference counterpart. */
int npme_get_table_error(const npme_handle *h, double err[2]);

/* ::ResetCoord (src/Interface.h:186-187; src/Interface.cpp:1958-1987) */
int npme_reset_coord(npme_handle *h, size_t nCharge, const double *coord);
void npme_destroy(npme_handle *h);

/* ---- ::CalcV (src/Interface.h:373-375 real, :715-717 complex).
 * Host arrays.  coord must be the geometry given to create/reset_coord
 * (the device copy is used); charge and coord are not modified. -------- */
int npme_calcv_real(npme_handle *h, double *V, size_t nCharge,
                    const double *charge, const double *coord);
int npme_calcv_complex(npme_handle *h, double *V, size_t nCharge,
                       const double *charge, const double *coord);

/* ---- ::CalcV_exact (src/Interface.h:380-383, :722-725): exact O(N^2)
 * FP64 sum of the full kernel on the GPU. ------------------------------- */
int npme_calcv_exact_real(npme_handle *h, double *V, size_t nCharge,
                          const double *charge, const double *coord);
int npme_calcv_exact_complex(npme_handle *h, double *V, size_t nCharge,
                             const double *charge, const double *coord);
/* exact sum for the targets idx[0..nTarget) only; V is [nTarget][4]
 * (complex: [nTarget][4][2]).  No reference counterpart: the subset form of
 * CalcV_exact used to check large systems. */
int npme_calcv_exact_subset(npme_handle *h, double *V, size_t nTarget,
                            const long *idx, const double *charge);

/* ---- device-resident CalcV: dCharge and dV are device pointers
 * (charges in input order, FP64; V [n][4] FP64 or [n][4][2]).  dCharge
 * must be complete when the call is made (its producer synchronised): the
 * library reads it from its own streams.  No host traffic; returns when
 * the result is complete.  No reference counterpart: the same CalcV with
 * inputs already in HBM. ------------------------------------------------ */
int npme_calcv_device(npme_handle *h, double *dV, const double *dCharge);

/* ---- parameters: Get_N1 ... Get_n3, Get_BnOrder (src/Interface.h:89-115),
 * plus Nder and the device grid footprint in GB. ----------------------- */
int npme_get_params(const npme_handle *h, long N[3], long nblk[3], long *Bn,
                    int *Nder, double *Rdir, double *deviceGB);
/* ::Print (src/Interface.h:385): human-readable summary into buf */
int npme_print(const npme_handle *h, char *buf, size_t buflen);

/* ---- number of ordered pairs i != j closer than Rdir, the work of the
 * direct sum (the reference logs it as nInteract,
 * src/Interface.cpp:1329-1334). ---------------------------------------- */
int npme_count_pairs(npme_handle *h, double *pairs);
/* the same count without any device allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  No counterpart in the reference.  Its partial counts
 * live in a static device array of 1 MiB per GPU (131,072 workgroup
 * partials), so it accepts sets of at most 33,554,432 charges and returns
 * NPME_ERR_ARG above that.  All calls on one GPU share that array: calls
 * that overlap on one GPU race. */
int npme_count_pairs_noalloc(npme_handle *h, double *pairs);

/* ---- integer indexing of the last set-up, per charge in input order:
 * clusterKey[i] and cellKey[i] are the embedded cluster and cell indices of
 * NPME_PointIntoEmbeddedCell (src/PartitionEmbeddedBox.cpp:266-455, cell
 * index of src/PartitionBox.cpp:311-351); mStart[i][3] is the first grid
 * index m of the charge's spline window per axis
 * (NPME_GetQpme_DFT_StartIndex, src/RecSumSupportFunctions.h:123-127, with
 * wX = N1/L1 (x - R0x) of src/RecSumQ.cpp:106-108); order[s] is the input
 * index of the s-th charge of the working order (a stable sort by cellKey).
 * Any output may be NULL.  No reference counterpart (the reference keeps
 * these inside the interface object); it exposes the set-up's integer
 * decisions for bit-exact checks. ---------------------------------------- */
int npme_get_indices(npme_handle *h, size_t nCharge, unsigned long long *clusterKey,
                     unsigned long long *cellKey, long *mStart, long *order);

/* ---- CalcV launch mode: on (default), the second and later CalcV calls of
 * one set-up with the same charge and V pointers replay a HIP graph captured
 * from the call's launches (the direct-sum and reciprocal branches joined on
 * the device); off, every call launches the kernels.  Results are the same.
 * No reference counterpart. ------------------------------------------------ */
int npme_set_calcv_graph(npme_handle *h, int on);

/* ---- two-box (translated-kernel) grids ------------------------------------
 * The reference declares NPME_RecSumGrid_FullTrans_Func (src/RecSumGrid.h:
 * 99-112, body src/RecSumGrid.cpp:663-794) and
 * NPME_RSTP_CalcTheta_FullF_ReverseSign (src/RecSumV.h:102-113, body
 * src/RecSumV.cpp:943-1009) for box-box interactions and calls neither from
 * CalcV.  The two functions below are their device forms on host arrays in
 * the reference's layouts; complex arrays are interleaved (re, im).
 *
 * npme_recsumgrid_fulltrans_func: f(r) = func(|r + R0|) on the full
 * N1 x N2 x N3 grid of the box [-X - del1, X + del1) x ... (sample n at
 * s L/N, s = n or n - N for n >= N/2; f[(n1 N2 + n2) N3 + n3]), summed over
 * the 8 overlapped images with the top-hat weights of parameter a1..a3.  The
 * kernel is the full kernel named by the keywords funcType (Laplace, Ralpha
 * with alpha, Helmholtz with k0_r, k0_i), evaluated on the GPU in FP64;
 * _user takes exactly one of func / cfunc (NPME_KfuncReal /
 * NPME_KfuncComplex callbacks) and samples it on the host.  Where the
 * reference prints rMin and exits (separation check, src/RecSumGrid.cpp:
 * 694-705), these return NPME_ERR_ARG.
 *
 * npme_rstp_calctheta_fullF_reversesign: T = compact theta [N1/2][N2/2][N3/2]
 * (DFT order, m < 0 at m + N/2) from the compact charge grid Qr (same
 * layout) and F, the full forward DFT [N1][N2][N3] of the kernel grid, with
 * the |lambda|^2 arrays of the axes (N_d values, DFT order); reverseSignX
 * etc. use F as the grid of R0 with that component negated.  The reference's
 * twiddle arrays C1..C3 (fixed by N) and its scratch X are not arguments; the
 * convolution runs in FP64 on the GPU. */
int npme_recsumgrid_fulltrans_func(double *f, long N1, double a1, double del1, long N2, double a2,
                                   double del2, long N3, double a3, double del3, double X, double Y,
                                   double Z, const npme_keywords *kernel, const double R0[3]);
int npme_recsumgrid_fulltrans_func_user(double *f, long N1, double a1, double del1, long N2,
                                        double a2, double del2, long N3, double a3, double del3,
                                        double X, double Y, double Z, const npme_kfunc_real *func,
                                        const npme_kfunc_complex *cfunc, const double R0[3]);
int npme_rstp_calctheta_fullF_reversesign(double *T, const double *F, const double *Qr, long N1,
                                          long N2, long N3, const double *lamda1_2,
                                          const double *lamda2_2, const double *lamda3_2,
                                          int reverseSignX, int reverseSignY, int reverseSignZ);

/* ---- box pairs: V_t = sum_s q_s f(|r_t - r_s|) and its gradient at the
 * targets t of one box from the sources s of another, entirely on the kernel
 * grid (the reference has the two functions above and no driver: this
 * pairing is the library's).  Each set is taken relative to the centre of
 * its own bounding box, so the grid spans the larger box, not the pair; the
 * kernel grid is the full kernel translated by the centre separation.  The
 * keywords are those of npme_create (funcType, ..., BnOrder, precision);
 * N1, N2, N3 are required (no error model), n1..n3, del1..3 (default
 * Rdir, 8), a1..3 default as NPME_AddMissingKeywords.  Accuracy needs the
 * grid box of edges 2 (X + del) to stay clear of the separation; set-up
 * returns NPME_ERR_ARG when its distance rMin is below 1e-6.
 * npme_boxpair_calcv_*: V [nTarget][4] (complex: [nTarget][4][2]);
 * reverseSign (NULL: none) evaluates the pair with the separation negated
 * along the flagged axes (the source box mirrored through the target box
 * centre) with the same kernel grid.  npme_boxpair_separation gives R0 =
 * target centre - source centre (f(r) = func(|r + R0|), r = the local
 * target-minus-source separation) and rMin.  Box-pair handles reject
 * npme_calcv_* and npme_reset_coord; npme_destroy releases them. */
int npme_create_boxpair(npme_handle **out, const npme_keywords *kw, size_t nTarget,
                        const double *coordTarget, size_t nSource, const double *coordSource);
int npme_create_boxpair_genfunc_real(npme_handle **out, const npme_keywords *kw, size_t nTarget,
                                     const double *coordTarget, size_t nSource,
                                     const double *coordSource, const npme_kfunc_real *func);
int npme_create_boxpair_genfunc_complex(npme_handle **out, const npme_keywords *kw, size_t nTarget,
                                        const double *coordTarget, size_t nSource,
                                        const double *coordSource, const npme_kfunc_complex *func);
int npme_boxpair_reset_coord(npme_handle *h, size_t nTarget, const double *coordTarget,
                             size_t nSource, const double *coordSource);
int npme_boxpair_calcv_real(npme_handle *h, double *V, size_t nTarget, size_t nSource,
                            const double *chargeSource, const int reverseSign[3]);
int npme_boxpair_calcv_complex(npme_handle *h, double *V, size_t nTarget, size_t nSource,
                               const double *chargeSource, const int reverseSign[3]);
int npme_boxpair_separation(const npme_handle *h, double R0[3], double *rMin);

/* ---- per-stage device times of the last CalcV when profiling is on:
 * ms[0] gather (the charges in sorted order for both sums), [1] direct
 * sum, [2] spread, [3..7] FFT passes A..E, [8] interpolate+combine (a box
 * pair, which has no direct sum, omits [1]); returns the number of stages
 * written. ------- */
int npme_set_profiling(npme_handle *h, int on);
int npme_get_stage_times(const npme_handle *h, double *ms, int n);

/* ---- serving: a persistent loop of nSlots worker slots on the current
 * GPU that evaluates independent CalcV requests (any size and kernel)
 * concurrently.  A request is admitted first come first served when a slot
 * is free and a provisional memory estimate fits the budget (memGB; <= 0
 * means 80% of the free device memory); once NPME_AddMissingKeywords has
 * chosen its grid (RSEM for its tol) its reservation becomes its own
 * footprint (grids, per-charge arrays, neighbour list at its density),
 * waiting for memory if that is larger, and after the device set-up the
 * memory its set-up holds; then CalcV runs and V is written.  coord,
 * charge and V must stay valid until npme_server_wait returns for its
 * ticket.  A result does not depend on the other requests in flight.  No
 * reference counterpart: the reference evaluates one charge set per call
 * (src/Interface.h:373-375). ---------------------------------------------- */
typedef struct npme_server npme_server;
typedef struct {
  const char *funcType;   /* "Laplace", "Ralpha" or "Helmholtz"            */
  double alpha;           /* r^alpha exponent (Ralpha)                     */
  double k0_r, k0_i;      /* wave number (Helmholtz)                       */
  double tol;             /* requested relative RMS error of V (RSEM)      */
  double Rdir;            /* short-range cutoff                            */
  long BnOrder;           /* B-spline order; 0 = the library default       */
  int fp64;               /* 0 = mixed precision, 1 = FP64                 */
  size_t nCharge;
  const double *coord;    /* [nCharge][3]                                  */
  const double *charge;   /* [nCharge] real or [nCharge][2] complex        */
  double *V;              /* [nCharge][4] or [nCharge][4][2], output       */
  double tArrival;        /* arrival time on the server clock, or < 0 for
                             the time of submission                       */
} npme_request;
typedef struct {
  int status;                              /* NPME_OK or an error code     */
  double tArrival, tSubmit, tStart, tEnd;  /* server clock, seconds        */
  long N[3];                               /* FFT grid used                */
  double deviceGB;                         /* device footprint             */
  /* wall time of the request's phases on its slot, milliseconds: grid
   * choice (RSEM keywords), set-up, charges to the device, CalcV, V to the
   * host */
  double msKeywords, msSetup, msCopyIn, msCalc, msCopyOut;
  /* device memory reserved for the request once its grid was chosen (its
   * estimated footprint), GB; deviceGB is what its set-up holds */
  double reservedGB;
  /* number of requests whose CalcVs ran in the same graph launch as this
   * one (1: it ran alone) */
  int batch;
} npme_request_info;
int npme_server_create(npme_server **s, int nSlots, double memGB);
/* queues a request; returns at once with its ticket */
int npme_server_submit(npme_server *s, const npme_request *r, long *ticket);
/* blocks until the request is done; returns its status (each ticket is
 * waited for once) */
int npme_server_wait(npme_server *s, long ticket, npme_request_info *info);
/* seconds since the server was created */
double npme_server_clock(const npme_server *s);
/* small requests queued together are batched on one slot: up to maxBatch
 * requests of at most maxCharges charges each run their set-ups in turn on
 * handles of the slot, then their CalcVs as one captured HIP graph (one
 * launch; the CalcVs run concurrently on the device), then their copies
 * out.  Each result is the same as alone.  Defaults 8 and 32768; maxBatch 1
 * turns batching off. */
int npme_server_set_batching(npme_server *s, int maxBatch, size_t maxCharges);
/* finishes the queued requests, then stops the slots */
void npme_server_destroy(npme_server *s);
/* page-locked host memory for request buffers (coord, charge, V): the
 * copies to and from the device then run at DMA speed.  No counterpart in
 * the reference.  npme_host_free may wait for the device: free at shutdown,
 * not per request. */
int npme_host_alloc(size_t bytes, void **p);
void npme_host_free(void *p);

/* ---- on-disk formats of the command-line tools (src/ReadPrint.cpp) and
 * the compareV metric.  Arrays returned through double ** are allocated by
 * the library and released with npme_free. ------------------------------ */
/* NPME_ReadSingleBoxChargeCoord (src/ReadPrint.cpp:282-440): coord [n][3],
 * charge [n] or [n][2] (interleaved re, im) */
int npme_read_coord_charge(const char *path, size_t *nCharge, int *isChargeReal,
                           double **coord, double **charge);
/* NPME_PrintSingleBoxChargeCoord (src/ReadPrint.cpp:234-280) */
int npme_write_coord_charge(const char *path, size_t nCharge, int isChargeReal,
                            const double *coord, const double *charge);
/* NPME_PrintSingleBox_V (src/ReadPrint.cpp:487-560) */
int npme_write_v(const char *path, size_t nCharge, int isChargeReal, const double *V,
                 int highPrecision);
/* NPME_ReadSingleBox_V (src/ReadPrint.cpp:565-739): V [n][4] or [n][4][2] */
int npme_read_v(const char *path, size_t *nCharge, int *isChargeReal, double **V);
void npme_free(void *p);
/* NPME_CompareV1Arrays + NPME_CalcV1AvgVecMag (src/SupportFunctions.cpp:
 * 88-305): out = {errorV, errordVdr, V_mag(Vref), dVdr_mag(Vref)}; the
 * relative errors are out[0]/out[2] and out[1]/out[3] */
int npme_compare_v(size_t nCharge, int isChargeReal, const double *V, const double *Vref,
                   double out[4]);

/* ---- testing hook: the calling thread's k-th next device allocation of
 * the library (k >= 0, counted from 0) fails as if the device were out of
 * memory; k < 0 disables it.  No reference counterpart: it exercises the
 * out-of-memory paths of set-up and CalcV. ----------------------------- */
void npme_debug_fail_alloc(long k);

const char *npme_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
