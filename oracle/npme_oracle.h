/*
 * npme_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 CPU restatement of the NPME potential evaluation (CalcV)
 * of ElkingD/npme, used as the checker for the HIP library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only to check or to time a CPU baseline; the product never links it.
 *
 * Every function names the reference file:line it restates.  Pinning (what
 * checks each function against something other than itself) is listed per
 * function; a function marked "parity unpinned" has no such check.
 *
 * Complex numbers are passed as interleaved (re, im) doubles, which is the
 * memory layout of C99 `_Complex double` used by the reference.
 */
#ifndef NPME_ORACLE_H
#define NPME_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kernel kinds */
#define NPO_LAPLACE   0   /* f = 1/r                         */
#define NPO_RALPHA    1   /* f = r^alpha                     */
#define NPO_HELMHOLTZ 2   /* f = exp(i k0 r)/r (complex)     */
#define NPO_LAPLACE_ERF 3 /* f = 1/r with the Ewald split erf(beta r)/r + erfc(beta r)/r */
/* f = c r with the split c r erf(beta r) + c r erfc(beta r): the user-defined
 * kernel of app/userDefineFunc.cpp:17-20,180-215 (rFunc), c = k0r + i k0i;
 * NPO_RFUNC is real (c = k0r), NPO_RFUNC_CX complex */
#define NPO_RFUNC    4
#define NPO_RFUNC_CX 5
#define NPO_IS_COMPLEX(kind) ((kind) == NPO_HELMHOLTZ || (kind) == NPO_RFUNC_CX)

/* kernel parts */
#define NPO_FULL 0
#define NPO_LR   1
#define NPO_SR   2

typedef struct {
  int    kind;       /* NPO_LAPLACE / NPO_RALPHA / NPO_HELMHOLTZ      */
  double alpha;      /* r^alpha exponent (Laplace uses -1)            */
  double k0r, k0i;   /* Helmholtz wave number                         */
  int    Nder;       /* derivative-match order                        */
  double Rdir;       /* direct-space cutoff                           */
  double a[2*31];    /* DM even-series coefficients (complex: re,im)  */
  double b[2*31];    /* (1/r) d/dr coefficients                       */
  double beta;       /* Ewald split parameter (NPO_LAPLACE_ERF)       */
} npo_kernel;

/* ---- B-splines (src/Bspline.cpp:16714, :16766) ------------------------ */
double npo_bspline(int n, double w);
double npo_bspline_dw(int n, double w, double *dBdw);

/* ---- DM split (src/FunctionDerivMatch.cpp:42-91, :226-298, :364-450) --- */
int  npo_kernel_init(npo_kernel *k, int kind, double alpha, double k0r,
                     double k0i, int Nder, double Rdir);
/* Ewald (erf/erfc) split of 1/r, EwaldSplit = LaplaceOrig
 * (src/KernelFunctionLaplace.cpp:41-72, :1087-1116, :1344-1373):
 * LR = erf(beta r)/r, SR = erfc(beta r)/r, beta from Rdir and tol. */
double npo_rdir2beta(double Rdir, double tol);
int  npo_kernel_init_erf(npo_kernel *k, double Rdir, double beta);
/* the rFunc user kernel (NPO_RFUNC / NPO_RFUNC_CX) with coefficient c and
 * split parameter beta; the short-range part is cut at r > Rdir */
int  npo_kernel_init_rfunc(npo_kernel *k, int isComplex, double cr, double ci,
                           double Rdir, double beta);
/* evaluate one part of the kernel at r2: f0 and f1 = (1/r) df/dr.
 * Real kernels write f0[0], f1[0]; Helmholtz writes (re, im) pairs. */
void npo_kernel_eval(const npo_kernel *k, int part, double r2,
                     double *f0, double *f1);

/* ---- FFT size / block size picker (src/SupportFunctions.cpp:384-455) -- */
long npo_optimal_grid_size(long n);
void npo_fft_size_block(long Nideal, long nideal, long *N, long *nblk);

/* ---- geometry (src/RecSumSupportFunctions.cpp:137-200, :276-296) ------ */
void npo_box_dim_center(size_t n, const double *coord, double X0[3],
                        double R0[3]);
void npo_pme_corr_box(const double X0[3], const double R0[3], const long N[3],
                      const double del[3], long Bn, double X[3], double R[3]);

/* ---- RSEM grid chooser (src/EstimateParm.cpp:62-181, :209-558,
 *      :2571-2888) ---------------------------------------------------- */
/* real charges, f = r^alpha; nDerivIn < 0 chooses Nder (smallest grid).
 * returns 0 ok, -1 no table row. */
int npo_rsem_static(size_t n, const double *q, const double *coord, long Bn,
                    double alpha, double tol, double Rdir, int nDerivIn,
                    long N[3], long nblk[3], int *NderOut);
int npo_rsem_static_pred_error(size_t n, const double *q, const double *coord,
                               long Bn, double alpha, int Nder, double Rdir,
                               const long N[3], double *eps, int *reliable);
int npo_rsem_helmholtz(size_t n, const double *qc, const double *coord,
                       long Bn, double k0r, double k0i, double tol,
                       double Rdir, int nDerivIn, long N[3], long nblk[3],
                       int *NderOut);

/* ---- cell partition (src/PartitionEmbeddedBox.cpp:266-455,
 *      src/PartitionBox.cpp:146-194, :311-351) -------------------------- */
/* writes per-point embedded cell indices; returns 0 ok, -1 out-of-box */
int npo_cell_index(size_t n, const double *coord, double Rdir, long nNeigh,
                   long nCellClust1D, long nCellA[3], unsigned long *totA,
                   unsigned long *totB);

/* first grid index m of every charge's spline window per axis,
 * mStart[n][3] (src/RecSumQ.cpp:106-108, src/RecSumSupportFunctions.h:
 * 123-149); returns -1 if a window leaves [-N/4, N/4-1] */
int npo_spline_start(size_t n, const double *coord, const long N[3],
                     const double L[3], long Bn, const double R[3], long *mStart);

/* ---- FFT (MKL DFTI boundary, src/ExtLibrary.cpp:930-1033) ------------ */
/* in-place complex FFT, no normalisation, sign = -1 forward, +1 backward */
void npo_fft1d(long N, double *z, int sign);
void npo_fft3d(long N1, long N2, long N3, double *z, int sign);

/* ---- reciprocal sum pieces ------------------------------------------- */
/* compact Fourier grid of f_LR (src/RecSumGrid.cpp:45-162, :388-651):
 * F[(N1/2+1)(N2/2+1)(N3/2+1)] complex */
int npo_fgrid_compact(const npo_kernel *k, const long N[3], const double a[3],
                      const double del[3], const double X[3], double *F);
/* |lambda|^2 (src/RecSumV.cpp:565-581) */
void npo_lambda2(long N, long Bn, double *lam2);
/* brute-force spreading (src/RecSumQ.cpp:93-168); complex charges are
 * spread as conj(q).  isComplex selects the charge type of q. */
int npo_spread_basic(size_t n, const double *q, int isComplex,
                     const double *coord, const long N[3], const double L[3],
                     long Bn, const double R[3], double *Q);
/* 8-parity compact convolution (src/RecSumV.cpp:639-840,
 * src/RecSumQ.cpp:456-522) */
void npo_theta_8parity(const long N[3], long Bn, const double *F,
                       const double *Q, double *T);
/* 8-parity compact convolution with the full Fourier grid F[N1][N2][N3]
 * and the optional sign reversal per axis (src/RecSumV.cpp:851-1009):
 * rev[d] != 0 uses F as the grid of R0 with component d negated */
void npo_theta_fullF_reverse(const long N[3], long Bn, const double *F,
                             const double *Q, const int rev[3], double *T);
/* translated kernel grid f(r) = part(|r + R0|) on the full grid
 * (src/RecSumGrid.cpp:663-794), real space, f[N1][N2][N3] complex;
 * 0 ok, -1 sample outside the extension, -2 R0 too close (rMin < 1e-6) */
int npo_fgrid_fulltrans(const npo_kernel *k, int part, const long N[3],
                        const double a[3], const double del[3], const double X[3],
                        const double R0[3], double *f);
/* min |r + R0| over the box of edges l centred at 0
 * (src/SupportFunctions.cpp:687-718) */
double npo_min_distance_r0(const double R0[3], const double l[3]);
/* brute-force interpolation (src/RecSumV.cpp:44-149): V1[n][4] complex */
int npo_interp_basic(size_t n, const double *coord, const double *theta,
                     const long N[3], const double L[3], long Bn,
                     const double R[3], double *V1);

/* ---- direct and exact sums ------------------------------------------- */
/* short-range DM direct sum, all ordered pairs with r inside the cutoff
 * (src/PotentialLaplace.cpp:201-264, :488-612; PotentialGenFunc.cpp:140-320;
 * PotentialHelmholtz.cpp:194-483).  V is [n][4] real or complex. */
void npo_direct_sr(const npo_kernel *k, size_t n, const double *coord,
                   const double *q, double *V);
/* direct sum over the reference's interacting cell pairs, every point pair
 * of them, no cutoff, for the short-range kernels without a cutoff of their
 * own: NPO_LAPLACE_ERF (src/PotentialLaplace.cpp:320-380) and the rFunc
 * user kernels (src/PotentialGenFunc.cpp:140-320), with the list of
 * src/PartitionBox.cpp:789-922; V[n][4] real or complex; -1 if a point
 * leaves the cell grid, -2 for another kernel kind */
int npo_direct_refpairs(const npo_kernel *k, size_t n, const double *coord,
                        const double *q, long nNeigh, long nCellClust1D,
                        double *V);
/* exact O(N^2) sum of the full kernel for the targets idx[0..nt)
 * (src/PotentialLaplace.cpp:40-132, PotentialHelmholtz.cpp:41-135,
 * PotentialGenFunc.cpp:321-522).  idx NULL means all targets. */
void npo_exact(const npo_kernel *k, size_t n, const double *coord,
               const double *q, size_t nt, const long *idx, double *V);
/* number of ordered pairs i != j with r_ij < Rdir (for FLOP accounting) */
double npo_count_pairs(size_t n, const double *coord, double Rdir);

/* ---- whole path ------------------------------------------------------ */
typedef struct {
  long   N[3], nblk[3];
  long   Bn;
  double del[3], a[3];
} npo_recsum_parm;

/* V[n][4] = NPME CalcV (src/Interface.cpp:2017-2100 Laplace,
 * :1719-1800 GenFunc, :2853-2938 Helmholtz) */
int npo_calcv(const npo_kernel *k, const npo_recsum_parm *p, size_t n,
              const double *q, const double *coord, double *V);

/* wall times of the last npo_calcv: [0] direct sum + self term,
 * [1] kernel-grid set-up (F), [2] spreading + convolution + interpolation */
void npo_last_calcv_times(double t[3]);

/* compareV metric (src/SupportFunctions.cpp:88-305):
 * out[0] = errorV, out[1] = errordVdr, out[2] = mean|V_ref|,
 * out[3] = mean over p=0..2 of |V_ref[4i+p]| (reference quirk),
 * out[4] = max_i |dV_i| / |V_ref_i| */
void npo_compare_v(size_t n, int isComplex, const double *Vref,
                   const double *V, double out[5]);

/* SplitMix64 stream as doubles in [0, 1) */
void npo_splitmix_uniform(unsigned long long seed, size_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif
