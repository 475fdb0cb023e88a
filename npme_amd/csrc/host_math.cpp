// Host-side FP64 mathematics of the NPME set-up: B-splines, the
// derivative-match fit, the FFT size picker, box geometry and the RSEM grid
// chooser.  Each routine cites the reference code whose behaviour it keeps.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "npme_internal.h"
#include "rsem_table.inc"

namespace npme {

using cd = std::complex<double>;

// ---------------------------------------------------------------- B-splines
// Cardinal B-spline M_n on [0, n] (src/Bspline.cpp:16714), evaluated for all
// integer shifts of a fractional argument with the Cox-de Boor recursion.
void bspline_shifts(int n, double t, double *out)
{
  double m[64], nm[64];
  if (n == 1) {
    out[0] = 1.0;
    return;
  }
  m[0] = t;
  m[1] = 1.0 - t;
  for (int k = 3; k <= n; k++) {
    for (int j = 0; j < k; j++) {
      double left = (j < k - 1) ? m[j] : 0.0;
      double right = (j >= 1) ? m[j - 1] : 0.0;
      nm[j] = ((t + j) * left + (k - t - j) * right) / (k - 1);
    }
    std::memcpy(m, nm, sizeof(double) * k);
  }
  std::memcpy(out, m, sizeof(double) * n);
}

double bspline(int n, double w)
{
  if (w < 0.0 || w > (double)n) return 0.0;
  if (n == 1) return (w < 1.0) ? 1.0 : 0.0;
  long j0 = (long)std::floor(w);
  if (j0 >= n) return 0.0;
  double sh[64];
  bspline_shifts(n, w - (double)j0, sh);
  return sh[j0];
}

// ---------------------------------------------------------------- DM split
// Radial derivatives at Rdir (src/FunctionDerivMatch.cpp:42-55 Helmholtz,
// :78-91 r^alpha) and the upper-triangular fit of the even series
// (src/FunctionDerivMatch.cpp:226-266); b_i = 2(i+1) a_{i+1}.
static double factorial(int n)
{
  double f = 1;
  for (int i = 2; i <= n; i++) f *= i;
  return f;
}

bool dm_kernel_init(DMKernel &k, int kind, double alpha, double k0r, double k0i, int Nder,
                    double Rdir, std::string &err)
{
  if (Nder < 2 || Nder > 30) {
    err = "Nder = " + std::to_string(Nder) + " outside [2, 30]";
    return false;
  }
  if (!(Rdir > 0)) {
    err = "Rdir must be positive";
    return false;
  }
  k = DMKernel();
  k.kind = kind;
  k.alpha = (kind == KIND_LAPLACE) ? -1.0 : alpha;
  k.k0r = k0r;
  k.k0i = k0i;
  k.Nder = Nder;
  k.Rdir = Rdir;
  cd f[32], a[32];
  const double r2 = Rdir * Rdir;
  if (kind == KIND_HELMHOLTZ) {
    const cd k0(k0r, k0i), I(0, 1);
    const cd C = std::exp(I * k0 * Rdir);
    f[0] = C / Rdir;
    f[1] = -C / (Rdir * r2) + I * k0 * C / r2;
    for (int n = 2; n <= Nder; n++) f[n] = -(2.0 * n - 1.0) / r2 * f[n - 1] - k0 * k0 / r2 * f[n - 2];
  } else {
    f[0] = std::pow(Rdir, k.alpha);
    for (int n = 1; n <= Nder; n++) f[n] = f[n - 1] * (k.alpha - 2 * n + 2) / r2;
  }
  // row n of the system: sum_{p>=n} a_p 2^n p!/(p-n)! R^{2(p-n)} = f_n
  for (int n = Nder; n >= 0; n--) {
    cd s = f[n];
    for (int p = n + 1; p <= Nder; p++)
      s -= std::pow(r2, p - n) * std::pow(2.0, n) * factorial(p) / factorial(p - n) * a[p];
    a[n] = s / (std::pow(2.0, n) * factorial(n));
  }
  for (int i = 0; i <= Nder; i++) {
    k.a[2 * i] = a[i].real();
    k.a[2 * i + 1] = a[i].imag();
    cd bi = (i < Nder) ? 2.0 * (i + 1) * a[i + 1] : cd(0);
    k.b[2 * i] = bi.real();
    k.b[2 * i + 1] = bi.imag();
  }
  return true;
}

double rdir2beta(double Rdir, double tol)
// x = beta Rdir solves x^2 + ln x + ln(sqrt(pi) tol Rdir) = 0, the
// asymptotic form of erfc(x)/Rdir = tol; Newton from x = 0.1 to |F| < 1e-12
// (src/KernelFunctionLaplace.cpp:41-72)
{
  const double C = std::log(std::sqrt(M_PI) * tol * Rdir);
  double x = 0.1;
  for (int it = 0; it < 1000; it++) {
    const double F0 = x * x + std::log(x) + C, F1 = 2 * x + 1.0 / x;
    if (std::fabs(F0) < 1e-12) return x / Rdir;
    x -= F0 / F1;
  }
  throw Error{NPME_ERR_KEYWORD, "EwaldSplit LaplaceOrig: no beta for Rdir and tol"};
}

void erf_kernel_init(DMKernel &k, int Nder, double Rdir, double beta)
{
  k = DMKernel();
  k.kind = KIND_LAPLACE_ERF;
  k.alpha = -1.0;
  k.Nder = Nder;
  k.Rdir = Rdir;
  k.beta = beta;
}

void dm_eval_lr(const DMKernel &k, double r2, double f0[2], double f1[2])
// f_LR: even series inside Rdir, the full kernel outside
// (src/KernelFunctionLaplace.cpp:394-437, KernelFunctionRalpha.cpp:343-385,
// KernelFunctionHelmholtz.cpp:506-560)
{
  const double r = std::sqrt(std::fabs(r2));
  if (k.kind == KIND_LAPLACE_ERF) {
    // erf(beta r)/r and (1/r) d/dr of it, by series below x = beta r = 1e-4
    // (src/KernelFunctionLaplace.cpp:1087-1116)
    const double x = k.beta * r, c = 2.0 / std::sqrt(M_PI);
    if (x < 1e-4) {
      f0[0] = k.beta * c * (1.0 - x * x / 3.0 + x * x * x * x / 10.0);
      f1[0] = k.beta * k.beta * k.beta * c * (-2.0 / 3.0 + 0.4 * x * x);
    } else {
      f0[0] = std::erf(x) / r;
      f1[0] = (k.beta * c * std::exp(-x * x) - f0[0]) / r2;
    }
    f0[1] = f1[1] = 0;
    return;
  }
  bool inside = (k.kind == KIND_RALPHA) ? !(r2 > k.Rdir * k.Rdir) : !(r > k.Rdir);
  cd v, d;
  if (inside) {
    v = cd(k.a[2 * k.Nder], k.a[2 * k.Nder + 1]);
    for (int i = k.Nder - 1; i >= 0; i--) v = v * r2 + cd(k.a[2 * i], k.a[2 * i + 1]);
    d = cd(k.b[2 * (k.Nder - 1)], k.b[2 * (k.Nder - 1) + 1]);
    for (int i = k.Nder - 2; i >= 0; i--) d = d * r2 + cd(k.b[2 * i], k.b[2 * i + 1]);
  } else if (k.kind == KIND_LAPLACE) {
    v = 1.0 / r;
    d = -1.0 / (r * r2);
  } else if (k.kind == KIND_RALPHA) {
    v = std::pow(r2, k.alpha / 2);
    d = k.alpha * v / r2;
  } else {
    const cd k0(k.k0r, k.k0i), I(0, 1);
    v = std::exp(I * k0 * r) / r;
    d = (I * k0 * v - v / r) / r;
  }
  f0[0] = v.real();
  f0[1] = v.imag();
  f1[0] = d.real();
  f1[1] = d.imag();
}

// ---------------------------------------------------------------- sizes
// smallest multiple of 4 >= n with prime factors in {2,3,5,7}; block size
// closest to the ideal with N % 2m == 0, m <= N/4, search seeded with
// minDiff = N (src/SupportFunctions.cpp:384-455)
static bool smooth2357(long n)
{
  for (long p : {2L, 3L, 5L, 7L})
    while (n % p == 0) n /= p;
  return n == 1;
}

long optimal_grid_size(long n)
{
  for (long N = 4; N <= 2 * n + 4; N += 4)
    if (N >= n && smooth2357(N)) return N;
  return -1;
}

void fft_size_block(long Nideal, long nideal, long &N, long &nblk)
{
  N = optimal_grid_size(Nideal);
  long minDiff = N, best = 1;
  for (long m = 2; m < N; m++)
    if (N % (2 * m) == 0 && m <= N / 4) {
      long d = std::labs(m - nideal);
      if (minDiff > d) {
        minDiff = d;
        best = m;
      }
    }
  nblk = best;
}

// ---------------------------------------------------------------- geometry
// src/RecSumSupportFunctions.cpp:149-200 and :137-147, :276-296
void box_dim_center(size_t n, const double *coord, double X0[3], double R0[3])
{
  // two charges (six values) per step: the six running extremes are
  // independent chains.  Minima and maxima are exact, so the grouping does
  // not change them.
  double mx6[6], mn6[6];
  for (int k = 0; k < 6; k++) {
    mx6[k] = -1e12;
    mn6[k] = 1e12;
  }
  size_t i = 0;
  for (; i + 2 <= n; i += 2)
    for (int k = 0; k < 6; k++) {
      const double x = coord[3 * i + k];
      mx6[k] = std::max(mx6[k], x);
      mn6[k] = std::min(mn6[k], x);
    }
  for (; i < n; i++)
    for (int d = 0; d < 3; d++) {
      mx6[d] = std::max(mx6[d], coord[3 * i + d]);
      mn6[d] = std::min(mn6[d], coord[3 * i + d]);
    }
  for (int d = 0; d < 3; d++) {
    const double mx = std::max(mx6[d], mx6[d + 3]), mn = std::min(mn6[d], mn6[d + 3]);
    X0[d] = mx - mn;
    R0[d] = 0.5 * (mx + mn);
  }
}

void box_from_bounds(const double mn[3], const double mx[3], double X0[3], double R0[3])
{
  for (int d = 0; d < 3; d++) {
    const double hi = std::max(-1.0e12, mx[d]), lo = std::min(1.0e12, mn[d]);
    X0[d] = hi - lo;
    R0[d] = 0.5 * (hi + lo);
  }
}

// sum of squares of m values in four interleaved partial sums, so the
// loop is not bound by the latency of a single chain of additions
double sum_squares(size_t m, const double *v)
{
  double s[4] = {0, 0, 0, 0};
  size_t i = 0;
  for (; i + 4 <= m; i += 4)
    for (int k = 0; k < 4; k++) s[k] += v[i + k] * v[i + k];
  for (; i < m; i++) s[0] += v[i] * v[i];
  return (s[0] + s[1]) + (s[2] + s[3]);
}

void pme_corr_box(const double X0[3], const double R0[3], const long N[3], const double del[3],
                  long Bn, double X[3], double R[3])
{
  for (int d = 0; d < 3; d++) {
    const double a1 = 1.0 - 2.0 * (Bn + 1) / N[d];
    const double a2 = 1.0 - 4.0 * (Bn + 1) / N[d];
    double x = X0[d] / a1 - del[d] / 2.0 * a2 / a1;
    if (X0[d] > x) x = X0[d];
    X[d] = x;
    R[d] = R0[d] + 0.5 * (X0[d] - x);
  }
}

// ---------------------------------------------------------------- RSEM
// Error model eps = Q_d Rdir^alpha (Rdir/L_d)^1.5 exp(P(u_d)), u = N Rdir/L,
// 10 fixed-point iterations from N = 100 (src/EstimateParm.cpp:62-102,
// :110-181, :209-417, :421-558; Helmholtz :2571-2888).
static void find_u(double &u, double y0, const double a[5], double uMin, double uMax)
{
  const int NumTrial = 1000;
  const double del = (uMax - uMin) / (NumTrial - 1);
  double minDiff = 1.0e6, yPrev = 1.0e6;
  u = uMin;
  for (int n = 0; n < NumTrial; n++) {
    double ut = uMin + n * del, y = 0, up = 1;
    for (int p = 0; p <= 4; p++) {
      y += a[p] * up;
      up *= ut;
    }
    double diff = std::fabs(y - y0);
    if (minDiff > diff) {
      minDiff = diff;
      u = ut;
    }
    if (minDiff < 0.1) break;
    if (y > yPrev) break;
    yPrev = y;
  }
}

// X0, R0: box size and centre of the charges (box_dim_center)
static void rsem_iterate(const double X0[3], const double R0[3], double Q0, long Bn,
                         double alphaExp, double tol, double Rdir, const double a[5], double uMin,
                         double uMax, long N[3], long nblk[3])
{
  const double del = Rdir;
  N[0] = N[1] = N[2] = 100;
  for (int it = 0; it < 10; it++) {
    double X[3], R[3], dl[3] = {del, del, del}, L[3];
    if (it == 0)
      for (int d = 0; d < 3; d++) X[d] = X0[d];
    else
      pme_corr_box(X0, R0, N, dl, Bn, X, R);
    for (int d = 0; d < 3; d++) L[d] = 2 * (X[d] + del);
    const double V = L[0] * L[1] * L[2];
    const double C = tol * std::pow(Rdir, -alphaExp);
    for (int d = 0; d < 3; d++) {
      double Qd = Q0 * std::sqrt(L[d] * L[d] * L[d] / V);
      double t = C / Qd * std::pow(L[d] / Rdir, 1.5);
      double u;
      find_u(u, std::log(t), a, uMin, uMax);
      N[d] = (long)(u * L[d] / Rdir) + 1;
      if (N[d] < 4 * Bn + 1) N[d] = 4 * Bn + 1;
    }
  }
  for (int d = 0; d < 3; d++) fft_size_block(N[d], 20, N[d], nblk[d]);
}

void charge_stats(size_t n, const double *coord, const double *q, bool complexQ, ChargeStats &st)
{
  box_dim_center(n, coord, st.X0, st.R0);
  st.Q0 = std::sqrt(sum_squares((complexQ ? 2 : 1) * n, q));
}

bool rsem_static(const ChargeStats &st, long Bn, double alpha, double tol, double Rdir,
                 int nDerivIn, long N[3], long nblk[3], int &Nder)
{
  const double Q0 = st.Q0;
  const double *X0 = st.X0, *R0 = st.R0;
  bool found = false;
  long best = 0;
  for (int r = 0; r < NPME_RSEM_N_STATIC; r++) {
    const double *row = npme_rsem_static[r];
    if ((long)row[0] != Bn || std::fabs(alpha - row[2]) >= 1e-6) continue;
    if (nDerivIn >= 0 && (int)row[1] != nDerivIn) continue;
    long Nt[3], nt[3];
    rsem_iterate(X0, R0, Q0, Bn, alpha, tol, Rdir, row + 5, row[3], row[4], Nt, nt);
    long sz = Nt[0] * Nt[1] * Nt[2];
    if (!found || best > sz) {
      best = sz;
      for (int d = 0; d < 3; d++) {
        N[d] = Nt[d];
        nblk[d] = nt[d];
      }
      Nder = (int)row[1];
      found = true;
    }
    if (nDerivIn >= 0) break;
  }
  return found;
}

bool rsem_static_pred_error(size_t n, const double *q, const double *coord, long Bn, double alpha,
                            int Nder, double Rdir, const long N[3], double &eps, bool &reliable)
{
  const double *row = nullptr;
  for (int r = 0; r < NPME_RSEM_N_STATIC; r++) {
    const double *t = npme_rsem_static[r];
    if ((long)t[0] == Bn && (int)t[1] == Nder && std::fabs(alpha - t[2]) < 1e-6) {
      row = t;
      break;
    }
  }
  if (!row) return false;
  const double Q0 = std::sqrt(sum_squares(n, q));
  double X0[3], R0[3], X[3], R[3], dl[3] = {Rdir, Rdir, Rdir}, L[3];
  box_dim_center(n, coord, X0, R0);
  pme_corr_box(X0, R0, N, dl, Bn, X, R);
  for (int d = 0; d < 3; d++) L[d] = 2 * (X[d] + Rdir);
  const double V = L[0] * L[1] * L[2];
  reliable = true;
  eps = 0;
  for (int d = 0; d < 3; d++) {
    double u = N[d] * Rdir / L[d];
    if (u < row[3] || u > row[4]) reliable = false;
    double P = 0, up = 1;
    for (int p = 0; p <= 4; p++) {
      P += row[5 + p] * up;
      up *= u;
    }
    double Qd = Q0 * std::sqrt(L[d] * L[d] * L[d] / V);
    eps = std::max(eps, Qd * std::pow(Rdir, alpha) * std::pow(Rdir / L[d], 1.5) * std::exp(P));
  }
  return true;
}

bool rsem_helmholtz(const ChargeStats &st, long Bn, double k0r, double k0i, double tol,
                    double Rdir, int nDerivIn, long N[3], long nblk[3], int &Nder,
                    std::string &err)
{
  const double Q0 = st.Q0;
  const double *X0 = st.X0, *R0 = st.R0;
  const double w = std::hypot(k0r, k0i) * Rdir;
  bool found = false;
  long best = 0;
  for (int r = 0; r < NPME_RSEM_N_HELM; r++) {
    const double *row = npme_rsem_helm[r];
    if ((long)row[0] != Bn) continue;
    if (nDerivIn >= 0 && (int)row[1] != nDerivIn) continue;
    if (w < row[4] || w > row[5]) {
      std::ostringstream os;
      os << "Helmholtz RSEM: w = |k0| Rdir = " << w << " outside [" << row[4] << ", " << row[5]
         << "]";
      err = os.str();
      return false;
    }
    double b[5], wq[5];
    wq[0] = 1;
    for (int q = 1; q < 5; q++) wq[q] = wq[q - 1] * w;
    for (int p = 0; p < 5; p++) {
      b[p] = 0;
      for (int q = 0; q < 5; q++) b[p] += row[6 + p * 5 + q] * wq[q];
    }
    long Nt[3], nt[3];
    rsem_iterate(X0, R0, Q0, Bn, -1.0, tol, Rdir, b, row[2], row[3], Nt, nt);
    long sz = Nt[0] * Nt[1] * Nt[2];
    if (!found || best > sz) {
      best = sz;
      for (int d = 0; d < 3; d++) {
        N[d] = Nt[d];
        nblk[d] = nt[d];
      }
      Nder = (int)row[1];
      found = true;
    }
    if (nDerivIn >= 0) break;
  }
  if (!found) err = "Helmholtz RSEM: no parameters for BnOrder = " + std::to_string(Bn);
  return found;
}

// |lambda(k)|^2 (src/RecSumV.cpp:565-581)
void lambda2(long N, long Bn, double *lam2)
{
  const double PI = 3.14159265358979323846;
  double Mp[64];
  for (int p = 0; p <= (int)Bn; p++) Mp[p] = bspline((int)Bn, (double)p);
  for (long idx = 0; idx < N; idx++) {
    long kk = (idx >= N / 2) ? idx - N : idx;
    double z = (double)kk / (double)N;
    cd den = 0;
    for (int p = 0; p <= (int)Bn; p++) den += Mp[p] * std::exp(cd(0, -2.0 * PI * z * p));
    lam2[idx] = 1.0 / std::norm(den);
  }
}

// ---------------------------------------------------------------- keywords
bool Keywords::get(const std::string &k, std::string &v) const
{
  for (auto &p : kv)
    if (p.first == k) {
      v = p.second;
      return true;
    }
  return false;
}
bool Keywords::get(const std::string &k, double &v) const
{
  std::string s;
  if (!get(k, s)) return false;
  char *end;
  v = std::strtod(s.c_str(), &end);
  return end != s.c_str();
}
bool Keywords::get(const std::string &k, long &v) const
{
  double d;
  if (!get(k, d)) return false;
  v = (long)d;
  return true;
}
bool Keywords::get(const std::string &k, int &v) const
{
  double d;
  if (!get(k, d)) return false;
  v = (int)d;
  return true;
}
void Keywords::set(const std::string &k, const std::string &v)
{
  for (auto &p : kv)
    if (p.first == k) {
      p.second = v;
      return;
    }
  kv.emplace_back(k, v);
}
void Keywords::set(const std::string &k, double v)
{
  std::ostringstream os;
  os.precision(17);
  os << v;
  set(k, os.str());
}
void Keywords::set(const std::string &k, long v) { set(k, std::to_string(v)); }

// ---------------------------------------------------------------- FFT plan
// radices 8, 4, 2, 7, 5, 3 (largest first)
FFTPlan1D make_plan(int N)
{
  FFTPlan1D p;
  p.N = N;
  int n = N;
  while (n % 8 == 0) {
    p.radix[p.nst++] = 8;
    n /= 8;
  }
  while (n % 4 == 0) {
    p.radix[p.nst++] = 4;
    n /= 4;
  }
  while (n % 2 == 0) {
    p.radix[p.nst++] = 2;
    n /= 2;
  }
  for (int r : {7, 5, 3})
    while (n % r == 0) {
      p.radix[p.nst++] = r;
      n /= r;
    }
  if (n != 1) p.nst = -1;  // unsupported prime factor
  return p;
}

}  // namespace npme
