// User-defined kernels: the host side of NPME_InterfaceReal_GenFunc /
// NPME_InterfaceComplex_GenFunc with NPME_KfuncReal / NPME_KfuncComplex
// plugins (src/Interface.cpp:1606-1800, src/Interface.h:563-650).
//
// * the long-range part f_LR is sampled on the kernel grid through the
//   user's callback, with the same images and weights as the built-in
//   kernels' device sampling (src/RecSumGrid.cpp:45-162 evaluates funcLR at
//   the same points);
// * fself = f_LR(0, 0, 0) (src/Interface.cpp:1656-1663);
// * the short-range part f_SR must be radial; it is tabulated as piecewise
//   polynomials in r on [0, Rdir] for the device direct sum, and the table's
//   error at check points between its nodes is kept;
// * the exact sum evaluates the full kernel through its callback on the host,
//   as NPME_PotGenFunc_MacroSelf_V1 does (src/PotentialGenFunc.cpp:321-522).
#include <algorithm>
#include <cmath>
#include <sstream>
#include <thread>
#include <vector>

#include "npme_internal.h"

namespace npme {
namespace {

// runs fn(task, thread) for task in [0, nTask) on up to 16 host threads
template <typename F> void parallel_for(long nTask, F &&fn)
{
  const long nth = std::max<long>(
      1, std::min<long>({(long)std::max(1u, std::thread::hardware_concurrency()), 16L, nTask}));
  if (nth == 1) {
    for (long t = 0; t < nTask; t++) fn(t, 0);
    return;
  }
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> err(nth);
  for (long w = 0; w < nth; w++)
    th.emplace_back([&, w] {
      try {
        for (long t = w; t < nTask; t += nth) fn(t, (int)w);
      } catch (...) {
        err[w] = std::current_exception();
      }
    });
  for (auto &t : th) t.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
}

// f at n points: fr, fi (fi = 0 for a real kernel)
void eval_value(const UserKernels &u, int which, size_t n, const double *x, const double *y,
                const double *z, double *fr, double *fi)
{
  std::copy(x, x + n, fr);
  if (u.complex) {
    const npme_kfunc_complex &k = u.c[which];
    k.value(k.ctx, n, fr, fi, y, z);
  } else {
    const npme_kfunc_real &k = u.r[which];
    k.value(k.ctx, n, fr, y, z);
    std::fill(fi, fi + n, 0.0);
  }
}

// f and grad f at n points; g[0..5] = (d/dx re, im, d/dy re, im, d/dz re, im)
void eval_grad(const UserKernels &u, int which, size_t n, const double *x, const double *y,
               const double *z, double *fr, double *fi, double *const g[6])
{
  std::copy(x, x + n, g[0]);
  std::copy(y, y + n, g[2]);
  std::copy(z, z + n, g[4]);
  if (u.complex) {
    const npme_kfunc_complex &k = u.c[which];
    k.grad(k.ctx, n, fr, fi, g[0], g[1], g[2], g[3], g[4], g[5]);
  } else {
    const npme_kfunc_real &k = u.r[which];
    k.grad(k.ctx, n, fr, g[0], g[2], g[4]);
    std::fill(fi, fi + n, 0.0);
    std::fill(g[1], g[1] + n, 0.0);
    std::fill(g[3], g[3] + n, 0.0);
    std::fill(g[5], g[5] + n, 0.0);
  }
}

struct Arr {
  std::vector<double> x, y, z, fr, fi, g[6];
  double *gp[6];
  explicit Arr(size_t n) : x(n, 0.0), y(n, 0.0), z(n, 0.0), fr(n), fi(n)
  {
    for (int c = 0; c < 6; c++) {
      g[c].resize(n);
      gp[c] = g[c].data();
    }
  }
};

// Chebyshev interpolant of degree D on the nodes t_i = cos(pi (i + 1/2) /
// (D + 1)), as monomial coefficients in t (lowest degree first)
void cheb_to_monomial(const double *val, double *mono)
{
  constexpr int D = kUserDeg, P = D + 1;
  double c[P];
  for (int k = 0; k < P; k++) {
    double s = 0;
    for (int i = 0; i < P; i++) s += val[i] * std::cos(M_PI * k * (i + 0.5) / P);
    c[k] = s * 2.0 / P;
  }
  c[0] *= 0.5;
  // monomial coefficients of T_k by T_{k+1} = 2 t T_k - T_{k-1}
  double Tm[P][P] = {{0}};
  Tm[0][0] = 1;
  Tm[1][1] = 1;
  for (int k = 1; k + 1 < P; k++)
    for (int j = 0; j < P; j++)
      Tm[k + 1][j] = (j > 0 ? 2 * Tm[k][j - 1] : 0.0) - Tm[k - 1][j];
  for (int j = 0; j < P; j++) {
    double s = 0;
    for (int k = 0; k < P; k++) s += c[k] * Tm[k][j];
    mono[j] = s;
  }
}

double horner(const double *c, double t)
{
  double s = c[kUserDeg];
  for (int j = kUserDeg - 1; j >= 0; j--) s = s * t + c[j];
  return s;
}

}  // namespace

void user_fself(const UserKernels &u, double fself[2])
{
  double x = 0, y = 0, z = 0, fr = 0, fi = 0;
  eval_value(u, 1, 1, &x, &y, &z, &fr, &fi);
  if (!std::isfinite(fr) || !std::isfinite(fi))
    throw Error{NPME_ERR_ARG, "funcLR(0, 0, 0) is not finite"};
  fself[0] = fr;
  fself[1] = fi;
}

// funcSR must be radial: its value at r along three directions equals its
// value on the x axis, and its gradient is f'(r) times the direction
void user_check_radial(const UserKernels &u, double Rdir)
{
  const double rs[4] = {0.11, 0.29, 0.53, 0.87};
  const double dir[3][3] = {{1, 2, 3}, {-2, 1, 0.5}, {0.3, -0.9, -1.7}};
  const size_t n = 16;
  Arr a(n);
  for (int i = 0; i < 4; i++) {
    a.x[4 * i] = rs[i] * Rdir;
    for (int d = 0; d < 3; d++) {
      const double nn = std::sqrt(dir[d][0] * dir[d][0] + dir[d][1] * dir[d][1] + dir[d][2] * dir[d][2]);
      a.x[4 * i + 1 + d] = rs[i] * Rdir * dir[d][0] / nn;
      a.y[4 * i + 1 + d] = rs[i] * Rdir * dir[d][1] / nn;
      a.z[4 * i + 1 + d] = rs[i] * Rdir * dir[d][2] / nn;
    }
  }
  const std::vector<double> x = a.x, y = a.y, z = a.z;
  eval_grad(u, 2, n, x.data(), y.data(), z.data(), a.fr.data(), a.fi.data(), a.gp);
  for (int i = 0; i < 4; i++) {
    const int b = 4 * i;
    const double f0r = a.fr[b], f0i = a.fi[b];
    const double dr = a.g[0][b], di = a.g[1][b];  // f'(r) on the x axis
    const double fs = std::hypot(f0r, f0i), ds = std::hypot(dr, di);
    for (int d = 1; d <= 3; d++) {
      const int p = b + d;
      const double r = rs[i] * Rdir;
      const double ux = x[p] / r, uy = y[p] / r, uz = z[p] / r;
      const double ef = std::hypot(a.fr[p] - f0r, a.fi[p] - f0i);
      double eg = 0;
      const double gx[2] = {dr * ux, di * ux}, gy[2] = {dr * uy, di * uy}, gz[2] = {dr * uz, di * uz};
      eg = std::max(eg, std::hypot(a.g[0][p] - gx[0], a.g[1][p] - gx[1]));
      eg = std::max(eg, std::hypot(a.g[2][p] - gy[0], a.g[3][p] - gy[1]));
      eg = std::max(eg, std::hypot(a.g[4][p] - gz[0], a.g[5][p] - gz[1]));
      if (!(ef <= 1e-9 * fs + 1e-300) || !(eg <= 1e-8 * ds + 1e-300)) {
        std::ostringstream os;
        os << "funcSR is not radial (at r = " << r << " its value or gradient depends on the "
           << "direction): the device direct sum tabulates f_SR(r)";
        throw Error{NPME_ERR_UNSUPPORTED, os.str()};
      }
    }
  }
}

// the short-range table: S pieces per octave, S doubled from 4 until the
// error at the check points is below 1e-13 (f) and 1e-12 (g) relative to the
// octave's largest |f| and |g|, or S = 512
void user_build_table(const UserKernels &u, double Rdir, std::vector<double> &coef, int &S,
                      double err[2])
{
  constexpr int K = kUserOct, P = kUserDeg + 1, NC = 17;
  const int nq = u.complex ? 4 : 2;  // functions per piece
  double tn[P], tc[NC];
  for (int i = 0; i < P; i++) tn[i] = std::cos(M_PI * (i + 0.5) / P);
  for (int m = 0; m < NC; m++) tc[m] = -1.0 + 2.0 * m / (NC - 1);
  for (S = 4;; S *= 2) {
    const size_t nPiece = (size_t)K * S, per = P + NC, n = nPiece * per;
    Arr a(n);
    for (int k = 0; k < K; k++)
      for (int j = 0; j < S; j++) {
        const double o = Rdir * std::ldexp(1.0, -(k + 1));
        const double lo = o * (1.0 + (double)j / S), hi = o * (1.0 + (double)(j + 1) / S);
        const double mid = 0.5 * (lo + hi), half = 0.5 * (hi - lo);
        const size_t b = ((size_t)k * S + j) * per;
        for (int i = 0; i < P; i++) a.x[b + i] = mid + half * tn[i];
        for (int m = 0; m < NC; m++) a.x[b + P + m] = mid + half * tc[m];
      }
    const std::vector<double> r = a.x;
    eval_grad(u, 2, n, r.data(), a.y.data(), a.z.data(), a.fr.data(), a.fi.data(), a.gp);
    // the four functions per point: Re f, Im f, Re g, Im g with g = f'(r) / r
    auto fun = [&](int q, size_t p) {
      switch (q) {
        case 0: return a.fr[p];
        case 1: return u.complex ? a.fi[p] : a.g[0][p] / r[p];
        case 2: return a.g[0][p] / r[p];
        default: return a.g[1][p] / r[p];
      }
    };
    // q index of the table order: real (f, g), complex (Re f, Im f, Re g, Im g)
    auto qmap = [&](int q) { return u.complex ? q : (q == 0 ? 0 : 2); };
    coef.assign(nPiece * nq * P, 0.0);
    double ef = 0, eg = 0;
    for (int k = 0; k < K; k++) {
      double sf = 0, sg = 0, df = 0, dg = 0;
      for (int j = 0; j < S; j++) {
        const size_t piece = (size_t)k * S + j, b = piece * per;
        double c[4][P];
        for (int q = 0; q < nq; q++) {
          double v[P];
          for (int i = 0; i < P; i++) {
            v[i] = fun(qmap(q), b + i);
            if (!std::isfinite(v[i])) {
              std::ostringstream os;
              os << "funcSR or its gradient is not finite at r = " << r[b + i];
              throw Error{NPME_ERR_UNSUPPORTED, os.str()};
            }
          }
          cheb_to_monomial(v, c[q]);
          std::copy(c[q], c[q] + P, coef.begin() + (piece * nq + q) * P);
        }
        for (int m = 0; m < NC; m++) {
          const size_t p = b + P + m;
          double vf[2] = {0, 0}, vg[2] = {0, 0}, tf[2] = {0, 0}, tg[2] = {0, 0};
          if (u.complex) {
            for (int c2 = 0; c2 < 2; c2++) {
              vf[c2] = fun(c2, p);
              vg[c2] = fun(2 + c2, p);
              tf[c2] = horner(c[c2], tc[m]);
              tg[c2] = horner(c[2 + c2], tc[m]);
            }
          } else {
            vf[0] = fun(0, p);
            vg[0] = fun(2, p);
            tf[0] = horner(c[0], tc[m]);
            tg[0] = horner(c[1], tc[m]);
          }
          sf = std::max(sf, std::hypot(vf[0], vf[1]));
          sg = std::max(sg, std::hypot(vg[0], vg[1]));
          df = std::max(df, std::hypot(tf[0] - vf[0], tf[1] - vf[1]));
          dg = std::max(dg, std::hypot(tg[0] - vg[0], tg[1] - vg[1]));
        }
      }
      if (!std::isfinite(df) || !std::isfinite(dg))
        throw Error{NPME_ERR_UNSUPPORTED, "funcSR is not finite on [0, Rdir]"};
      ef = std::max(ef, sf > 0 ? df / sf : df);
      eg = std::max(eg, sg > 0 ? dg / sg : dg);
    }
    err[0] = ef;
    err[1] = eg;
    if ((ef <= 1e-13 && eg <= 1e-12) || S >= 512) break;
  }
}

// f_0 of the user's func (which = 0) or funcLR (which = 1) on a grid given by
// per-axis tables [x1 | x2 | T1 | T2] of n[d] points: the same images and
// weights as the built-in kernels' device sampling; F is [n0][n1][n2]
// doubles (real) or (re, im) pairs (a complex kernel, or cxOut)
void user_fgrid_host_n(const UserKernels &u, int which, const long n[3], const std::vector<double> &tab,
                       bool cxOut, std::vector<double> &F)
{
  const double *t0 = tab.data(), *t1 = t0 + 4 * n[0], *t2 = t1 + 4 * n[1];
  const int cw = (u.complex || cxOut) ? 2 : 1;
  F.assign((size_t)n[0] * n[1] * n[2] * cw, 0.0);
  const long rows = n[0] * n[1];
  const size_t maxPts = 8 * (size_t)n[2];
  struct Buf {
    std::vector<double> x, y, z, fr, fi, w;
    std::vector<long> at;
  };
  const long nth = std::min<long>({(long)std::max(1u, std::thread::hardware_concurrency()), 16L, rows});
  std::vector<Buf> bufs(std::max<long>(nth, 1));
  for (Buf &b : bufs) {
    b.x.resize(maxPts);
    b.y.resize(maxPts);
    b.z.resize(maxPts);
    b.fr.resize(maxPts);
    b.fi.resize(maxPts);
    b.w.resize(maxPts);
    b.at.resize(maxPts);
  }
  parallel_for(rows, [&](long row, int th) {
    Buf &b = bufs[th];
    const long i1 = row / n[1], i2 = row % n[1];
    const double X[2] = {t0[i1], t0[n[0] + i1]}, TX[2] = {t0[2 * n[0] + i1], t0[3 * n[0] + i1]};
    const double Y[2] = {t1[i2], t1[n[1] + i2]}, TY[2] = {t1[2 * n[1] + i2], t1[3 * n[1] + i2]};
    size_t np = 0;
    for (long i3 = 0; i3 < n[2]; i3++) {
      const double Z[2] = {t2[i3], t2[n[2] + i3]}, TZ[2] = {t2[2 * n[2] + i3], t2[3 * n[2] + i3]};
      for (int ax = 0; ax < 2; ax++)
        for (int ay = 0; ay < 2; ay++)
          for (int az = 0; az < 2; az++) {
            // outside an axis' extension band its second weight is 0: the
            // term adds exactly zero and is not evaluated
            const double w = TX[ax] * TY[ay] * TZ[az];
            if (w == 0.0) continue;
            b.x[np] = X[ax];
            b.y[np] = Y[ay];
            b.z[np] = Z[az];
            b.w[np] = w;
            b.at[np] = i3;
            np++;
          }
    }
    if (np == 0) return;
    const std::vector<double> x(b.x.begin(), b.x.begin() + np);
    eval_value(u, which, np, x.data(), b.y.data(), b.z.data(), b.fr.data(), b.fi.data());
    double *Frow = F.data() + (size_t)row * n[2] * cw;
    for (size_t p = 0; p < np; p++) {
      Frow[b.at[p] * cw] += b.w[p] * b.fr[p];
      if (cw == 2) Frow[b.at[p] * 2 + 1] += b.w[p] * b.fi[p];
    }
  });
  for (double v : F)
    if (!std::isfinite(v))
      throw Error{NPME_ERR_UNSUPPORTED,
                  std::string(which == 0 ? "func" : "funcLR") + " is not finite on the kernel grid"};
}

// f_LR on the compact real-space kernel grid (per axis d the table holds x1,
// x2, T1, T2 of Mf[d] points each)
void user_fgrid_host(const UserKernels &u, const long Mf[3], const std::vector<double> &tab,
                     std::vector<double> &F)
{
  user_fgrid_host_n(u, 1, Mf, tab, false, F);
}

// V[t] = sum over j != i of f(r_i - r_j) q_j and its gradient for the
// targets i = idx[t] (all charges when idx is NULL), with the user's full
// kernel on the host; V is [nt][4] (complex [nt][4][2]); q real [n] or
// complex [n][2]
void user_exact_host(const UserKernels &u, size_t n, const double *coord, const double *q,
                     size_t nt, const long *idx, double *V)
{
  parallel_for((long)nt, [&](long t, int) {
    thread_local std::vector<double> x, y, z, fr, fi, g[6];
    const size_t i = idx ? (size_t)idx[t] : (size_t)t;
    x.resize(n);
    y.resize(n);
    z.resize(n);
    fr.resize(n);
    fi.resize(n);
    double *gp[6];
    for (int c = 0; c < 6; c++) {
      g[c].resize(n);
      gp[c] = g[c].data();
    }
    size_t m = 0;
    for (size_t j = 0; j < n; j++) {
      if (j == i) continue;
      x[m] = coord[3 * i] - coord[3 * j];
      y[m] = coord[3 * i + 1] - coord[3 * j + 1];
      z[m] = coord[3 * i + 2] - coord[3 * j + 2];
      m++;
    }
    const std::vector<double> xx(x.begin(), x.begin() + m);
    eval_grad(u, 0, m, xx.data(), y.data(), z.data(), fr.data(), fi.data(), gp);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    size_t k = 0;
    for (size_t j = 0; j < n; j++) {
      if (j == i) continue;
      if (u.complex) {
        const double qr = q[2 * j], qi = q[2 * j + 1];
        const double fre[4] = {fr[k], g[0][k], g[2][k], g[4][k]};
        const double fim[4] = {fi[k], g[1][k], g[3][k], g[5][k]};
        for (int c = 0; c < 4; c++) {
          acc[2 * c] += fre[c] * qr - fim[c] * qi;
          acc[2 * c + 1] += fre[c] * qi + fim[c] * qr;
        }
      } else {
        const double qj = q[j];
        acc[0] += fr[k] * qj;
        acc[1] += g[0][k] * qj;
        acc[2] += g[2][k] * qj;
        acc[3] += g[4][k] * qj;
      }
      k++;
    }
    const int w = u.complex ? 8 : 4;
    for (int c = 0; c < w; c++) V[(size_t)t * w + c] = acc[c];
  });
}

}  // namespace npme
