"""Pins the CPU oracle's building blocks against facts that do not come from
the oracle: closed forms, textbook values, the preprint's Table 1 and
independent numpy/scipy computations."""
import math

import numpy as np
import pytest
from scipy.interpolate import BSpline

from conftest import uniform_box


# ------------------------------------------------------------ B-splines
def test_bspline_integer_values(oracle):
    # Cardinal B-spline values at the integers (Euler-Frobenius numbers):
    # M_4: 1,4,1 /6 ; M_6: 1,26,66,26,1 /120 ; M_8: 1,120,1191,2416,1191,120,1 /5040.
    # M_6(3) = 0.55 is also the constant term of the reference's piece B_6_3.
    exp = {4: [0, 1, 4, 1, 0], 6: [0, 1, 26, 66, 26, 1, 0],
           8: [0, 1, 120, 1191, 2416, 1191, 120, 1, 0]}
    den = {4: 6.0, 6: 120.0, 8: 5040.0}
    for n, vals in exp.items():
        got = [oracle.bspline(n, w) for w in range(n + 1)]
        np.testing.assert_allclose(got, np.array(vals) / den[n], atol=1e-15)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 12, 16])
def test_bspline_matches_scipy_and_identities(oracle, n):
    ref = BSpline.basis_element(np.arange(n + 1), extrapolate=False)
    rng = np.random.default_rng(n)
    ws = np.concatenate([rng.uniform(0, n, 50), [0.0, 0.5, n - 1e-9]])
    for w in ws:
        v, d = oracle.bspline_dw(n, w)
        assert abs(v - float(np.nan_to_num(ref(w)))) < 1e-13
        # derivative against the analytic derivative of the scipy spline
        dref = float(np.nan_to_num(ref.derivative()(w))) if n > 2 else None
        if dref is not None and 0 < w < n and abs(w - round(w)) > 1e-6:
            assert abs(d - dref) < 1e-12
    # partition of unity, symmetry, zero outside the support
    for t in rng.uniform(0, 1, 20):
        assert abs(sum(oracle.bspline(n, t + j) for j in range(n)) - 1.0) < 1e-14
        assert abs(oracle.bspline(n, t + 1) - oracle.bspline(n, n - t - 1)) < 1e-14
    assert oracle.bspline(n, -1e-9) == 0.0 and oracle.bspline(n, n + 1e-9) == 0.0
    assert abs(oracle.bspline(n, n)) < 1e-15


# ------------------------------------------------------------ DM split
def test_dm_table1_known_answer(oracle):
    # preprint Table 1 (p.15): 1/r, Nder = 4, Rdir = 1
    a, b = oracle.kernel(oracle.LAPLACE, 4, 1.0).coeffs()
    np.testing.assert_allclose(a, [2.4609375, -3.28125, 2.953125, -1.40625, 0.2734375],
                               rtol=0, atol=1e-14)
    np.testing.assert_allclose(b[:4], [2 * (i + 1) * a[i + 1] for i in range(4)], atol=1e-14)
    assert b[4] == 0.0


@pytest.mark.parametrize("alpha,Nder,R", [(-1.0, 6, 8.0), (-1.0, 5, 8.0), (0.5, 6, 8.0),
                                          (1.5, 8, 7.0), (-3.0, 6, 5.0)])
def test_dm_scaling_and_derivative_match(oracle, alpha, Nder, R):
    kind = oracle.LAPLACE if alpha == -1.0 else oracle.RALPHA
    a1, _ = oracle.kernel(kind, Nder, 1.0, alpha=alpha).coeffs()
    aR, _ = oracle.kernel(kind, Nder, R, alpha=alpha).coeffs()
    # scaling law a_n(R) = a_n(1) R^(alpha - 2n)  (preprint Eq. 30)
    np.testing.assert_allclose(aR, a1 * R ** (alpha - 2 * np.arange(Nder + 1)), rtol=1e-11)
    # defining identity: (1/r d/dr)^k P (R) = (1/r d/dr)^k r^alpha (R), k <= Nder
    for kk in range(Nder + 1):
        lhs = sum(aR[p] * 2 ** kk * math.factorial(p) / math.factorial(p - kk) * R ** (2 * (p - kk))
                  for p in range(kk, Nder + 1))
        rhs = np.prod([alpha - 2 * j for j in range(kk)]) * R ** (alpha - 2 * kk)
        assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(rhs))


def _helmholtz_radial_derivs(k0, R, nmax):
    """(1/r d/dr)^n e^{ik0 r}/r at r = R by Taylor composition in t = (r^2-R^2)/2
    (so that 1/r d/dr = d/dt): r = R sqrt(1 + 2t/R^2), expand everything as a
    truncated power series in t -- no recurrence from the oracle."""
    K = nmax + 2
    # h(t) = r - R = R (sqrt(1 + 2t/R^2) - 1), binomial series
    h = np.zeros(K, complex)
    c = 1.0
    for m in range(1, K):
        c *= (0.5 - (m - 1)) / m
        h[m] = R * c * (2.0 / R ** 2) ** m

    def mul(x, y):
        return np.convolve(x, y)[:K]

    def series_of(fcoef):   # sum_j fcoef[j] h^j
        out = np.zeros(K, complex)
        p = np.zeros(K, complex)
        p[0] = 1
        for j in range(K):
            out += fcoef[j] * p
            p = mul(p, h)
        return out
    e = [np.exp(1j * k0 * R) * (1j * k0) ** j / math.factorial(j) for j in range(K)]
    g = [(1.0 / R) * (-1.0 / R) ** j for j in range(K)]
    s = mul(series_of(e), series_of(g))
    return [s[n] * math.factorial(n) for n in range(nmax + 1)]


@pytest.mark.parametrize("k0,Nder,R", [(1.1, 6, 4.0), (0.5, 8, 10.0), (0.3 + 0.05j, 6, 8.0)])
def test_dm_helmholtz_derivative_match(oracle, k0, Nder, R):
    k = oracle.kernel(oracle.HELMHOLTZ, Nder, R, k0=k0)
    a, b = k.coeffs()
    fr = _helmholtz_radial_derivs(k0, R, Nder)
    for kk in range(Nder + 1):
        lhs = sum(a[p] * 2 ** kk * math.factorial(p) / math.factorial(p - kk) * R ** (2 * (p - kk))
                  for p in range(kk, Nder + 1))
        assert abs(lhs - fr[kk]) <= 1e-9 * max(1e-3, abs(fr[kk]))


@pytest.mark.parametrize("kind,alpha,k0", [(0, -1.0, 0), (1, 0.5, 0), (1, 1.5, 0), (2, -1.0, 1.1)])
def test_kernel_parts_consistent(oracle, kind, alpha, k0):
    R = 6.0
    k = oracle.kernel(kind, 6, R, alpha=alpha, k0=k0)
    for r in [0.3, 1.0, 2.5, 5.99, 6.01, 9.0]:
        f, f1 = oracle.kernel_eval(k, oracle.FULL, r * r)
        l, l1 = oracle.kernel_eval(k, oracle.LR, r * r)
        s, s1 = oracle.kernel_eval(k, oracle.SR, r * r)
        # closed forms of the full kernel
        if kind == 0:
            ef, ef1 = 1 / r, -1 / r ** 3
        elif kind == 1:
            ef, ef1 = r ** alpha, alpha * r ** (alpha - 2)
        else:
            ef = np.exp(1j * k0 * r) / r
            ef1 = (1j * k0 * ef - ef / r) / r
        assert abs(f - ef) < 1e-13 * max(1, abs(ef)) and abs(f1 - ef1) < 1e-12 * max(1, abs(ef1))
        assert abs(l + s - f) < 1e-12 * max(1, abs(f))
        assert abs(l1 + s1 - f1) < 1e-11 * max(1, abs(f1))
        if r > R:
            assert s == 0 and s1 == 0
    # LR is continuous through Rdir (matched value and first derivative)
    lin, lin1 = oracle.kernel_eval(k, oracle.LR, (R * (1 - 1e-9)) ** 2)
    lout, lout1 = oracle.kernel_eval(k, oracle.LR, (R * (1 + 1e-9)) ** 2)
    assert abs(lin - lout) < 1e-7 * max(1, abs(lout))
    assert abs(lin1 - lout1) < 1e-6 * max(1, abs(lout1))


# ------------------------------------------------------------ size picker
def test_fft_size_picker_bruteforce(oracle):
    def smooth(n):
        for p in (2, 3, 5, 7):
            while n % p == 0:
                n //= p
        return n == 1
    for n in list(range(1, 400)) + [513, 777, 961, 1001, 1153]:
        exp = next(N for N in range(4, 10 * n + 8, 4) if N >= n and smooth(N))
        assert oracle.optimal_grid_size(n) == exp
        N, blk = oracle.fft_size_block(n, 20)
        cands = [m for m in range(2, N) if N % (2 * m) == 0 and m <= N // 4]
        best = min(cands, key=lambda m: (abs(m - 20), m)) if cands else 1
        if not cands or abs(best - 20) >= N:
            best = 1   # src/SupportFunctions.cpp:440-452 starts the search at minDiff = N
        assert (N, blk) == (exp, best)


# ------------------------------------------------------------ RSEM
def test_rsem_grid_choices(oracle):
    coord, q = uniform_box(10000, 40.0)
    assert oracle.rsem_static(q, coord, 6, -1.0, 1e-6, 8.0) == ((120, 120, 120), (20, 20, 20), 5)
    eps, rel = oracle.rsem_static_pred_error(q, coord, 6, -1.0, 5, 8.0, (120, 120, 120))
    assert rel and 1e-7 < eps <= 1e-6
    n = 1000000
    L = (n / 0.1) ** (1 / 3)
    coord, q = uniform_box(n, L)
    assert oracle.rsem_static(q, coord, 6, -1.0, 1e-6, 8.0) == ((512,) * 3, (16,) * 3, 6)
    assert oracle.rsem_static(q, coord, 8, -1.0, 1e-6, 8.0) == ((432,) * 3, (18,) * 3, 7)
    with pytest.raises(ValueError):
        oracle.rsem_static(q, coord, 6, 0.5, 1e-6, 8.0)   # no alpha = 0.5 row


@pytest.mark.parametrize("alpha,kind,Bn,Nder", [(-1.0, "LAPLACE", 6, 4), (-1.0, "LAPLACE", 6, 6),
                                                 (-1.0, "LAPLACE", 8, 7), (1.0, "RALPHA", 6, 5),
                                                 (1.0, "RALPHA", 8, 6)])
def test_rsem_pred_error_tracks_measured(oracle, alpha, kind, Bn, Nder):
    """The RSEM static error model (CalcPredError_V, src/EstimateParm.cpp:
    421-558) predicts the absolute RMS error of V.  Over a sweep of grids
    inside the model's reliable range of u = N Rdir / L, its prediction
    must track the measured whole-path error of the oracle's CalcV against
    the exact O(N^2) sum within a factor 2.5 either way, for 1/r and r^1;
    and both must fall as the grid is refined.  A dropped factor (the
    Rdir^alpha or (Rdir/L)^1.5 scale, the charge norm) or a wrong sign in
    the fitted polynomial P(u) moves the ratio by orders of magnitude."""
    n, Lb, Rdir = 3000, 28.0, 6.0
    coord, q = uniform_box(n, Lb)
    k = oracle.kernel(getattr(oracle, kind), Nder, Rdir, alpha=alpha)
    Ve = oracle.exact(k, coord, q)
    preds, meas = [], []
    for Ng in (60, 72, 84, 96):
        eps, reliable = oracle.rsem_static_pred_error(q, coord, Bn, alpha, Nder, Rdir, (Ng,) * 3)
        assert reliable
        m = oracle.compare_v(Ve, oracle.calcv(k, (Ng,) * 3, Bn, coord, q))
        print(alpha, Bn, Nder, Ng, "predicted %.3e measured %.3e ratio %.2f" % (eps, m["errV"],
                                                                                m["errV"] / eps))
        assert 1 / 2.5 <= m["errV"] / eps <= 2.5
        preds.append(eps)
        meas.append(m["errV"])
    assert all(a > b for a, b in zip(preds, preds[1:]))
    assert all(a > b for a, b in zip(meas, meas[1:]))


def test_rsem_helmholtz_grid_choice(oracle):
    n = 500000
    coord, _ = uniform_box(n, 120.0, 3001, 3002)
    rng = np.random.default_rng(3003)
    qc = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert oracle.rsem_helmholtz(qc, coord, 8, 0.5, 1e-6, 10.0) == ((280,) * 3, (20,) * 3, 8)


# ------------------------------------------------------------ geometry
def _mstart(w):
    """NPME_GetQpme_DFT_StartIndex restated with numpy (C truncation)"""
    return np.where(w < 0, np.trunc(-w) + 1, np.trunc(-w)).astype(np.int64)


def _windows_fit(xmin, xmax, Xp, dl, N, Bn):
    """do the spline windows of every charge in [xmin, xmax] fit the compact
    grid [-N/4, N/4-1] when the half box is Xp (top-aligned: R + Xp/2 =
    xmax)?  Brute force over 20,001 positions including both ends."""
    R = xmax - 0.5 * Xp
    L = 2.0 * (Xp + dl)
    x = np.linspace(xmin, xmax, 20001)
    m = _mstart(N / L * (x - R))
    return m.min() >= -(N // 4) and m.max() + Bn - 1 <= N // 4 - 1


@pytest.mark.parametrize("case", range(6))
def test_pme_corr_box_is_minimal(oracle, case):
    """NPME_RecSumInterface_GetPMECorrBox (src/RecSumSupportFunctions.cpp:
    137-147, 276-296) against brute force: with the returned box every
    spline window of the charges fits [-N/4, N/4-1]; the box is not larger
    than the smallest one that fits (found by bisection over the half box
    with the same top alignment) by more than two grid spacings (the
    formula keeps a margin of one spline order); it never shrinks below
    the charges' extent X0; and its top edge is the charges' maximum."""
    rng = np.random.default_rng(100 + case)
    X0 = rng.uniform(5.0, 300.0, 3)
    R0 = rng.uniform(-50.0, 50.0, 3)
    Bn = int(rng.choice([4, 6, 8, 12]))
    # coarse grids or narrow extensions, where the correction is active
    # (with del = Rdir and grid spacings below del / (Bn + 1) the charges'
    # own extent already fits and X = X0)
    dl = np.full(3, float(rng.choice([0.5, 1.0, 2.0, 8.0])))
    N = np.array([4 * int(rng.integers(Bn + 1, max(Bn + 2, int(X0[d] / 2)))) for d in range(3)])
    X, R = oracle.pme_corr_box(X0, R0, N, dl, Bn)
    active = 0
    for d in range(3):
        xmin, xmax = R0[d] - X0[d] / 2, R0[d] + X0[d] / 2
        assert X[d] >= X0[d]
        assert abs((R[d] + X[d] / 2) - xmax) <= 1e-12 * (abs(xmax) + X[d])
        assert _windows_fit(xmin, xmax, X[d], dl[d], N[d], Bn)
        if X[d] == X0[d]:
            continue  # the charges' own extent already fits: minimal
        lo, hi = X0[d], X[d]
        if _windows_fit(xmin, xmax, lo, dl[d], N[d], Bn):
            Xmin = lo
        else:
            for _ in range(60):
                mid = 0.5 * (lo + hi)
                if _windows_fit(xmin, xmax, mid, dl[d], N[d], Bn):
                    hi = mid
                else:
                    lo = mid
            Xmin = hi
        h = 2.0 * (X[d] + dl[d]) / N[d]
        print(case, d, "N", N[d], "Bn", Bn, "X", X[d], "Xmin", Xmin, "slack/h", (X[d] - Xmin) / h)
        assert 0.0 <= X[d] - Xmin <= 2.0 * h
        active += 1
    assert active >= 1


def test_cell_index_containment(oracle):
    coord, _ = uniform_box(5000, (30.0, 45.0, 60.0))
    Rdir, nN, nC = 8.0, 2, 3
    nA, tA, tB = oracle.cell_index(coord, Rdir, nN, nC)
    csB = Rdir / nN
    csA = nC * csB
    mn = coord.min(0)
    L = coord.max(0) - mn
    assert list(nA) == [int(L[d] / csA) + 1 for d in range(3)]
    Nc = nC ** 3
    a = np.stack([tA // (nA[1] * nA[2]), (tA // nA[2]) % nA[1], tA % nA[2]], 1).astype(float)
    loc = tB % Nc
    b = np.stack([loc // (nC * nC), (loc // nC) % nC, loc % nC], 1).astype(float)
    assert np.all(tB // Nc == tA)
    x = coord - mn
    lo = a * csA + b * csB
    assert np.all(x >= lo - 1e-12) and np.all(x <= lo + csB + 1e-12)


# ------------------------------------------------------------ FFT
@pytest.mark.parametrize("N", [4, 8, 12, 30, 60, 105, 120, 256, 280, 432, 512])
def test_fft1d_matches_numpy(oracle, N):
    rng = np.random.default_rng(N)
    z = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    np.testing.assert_allclose(oracle.fft1d(z, -1), np.fft.fft(z), atol=1e-11 * N)
    np.testing.assert_allclose(oracle.fft1d(z, +1), np.fft.ifft(z) * N, atol=1e-11 * N)


def test_fft3d_matches_numpy(oracle):
    rng = np.random.default_rng(7)
    z = rng.standard_normal((12, 10, 14)) + 1j * rng.standard_normal((12, 10, 14))
    np.testing.assert_allclose(oracle.fft3d(z, -1), np.fft.fftn(z), atol=1e-10)


def test_lambda2_definition(oracle):
    for N, n in [(40, 6), (64, 8), (60, 4)]:
        M = BSpline.basis_element(np.arange(n + 1), extrapolate=False)
        Mp = np.nan_to_num(M(np.arange(n + 1).astype(float)))
        k = np.fft.fftfreq(N, 1.0 / N)
        den = np.array([np.sum(Mp * np.exp(-2j * np.pi * kk * np.arange(n + 1) / N)) for kk in k])
        np.testing.assert_allclose(oracle.lambda2(N, n), 1 / np.abs(den) ** 2, rtol=1e-12)


def test_compare_v_metric(oracle):
    rng = np.random.default_rng(1)
    Vr = rng.standard_normal((50, 4))
    V = Vr + 1e-3 * rng.standard_normal((50, 4))
    m = oracle.compare_v(Vr, V)
    d = V - Vr
    assert abs(m["errV"] - np.sqrt(np.mean(d[:, 0] ** 2))) < 1e-15
    assert abs(m["errdV"] - np.sqrt(np.mean(d[:, 1:] ** 2))) < 1e-15
    assert abs(m["magV"] - np.mean(np.abs(Vr[:, 0]))) < 1e-15
    # quirk of the reference: the "gradient magnitude" averages components 0..2
    assert abs(m["magdV"] - np.mean(np.abs(Vr[:, :3]))) < 1e-15


# ------------------------------------------------------------ Ewald (erfc) split
@pytest.mark.parametrize("R,tol", [(4.0, 1e-6), (6.0, 1e-6), (8.0, 1e-6), (6.0, 1e-4)])
def test_erf_split_beta(oracle, R, tol):
    """beta of EwaldSplit LaplaceOrig: erfc(beta R)/R is tol times the
    asymptotic factor 1 - 1/(2x^2) + 3/(4x^4), x = beta R (scipy's erfc), and
    beta R solves the defining equation x^2 + ln x + ln(sqrt(pi) tol R) = 0."""
    from scipy.special import erfc
    b = oracle.rdir2beta(R, tol)
    x = b * R
    assert abs(x * x + math.log(x) + math.log(math.sqrt(math.pi) * tol * R)) < 1e-11
    ratio = erfc(x) / R / tol
    assert 0.85 < ratio < 1.0
    assert ratio == pytest.approx(1 - 1 / (2 * x * x) + 3 / (4 * x ** 4), abs=0.01)


def test_erf_split_parts(oracle):
    """LR + SR = 1/r inside Rdir (SR = 0 outside), LR(0) = 2 beta/sqrt(pi),
    and f1 = (1/r) df/dr against central differences of f0, for both parts;
    the values against scipy's erf / erfc."""
    from scipy.special import erf, erfc
    R, b = 6.0, oracle.rdir2beta(6.0, 1e-6)
    k = oracle.kernel_erf(R, b)
    assert oracle.kernel_eval(k, oracle.LR, 0.0)[0] == pytest.approx(2 * b / math.sqrt(math.pi), rel=1e-14)
    for r in (1e-5, 0.3, 1.0, 2.5, 5.9, 7.0):
        full = oracle.kernel_eval(k, oracle.FULL, r * r)[0]
        lr, lr1 = oracle.kernel_eval(k, oracle.LR, r * r)
        sr, sr1 = oracle.kernel_eval(k, oracle.SR, r * r)
        assert lr == pytest.approx(erf(b * r) / r, rel=1e-12)
        if r < R:
            assert lr + sr == pytest.approx(full, rel=1e-12)
            assert sr == pytest.approx(erfc(b * r) / r, rel=1e-12)
        else:
            assert sr == 0.0
        h = 1e-5 * max(r, 1e-3)
        if r > 1e-3:
            for part, f1 in ((oracle.LR, lr1), (oracle.SR, sr1)):
                if part == oracle.SR and r >= R:
                    continue
                d = (oracle.kernel_eval(k, part, (r + h) ** 2)[0]
                     - oracle.kernel_eval(k, part, (r - h) ** 2)[0]) / (2 * h)
                assert f1 == pytest.approx(d / r, rel=1e-6)


def test_erf_split_whole_path_vs_exact(oracle):
    """The oracle's CalcV with the erfc split reaches the split's truncation
    floor (~tol) against the exact sum on a fine grid, and is worse than the
    derivative match on the grid RSEM picks (the DM grid, as in the
    reference's AddMissingKeywords)."""
    coord, q = uniform_box(3000, 30.0)
    R = 8.0
    N, nb, nd = oracle.rsem_static(q, coord, 6, -1.0, 1e-6, R)
    k = oracle.kernel_erf(R, oracle.rdir2beta(R, 1e-6))
    kd = oracle.kernel(oracle.LAPLACE, nd, R)
    Ve = oracle.exact(kd, coord, q)
    fine = tuple(int(n * 3 // 2 // 4 * 4) for n in N)
    e_fine = oracle.compare_v(Ve, oracle.calcv(k, fine, 6, coord, q))["relV"]
    e_rsem = oracle.compare_v(Ve, oracle.calcv(k, N, 6, coord, q))["relV"]
    e_dm = oracle.compare_v(Ve, oracle.calcv(kd, N, 6, coord, q))["relV"]
    print("erf fine", e_fine, "erf rsem", e_rsem, "dm", e_dm)
    assert e_fine < 2e-6 and e_dm < e_rsem < 1e-5
