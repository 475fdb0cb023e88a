"""Pins the oracle's two-box pieces: the translated full kernel grid
(NPME_RecSumGrid_FullTrans_Func, src/RecSumGrid.cpp:663-794) and the
full-F convolution with sign reversal (NPME_RSTP_CalcTheta_FullF_ReverseSign,
src/RecSumV.cpp:851-1009), and the box-pair potential built from them.

  * at R0 = 0 the DFT of the full grid of the long-range part equals the
    even extension of the compact grid (npo_fgrid_compact, pinned elsewhere);
  * with no reversal the full-F theta equals a zero-padded full-grid
    correlation computed with numpy's FFT, and with an even F it equals the
    compact 8-parity theta;
  * reversing an axis equals numpy's F at the negated frequency, and equals
    the grid sampled at R0 with that component negated;
  * the box-pair potential matches the exact O(N^2) sum between the sets."""
import numpy as np
import pytest

from conftest import uniform_box


def _fold(N):
    return [np.abs(np.fft.fftfreq(N[d], 1.0 / N[d])).astype(int) for d in range(3)]


def _neg(F, axes):
    """F at the negated frequency along the given axes"""
    for ax in axes:
        F = np.roll(np.flip(F, axis=ax), 1, axis=ax)
    return F


def _numpy_theta_full(N, Bn, F, Q, oracle):
    """theta(m) = 1/(N1N2N3) sum_k conj(Qhat(k)) F(k) lam2(k) e^{-2 pi i k.m/N}
    with Q zero-padded to the full grid, truncated to the compact window"""
    M = [n // 2 for n in N]
    m = [np.where(np.arange(M[d]) < M[d] // 2, np.arange(M[d]), np.arange(M[d]) - M[d]) for d in range(3)]
    full_idx = [np.where(m[d] < 0, m[d] + N[d], m[d]) for d in range(3)]
    Qf = np.zeros(N, complex)
    Qf[np.ix_(*full_idx)] = Q
    lam = [oracle.lambda2(N[d], Bn) for d in range(3)]
    G = F * np.einsum("i,j,k->ijk", *lam) / np.prod(N)
    th = np.fft.fftn(np.conj(np.fft.fftn(Qf)) * G)
    return th[np.ix_(*full_idx)]


def test_theta_fullF_equals_numpy_correlation_and_reversal(oracle):
    rng = np.random.default_rng(11)
    N, Bn = (24, 20, 28), 6
    M = [n // 2 for n in N]
    Q = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    F = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    T = oracle.theta_fullF_reverse(N, Bn, F, Q)
    np.testing.assert_allclose(T, _numpy_theta_full(N, Bn, F, Q, oracle), atol=1e-11)
    for rev in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1)]:
        axes = [d for d in range(3) if rev[d]]
        Tr = oracle.theta_fullF_reverse(N, Bn, F, Q, rev)
        np.testing.assert_allclose(Tr, _numpy_theta_full(N, Bn, _neg(F, axes), Q, oracle),
                                   atol=1e-11, err_msg=str(rev))


def test_theta_fullF_with_even_F_equals_compact_8parity(oracle):
    rng = np.random.default_rng(12)
    N, Bn = (16, 24, 20), 4
    M = [n // 2 for n in N]
    Q = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    Fc = rng.standard_normal([n // 2 + 1 for n in N]) + 1j * rng.standard_normal([n // 2 + 1 for n in N])
    Ffull = Fc[np.ix_(*_fold(N))]
    Tc = oracle.theta_8parity(N, Bn, Fc, Q)
    for rev in [(0, 0, 0), (1, 1, 0), (0, 1, 1)]:
        np.testing.assert_allclose(oracle.theta_fullF_reverse(N, Bn, Ffull, Q, rev), Tc, atol=1e-11)


def test_fulltrans_at_zero_shift_is_the_compact_grid(oracle):
    N, Rdir = (24, 28, 20), 5.0
    k = oracle.kernel(oracle.LAPLACE, 6, Rdir)
    X, dl, a = np.array([8.0, 10.0, 7.0]), np.full(3, Rdir), (4.0, 4.0, 4.0)
    f = oracle.fgrid_fulltrans(k, N, a, dl, X, (0.0, 0.0, 0.0), part=oracle.LR)
    Fc = oracle.fgrid_compact(k, N, a, dl, X)
    F = oracle.fft3d(f, -1)
    np.testing.assert_allclose(F, Fc[np.ix_(*_fold(N))], atol=1e-10 * np.abs(Fc).max())
    # grid point order: sample n at n L/N, or (n - N) L/N for n >= N/2
    L = 2 * (X + dl)
    x = (np.arange(N[0]) - N[0] * (np.arange(N[0]) >= N[0] // 2)) * L[0] / N[0]
    inside = np.abs(x) <= X[0]
    ref = np.array([oracle.kernel_eval(k, oracle.LR, float(xx * xx))[0] for xx in x])
    np.testing.assert_allclose(f[inside, 0, 0].real, ref[inside], rtol=1e-14, atol=1e-15)


def test_fulltrans_axis_reversal_identity(oracle):
    """the grid of R0 at negated frequencies is the grid of R0 with that
    component negated (the sampling is symmetric about n = 0)"""
    N = (20, 24, 16)
    k = oracle.kernel(oracle.LAPLACE, 6, 6.0)
    X, dl, a = np.array([6.0, 7.0, 5.0]), np.full(3, 6.0), (4.0, 4.0, 4.0)
    R0 = np.array([30.0, -4.0, 2.5])
    F = oracle.fft3d(oracle.fgrid_fulltrans(k, N, a, dl, X, R0), -1)
    for axes in [(0,), (1,), (2,), (0, 2)]:
        R1 = R0.copy()
        R1[list(axes)] *= -1
        F1 = oracle.fft3d(oracle.fgrid_fulltrans(k, N, a, dl, X, R1), -1)
        np.testing.assert_allclose(_neg(F, axes), F1, atol=1e-12 * np.abs(F).max())


def test_fulltrans_refuses_overlapping_boxes(oracle):
    k = oracle.kernel(oracle.LAPLACE, 6, 6.0)
    X, dl = np.array([6.0, 7.0, 5.0]), np.full(3, 6.0)
    # the grid box spans |x| <= X + del = 12 on axis 1: R0 inside it
    with pytest.raises(ValueError, match="separation"):
        oracle.fgrid_fulltrans(k, (20, 24, 16), (4.0, 4.0, 4.0), dl, X, (11.0, 3.0, 0.0))
    assert oracle.min_distance_r0((11.0, 3.0, 0.0), 2 * (X + dl)) == 0.0
    assert abs(oracle.min_distance_r0((15.0, 0.0, 20.0), 2 * (X + dl)) - np.hypot(3.0, 9.0)) < 1e-14


def _two_boxes(nT, nS, side, shift, seed=0):
    cT, _ = uniform_box(nT, side, seed_c=3001 + seed, seed_q=3002 + seed)
    cS, qS = uniform_box(nS, side, seed_c=3003 + seed, seed_q=3004 + seed)
    return cT, cS + np.asarray(shift, float), qS


@pytest.mark.parametrize("rev", [(0, 0, 0), (1, 0, 0), (0, 1, 1)])
def test_boxpair_oracle_matches_exact(oracle, rev):
    """V at targets from sources in a separate box (Laplace, full kernel)"""
    nT, nS, side = 150, 200, 8.0
    shift = np.array([28.0, 6.0, -4.0])
    cT, cS, qS = _two_boxes(nT, nS, side, shift)
    k = oracle.kernel(oracle.LAPLACE, 6, 6.0)
    # the error falls with the grid spacing against the box-box gap and with
    # the extension width: 2.4e-7 at 48^3 and del = 10 (5.5e-5 at 40^3, del 6)
    N, Bn = (48, 48, 48), 8
    V = oracle.boxpair(k, N, Bn, cT, cS, qS, dl=np.full(3, 10.0), rev=rev)
    # the reversed grid is the grid of the source box mirrored through the
    # target box centre on those axes
    _, ct = oracle.box_dim_center(cT)
    _, cs = oracle.box_dim_center(cS)
    cSm = cS.copy()
    for d in range(3):
        if rev[d]:
            cSm[:, d] = cS[:, d] - 2 * (cs[d] - ct[d])
    Vx = oracle.exact(k, np.vstack([cT, cSm]), np.concatenate([np.zeros(nT), qS]))[:nT]
    m = oracle.compare_v(Vx, V.real)
    assert m["relV"] < 1e-6 and m["reldV"] < 5e-6, m
    assert np.abs(V.imag).max() < 1e-12 * np.abs(V.real).max()
