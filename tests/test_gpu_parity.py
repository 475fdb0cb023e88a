"""GPU parity: the HIP CalcV path against the CPU oracle (same inputs, same
grid) and against the exact O(N^2) FP64 sum, through the C ABI.

Tolerances (relative RMS, compareV metric):
  fp64 GPU path vs oracle CalcV ............. 1e-11 (V and grad)
  GPU CalcV_exact vs oracle exact ........... 1e-12
  mixed GPU path vs oracle CalcV ............ 5e-7 (V), 5e-6 (grad)
  mixed GPU path vs exact sum ............... 1e-6 (V) for 1/r at tol 1e-6
  mixed GPU path vs oracle, every spline order 4..12 ... 1e-6 (V), 1e-5 (grad)
  (the mixed form refuses orders above 12)
The integer work (cell and cluster indices, working order, spline windows)
is compared bit for bit with the oracle in test_gpu_indices.py; here the
results must also be bit-identical run to run."""
import numpy as np
import pytest

from conftest import uniform_box

pytestmark = pytest.mark.gpu


def _kw(**k):
    import npme_amd as N
    kk = N.NPME_KeywordInput()
    for a, b in k.items():
        kk.AddKeyword(a, b)
    return kk


def _real_setup(coord, q, prec, funcType="Laplace", Rdir=5.0, Bn=6, **extra):
    import npme_amd as N
    print("case", funcType, prec, len(q), flush=True)
    k = _kw(calcType="pme", funcType=funcType, Rdir=Rdir, BnOrder=Bn, tol=1e-6,
            precision=prec, **extra)
    assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q), N.npme_last_error()
    h = N.NPME_Interface_SetUpRealKernel(k, len(q), coord)
    assert h is not None, N.npme_last_error()
    return h


def _grid(h):
    return (h.Get_N1(), h.Get_N2(), h.Get_N3())


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_laplace_config1(oracle, prec):
    """Config 1: N=10,000 uniform in 40^3, Bn=6, Rdir=8, RSEM grid at 1e-6."""
    coord, q = uniform_box(10000, 40.0)
    h = _real_setup(coord, q, prec, Rdir=8.0)
    assert _grid(h) == (120, 120, 120) and h.Get_Nder() == 5
    V = h.calcv(q)
    assert np.array_equal(V, h.calcv(q)), "CalcV is not deterministic"
    k = oracle.kernel(oracle.LAPLACE, h.Get_Nder(), 8.0)
    Vo = oracle.calcv(k, _grid(h), 6, coord, q)
    Ve = oracle.exact(k, coord, q)
    mo, me = oracle.compare_v(Vo, V), oracle.compare_v(Ve, V)
    print(prec, "vs oracle", mo, "vs exact", me)
    if prec == "fp64":
        assert mo["relV"] < 1e-11 and mo["reldV"] < 1e-11
    else:
        assert mo["relV"] < 5e-7 and mo["reldV"] < 5e-6
    assert me["relV"] <= 1e-6


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("Bn", [4, 8, 10, 12, 14, 16])
def test_spline_orders_match_oracle(oracle, prec, Bn):
    """Spreading and interpolation for each spline order the kernels are
    compiled for (4, 6, 8, 10) and the generic order (12, 14, 16), on a
    given 64^3 grid (80^3 for order 16, which needs N/4 > 16), against the
    oracle's reciprocal + direct sum.  The B-spline weights are computed in
    FP64 and rounded once.  The mixed form is held to the requested 1e-6 at
    every order it accepts (up to 12: measured 2.5e-7 at orders 4-12); at
    order 14 it measured 1.06e-6, so set-up refuses orders above 12."""
    import npme_amd as N
    coord, q = uniform_box(2500, 25.0)
    Ng = 80 if Bn == 16 else 64
    print("case spline order", Bn, prec, Ng, flush=True)
    k = _kw(calcType="pme", funcType="Laplace", Rdir=5.0, BnOrder=Bn, tol=1e-6, precision=prec,
            N1=Ng, N2=Ng, N3=Ng, n1=8, n2=8, n3=8, nDeriv=6)
    assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q), N.npme_last_error()
    h = N.NPME_Interface_SetUpRealKernel(k, len(q), coord)
    if prec == "mixed" and Bn > 12:
        assert h is None and "fp64" in N.npme_last_error()
        return
    assert h is not None, N.npme_last_error()
    V = h.calcv(q)
    ko = oracle.kernel(oracle.LAPLACE, h.Get_Nder(), 5.0)
    mo = oracle.compare_v(oracle.calcv(ko, (Ng, Ng, Ng), Bn, coord, q), V)
    print(Bn, prec, mo)
    if prec == "fp64":
        assert mo["relV"] < 1e-11 and mo["reldV"] < 1e-11
    else:
        assert mo["relV"] < 1e-6 and mo["reldV"] < 1e-5


def test_exact_matches_oracle(oracle):
    coord, q = uniform_box(3000, 20.0)
    h = _real_setup(coord, q, "mixed")
    Vg = np.zeros((3000, 4))
    assert h.CalcV_exact(Vg, 3000, q, coord)
    k = oracle.kernel(oracle.LAPLACE, 5, 5.0)
    Ve = oracle.exact(k, coord, q)
    assert oracle.compare_v(Ve, Vg)["relV"] < 1e-12
    idx = np.array([0, 17, 2999, 1500])
    np.testing.assert_allclose(h.calcv_exact_subset(q, idx), Vg[idx], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_ralpha_anisotropic(oracle, prec):
    """r^0.5, DM, anisotropic box, explicit grid (no RSEM row for 0.5).  The
    grid 72x72x168 has a method error of 2.5e-7 against the exact sum (the
    oracle on the same grid); a coarser 60x60x140 grid gives 1.09e-6."""
    coord, q = uniform_box(3000, (20.0, 20.0, 60.0))
    h = _real_setup(coord, q, prec, funcType="Ralpha", alpha=0.5, Rdir=6.0, nDeriv=6,
                    N1=72, N2=72, N3=168, n1=12, n2=12, n3=12)
    V = h.calcv(q)
    k = oracle.kernel(oracle.RALPHA, 6, 6.0, alpha=0.5)
    Vo = oracle.calcv(k, _grid(h), 6, coord, q)
    mo = oracle.compare_v(Vo, V)
    print(prec, mo)
    assert mo["relV"] < (1e-11 if prec == "fp64" else 1e-6)
    assert oracle.compare_v(oracle.exact(k, coord, q), V)["relV"] < 1e-6


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("compactF", [1, 0])
def test_laplace_orig_split(oracle, prec, compactF):
    """EwaldSplit LaplaceOrig (erf/erfc split, beta from Rdir and tol) on the
    RSEM grid: against the oracle's NPME_InterfaceReal_Laplace_Original
    CalcV (same beta, same grid, and the reference's direct-sum pair set:
    every point pair of the interacting cell pairs, no cutoff) and against
    the exact sum, which it meets less well than the derivative match on
    that grid (as in the reference).  The direct sum runs in FP64 in both
    forms.  useCompactF 0 gives the same result as 1 (the full F of a radial
    kernel is the compact F extended)."""
    coord, q = uniform_box(3000, 30.0)
    h = _real_setup(coord, q, prec, Rdir=8.0, EwaldSplit="LaplaceOrig", useCompactF=compactF)
    V = h.calcv(q)
    assert "LaplaceOrig" in h.Print() and "interacting cell pairs" in h.Print()
    k = oracle.kernel_erf(8.0, oracle.rdir2beta(8.0, 1e-6))
    Vo = oracle.calcv_laplace_orig(k, _grid(h), 6, coord, q, nNeigh=2, nCellClust1D=3)
    # the reference's pair set differs from a cutoff at Rdir by the tail of
    # the kept pairs beyond Rdir
    tail = oracle.compare_v(oracle.calcv(k, _grid(h), 6, coord, q), Vo)["relV"]
    assert tail > 1e-9
    mo = oracle.compare_v(Vo, V)
    me = oracle.compare_v(oracle.exact(oracle.kernel(oracle.LAPLACE, 6, 8.0), coord, q), V)
    print(prec, compactF, "vs oracle", mo, "vs exact", me)
    if prec == "fp64":
        assert mo["relV"] < 1e-11 and mo["reldV"] < 1e-11
    else:
        assert mo["relV"] < 1e-6 and mo["reldV"] < 5e-6
    assert me["relV"] < 1e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_helmholtz(oracle, prec):
    import npme_amd as N
    coord, _ = uniform_box(3000, 24.0)
    rng = np.random.default_rng(11)
    qc = rng.standard_normal(3000) + 1j * rng.standard_normal(3000)
    print("case helmholtz", prec, flush=True)
    k = _kw(calcType="pme", funcType="Helmholtz", k0_r=1.1, k0_i=0.0, Rdir=5.0, BnOrder=8,
            tol=1e-6, precision=prec)
    assert N.NPME_AddMissingKeywords(k, 3000, coord, chargeComplex=qc), N.npme_last_error()
    h = N.NPME_Interface_SetUpComplexKernel(k, 3000, coord)
    assert h is not None, N.npme_last_error()
    V = h.calcv(qc)
    ko = oracle.kernel(oracle.HELMHOLTZ, h.Get_Nder(), 5.0, k0=1.1)
    Vo = oracle.calcv(ko, _grid(h), 8, coord, qc)
    mo = oracle.compare_v(Vo, V)
    print(prec, mo)
    assert mo["relV"] < (1e-11 if prec == "fp64" else 1e-6)
    Vg = np.zeros_like(V)
    assert h.CalcV_exact(Vg, 3000, qc, coord)
    assert oracle.compare_v(oracle.exact(ko, coord, qc), Vg)["relV"] < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("Bn", [14, 16])
def test_helmholtz_high_orders(oracle, prec, Bn):
    """Complex charges at spline orders 14 and 16: the complex FP64 theta
    block of an interpolation tile needs tile edges 2 and 1 to stay inside
    64 KB of LDS (an 80^3 grid, N/4 > 16).  The mixed form is refused above
    order 12."""
    import npme_amd as N
    coord, _ = uniform_box(2000, 22.0)
    rng = np.random.default_rng(12)
    qc = rng.standard_normal(2000) + 1j * rng.standard_normal(2000)
    print("case helmholtz order", Bn, prec, flush=True)
    k = _kw(calcType="pme", funcType="Helmholtz", k0_r=0.8, k0_i=0.1, Rdir=5.0, BnOrder=Bn,
            tol=1e-6, precision=prec, N1=80, N2=80, N3=80, n1=8, n2=8, n3=8, nDeriv=8)
    assert N.NPME_AddMissingKeywords(k, 2000, coord, chargeComplex=qc), N.npme_last_error()
    h = N.NPME_Interface_SetUpComplexKernel(k, 2000, coord)
    if prec == "mixed":
        assert h is None and "fp64" in N.npme_last_error()
        return
    assert h is not None, N.npme_last_error()
    V = h.calcv(qc)
    ko = oracle.kernel(oracle.HELMHOLTZ, 8, 5.0, k0=0.8 + 0.1j)
    mo = oracle.compare_v(oracle.calcv(ko, (80, 80, 80), Bn, coord, qc, nblk=(8, 8, 8)), V)
    print(Bn, prec, mo)
    if prec == "fp64":
        assert mo["relV"] < 1e-11 and mo["reldV"] < 1e-11
    else:
        assert mo["relV"] < 1e-6 and mo["reldV"] < 1e-5


def test_linearity_and_reset_coord(oracle):
    coord, q = uniform_box(5000, 30.0)
    h = _real_setup(coord, q, "fp64", Rdir=6.0)
    V1 = h.calcv(q)
    q2 = np.roll(q, 3)
    V2 = h.calcv(q2)
    V12 = h.calcv(2.0 * q - 0.5 * q2)
    np.testing.assert_allclose(V12, 2.0 * V1 - 0.5 * V2, atol=1e-11)
    # ResetCoord to a translated geometry: potentials are translation invariant
    assert h.ResetCoord(5000, coord + np.array([3.0, -7.0, 11.0]))
    np.testing.assert_allclose(h.calcv(q), V1, atol=1e-9)


@pytest.mark.parametrize("mode", ["device", "host"])
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_setup_rejects_non_finite_coordinates(mode, bad):
    import npme_amd as N
    coord, q = uniform_box(3000, 25.0)
    k = _kw(calcType="pme", funcType="Laplace", Rdir=5.0, BnOrder=6, tol=1e-6, setup=mode)
    assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q), N.npme_last_error()
    c2 = coord.copy()
    c2[1234, 1] = bad
    assert N.NPME_Interface_SetUpRealKernel(k, len(q), c2) is None
    assert "finite" in N.npme_last_error()
    h = N.NPME_Interface_SetUpRealKernel(k, len(q), coord)
    assert h is not None
    assert not h.ResetCoord(len(q), c2) and "finite" in N.npme_last_error()
    # the handle refuses work until a set-up succeeds
    V = np.zeros((len(q), 4))
    assert not h.CalcV(V, len(q), q, coord) and "no valid set-up" in N.npme_last_error()
    assert h.ResetCoord(len(q), coord)
    assert np.all(np.isfinite(h.calcv(q)))


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_reset_coord_reuses_kernel_grid(prec):
    """ResetCoord to a translated geometry keeps the box size, grid and
    kernel, so G is reused; the result is bit-identical to a fresh set-up
    of the translated charges.  A stretched geometry recomputes G."""
    import npme_amd as N
    coord, q = uniform_box(5000, 30.0)
    # positions on a grid of 2^-20, so the translated box has exactly the
    # same size (the shifted coordinates are exact)
    coord = np.round(coord * 2.0 ** 20) / 2.0 ** 20
    h = _real_setup(coord, q, prec, Rdir=6.0)
    assert "kernel grid G: computed" in h.Print()
    c2 = coord + np.array([5.0, -3.0, 2.5])
    assert h.ResetCoord(len(q), c2)
    assert "kernel grid G: reused" in h.Print()
    V = h.calcv(q)
    k = _kw(calcType="pme", funcType="Laplace", Rdir=6.0, BnOrder=6, tol=1e-6, precision=prec)
    assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q)
    fresh = N.NPME_Interface_SetUpRealKernel(k, len(q), c2)
    assert fresh is not None and "kernel grid G: computed" in fresh.Print()
    assert _grid(fresh) == _grid(h)
    assert np.array_equal(V, fresh.calcv(q))
    c3 = coord * np.array([1.0, 1.0, 1.05])
    assert h.ResetCoord(len(q), c3)
    assert "kernel grid G: computed" in h.Print()
    fresh3 = N.NPME_Interface_SetUpRealKernel(k, len(q), c3)
    assert np.array_equal(h.calcv(q), fresh3.calcv(q))


def test_calcv_rejects_wrong_size():
    import npme_amd as N
    coord, q = uniform_box(2000, 20.0)
    h = _real_setup(coord, q, "mixed")
    V = np.zeros((1999, 4))
    assert not h.CalcV(V, 1999, q[:1999], coord[:1999])
    assert "nCharge" in N.npme_last_error()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("n,box,nNeigh,nClust", [(6000, 35.0, 2, 3), (220000, 130.0, 2, 3),
                                                 (6000, 35.0, 3, 3), (60000, 84.0, 3, 3),
                                                 (6000, 35.0, 4, 4)])
def test_direct_sum_paths_agree(oracle, prec, n, box, nNeigh, nClust):
    """The neighbour-list and the cell-row direct sums give the same V.  The
    cases build the list with each walk: a wavefront per target (up to
    32,768 charges), a wavefront per cell (above), both also with the
    widest stencil they take (nNeigh 3), and a thread per target (nNeigh 4).
    In FP64 the two sums agree to rounding.  In the mixed form each is
    checked against the FP64 path: the two FP32 sums round differently, so
    their mutual difference is up to the sum of their errors (measured at
    60000 charges in 84^3: list 1.8e-7 / 1.04e-6, cell 1.8e-7 / 1.02e-6 in
    V / grad V)."""
    print("case direct paths", n, prec, nNeigh, flush=True)
    coord, q = uniform_box(n, box)
    kw = dict(nNeigh=nNeigh, nCellClust1D=nClust)
    hl = _real_setup(coord, q, prec, Rdir=7.0, neighborList="on", **kw)
    hc = _real_setup(coord, q, prec, Rdir=7.0, neighborList="off", **kw)
    assert "neighbour list" in hl.Print() and "cell rows" in hc.Print()
    Vl, Vc = hl.calcv(q), hc.calcv(q)
    if prec == "fp64":
        m = oracle.compare_v(Vl, Vc)
        assert m["relV"] < 1e-13 and m["reldV"] < 1e-13, m
        return
    ref = _real_setup(coord, q, "fp64", Rdir=7.0, neighborList="on", **kw).calcv(q)
    for V in (Vl, Vc):
        m = oracle.compare_v(ref, V)
        assert m["relV"] < 5e-7 and m["reldV"] < 1.5e-6, m


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_direct_sum_paths_agree_dense_cells(oracle, prec):
    """Dense blobs above the small-set size: cells hold more than 64 charges,
    so the wavefront-per-cell list walk takes its targets in several
    batches; the list sum equals the cell-row sum."""
    coord, q = _blobs(40000, 8, 60.0, 3.0, 43)
    print("case direct paths dense cells", prec, flush=True)
    hl = _real_setup(coord, q, prec, Rdir=5.0, neighborList="on")
    hc = _real_setup(coord, q, prec, Rdir=5.0, neighborList="off")
    assert "neighbour list" in hl.Print() and "cell rows" in hc.Print()
    m = oracle.compare_v(hl.calcv(q), hc.calcv(q))
    assert m["relV"] < (1e-13 if prec == "fp64" else 2e-7) and m["reldV"] < (1e-13 if prec == "fp64" else 1e-6), m


def _blobs(n, nblob, box, sigma, seed):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(0, box, (nblob, 3))
    coord = cen[rng.integers(0, nblob, n)] + rng.normal(0, sigma, (n, 3))
    q = rng.uniform(-1, 1, n)
    return coord, q - q.mean()


@pytest.mark.parametrize("nl", ["on", "off"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["uniform", "blobs", "tiny", "ralpha_aniso", "helmholtz"])
def test_device_setup_matches_host_setup(prec, case, nl):
    """The set-up on the device (cell keys, stable sort, neighbour list or
    cell-row segments) reproduces the host set-up exactly: CalcV results
    are bit-identical."""
    import npme_amd as N
    extra, cx = {}, False
    if case == "uniform":
        coord, q = uniform_box(6000, 35.0)
        kw = dict(funcType="Laplace", Rdir=7.0, BnOrder=6)
    elif case == "blobs":
        coord, q = _blobs(8000, 5, 40.0, 2.0, 21)
        kw = dict(funcType="Laplace", Rdir=6.0, BnOrder=6)
    elif case == "tiny":
        coord, q = uniform_box(5, 6.0)
        kw = dict(funcType="Laplace", Rdir=4.0, BnOrder=4, N1=32, N2=32, N3=32, n1=8, n2=8, n3=8)
    elif case == "ralpha_aniso":
        coord, q = uniform_box(3000, (20.0, 20.0, 60.0))
        kw = dict(funcType="Ralpha", alpha=0.5, Rdir=6.0, BnOrder=6, nDeriv=6,
                  N1=60, N2=60, N3=140, n1=15, n2=15, n3=14)
    else:
        coord, _ = uniform_box(3000, 24.0)
        rng = np.random.default_rng(5)
        q = rng.standard_normal(3000) + 1j * rng.standard_normal(3000)
        kw = dict(funcType="Helmholtz", k0_r=0.7, k0_i=0.05, Rdir=5.0, BnOrder=8)
        cx = True
    print("case device-vs-host set-up", case, prec, flush=True)
    out = {}
    for mode in ("device", "host"):
        k = _kw(calcType="pme", tol=1e-6, precision=prec, setup=mode, neighborList=nl, **kw)
        if cx:
            assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeComplex=q), N.npme_last_error()
            h = N.NPME_Interface_SetUpComplexKernel(k, len(q), coord)
        else:
            assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q), N.npme_last_error()
            h = N.NPME_Interface_SetUpRealKernel(k, len(q), coord)
        assert h is not None, N.npme_last_error()
        assert ("set-up on the " + mode) in h.Print()
        out[mode] = h.calcv(q)
        out[mode + "_direct"] = [l for l in h.Print().splitlines() if "direct sum:" in l]
    assert np.array_equal(out["device"], out["host"])
    # the same structure: list entries or cell-row segments
    assert out["device_direct"] == out["host_direct"]


@pytest.mark.parametrize("setup", ["device", "host"])
def test_mixed_close_pairs(setup):
    """Pairs far closer than the fixed-point unit matters for: the mixed
    direct sum must keep FP32 relative accuracy in r for them (checked
    against the FP64 GPU path on the same grid)."""
    import npme_amd as N
    coord, q = uniform_box(3000, 20.0)
    rng = np.random.default_rng(9)
    seps = [1e-3, 3e-3, 1e-2, 3e-2, 0.1, 0.3]
    for k, r in enumerate(seps):
        u = rng.normal(size=3)
        coord[2 * k + 1] = coord[2 * k] + r * u / np.linalg.norm(u)
    print("case close pairs", setup, flush=True)
    V = {}
    for prec in ("fp64", "mixed"):
        h = _real_setup(coord, q, prec, Rdir=5.0, setup=setup, neighborList="on")
        V[prec] = h.calcv(q)
    d = V["mixed"] - V["fp64"]
    meanV = np.mean(np.abs(V["fp64"][:, 0]))
    planted = np.arange(2 * len(seps))
    rest = np.arange(2 * len(seps), len(q))
    # a planted target's V is dominated by its close pair (up to 1e3 x the
    # mean): its error is measured against its own |V|, the FP32 rounding
    # level of that pair term; the other targets against the mean |V|
    relOwn = np.abs(d[planted, 0]) / np.maximum(np.abs(V["fp64"][planted, 0]), meanV)
    relRest = np.abs(d[rest, 0]) / meanV
    relG = np.linalg.norm(d[planted, 1:], axis=1) / np.linalg.norm(V["fp64"][planted, 1:], axis=1)
    print("planted relV", relOwn.max(), "planted relG", relG.max(), "other max relV", relRest.max())
    # a position error of one fixed-point unit (2e-8 A) would give ~2e-5 in V
    # and ~8e-5 in the gradient at the 1e-3 A pair
    assert relOwn.max() < 2e-6 and relG.max() < 5e-6
    assert relRest.max() < 1e-5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fft_pow8_kernels_match_generic(prec):
    """N = 512: the compile-time radix-8 FFT kernels give the same V as the
    general Stockham kernels (keyword fftKernels generic)."""
    coord, q = uniform_box(3000, 30.0)
    grid = dict(N1=512, N2=512, N3=512, n1=16, n2=16, n3=16, nDeriv=6)
    V = {}
    for mode in ("auto", "generic"):
        h = _real_setup(coord, q, prec, Rdir=6.0, fftKernels=mode, **grid)
        V[mode] = h.calcv(q)
    d = V["auto"] - V["generic"]
    rel = np.sqrt(np.mean(d[:, 0] ** 2)) / np.mean(np.abs(V["generic"][:, 0]))
    relg = np.sqrt(np.mean(d[:, 1:] ** 2)) / np.mean(np.abs(V["generic"][:, 1:]))
    print(prec, rel, relg)
    # FP64: the same transform to rounding; FP32: a different operation
    # order (pruned first stage, loads straight into the butterflies)
    assert rel < (1e-13 if prec == "fp64" else 2e-7) and relg < (1e-13 if prec == "fp64" else 1e-6)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_grid64_matches_oracle(oracle, prec):
    """N = 64 per axis (the compile-time two-stage kernels) against the
    oracle on the same grid."""
    coord, q = uniform_box(2000, 20.0)
    h = _real_setup(coord, q, prec, Rdir=6.0, N1=64, N2=64, N3=64, n1=8, n2=8, n3=8, nDeriv=6)
    V = h.calcv(q)
    k = oracle.kernel(oracle.LAPLACE, 6, 6.0)
    Vo = oracle.calcv(k, (64, 64, 64), 6, coord, q)
    mo = oracle.compare_v(Vo, V)
    print(prec, mo)
    assert mo["relV"] < (1e-11 if prec == "fp64" else 5e-7)
