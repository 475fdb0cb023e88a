"""Two-box (translated-kernel) grids on the GPU, through the C ABI, against
the oracle's restatement (tests/test_oracle_twobox.py pins it) and the exact
sum between the two sets:

  npme_recsumgrid_fulltrans_func vs oracle (FP64) ........ 1e-13 relative
  npme_rstp_calctheta_fullF_reversesign vs oracle ......... 1e-12 relative,
                                                          every reversal
  box pair fp64 vs the oracle's box pair .................. 1e-10 (V, grad)
  box pair vs the exact two-set sum ....................... relV < 1e-6 at
                                                          48^3, del 10 (the
                                                          grid's own error
                                                          there is 2.4e-7)
  reversal = the mirrored source box, same kernel grid
  graph replay and eager launches give identical V"""
import numpy as np
import pytest

from conftest import uniform_box

pytestmark = pytest.mark.gpu

NG, BN, DEL = (48, 48, 48), 8, 10.0


def _kw(funcType="Laplace", prec="fp64", **extra):
    import npme_amd as N
    k = N.NPME_KeywordInput(funcType=funcType, BnOrder=BN, precision=prec, N1=NG[0], N2=NG[1],
                            N3=NG[2], del1=DEL, del2=DEL, del3=DEL, Rdir=6.0, **extra)
    return k


def _boxes(nT=150, nS=200, side=8.0, shift=(28.0, 6.0, -4.0), complex_q=False):
    cT, _ = uniform_box(nT, side, seed_c=3001, seed_q=3002)
    cS, qS = uniform_box(nS, side, seed_c=3003, seed_q=3004)
    cS = cS + np.asarray(shift, float)
    if complex_q:
        qS = qS + 1j * np.random.default_rng(7).standard_normal(nS)
    return cT, cS, qS


def _mirror(cT, cS, rev, oracle):
    _, ct = oracle.box_dim_center(cT)
    _, cs = oracle.box_dim_center(cS)
    out = cS.copy()
    for d in range(3):
        if rev[d]:
            out[:, d] = cS[:, d] - 2 * (cs[d] - ct[d])
    return out


def _exact(oracle, k, cT, cS, qS):
    nT = len(cT)
    q = np.concatenate([np.zeros(nT, qS.dtype), qS])
    return oracle.exact(k, np.vstack([cT, cS]), q)[:nT]


@pytest.mark.parametrize("funcType,extra", [("Laplace", {}), ("Ralpha", {"alpha": 0.5}),
                                            ("Helmholtz", {"k0_r": 0.5, "k0_i": 0.05})])
def test_fulltrans_func_matches_oracle(oracle, funcType, extra):
    import npme_amd as N
    print("case fulltrans", funcType, flush=True)
    Ng, a, dl, X, R0 = (20, 16, 24), (4.0, 4.0, 4.0), (5.0, 5.0, 5.0), (6.0, 5.0, 7.0), (26.0, -3.0, 1.5)
    kw = N.NPME_KeywordInput(funcType=funcType, **extra)
    f = N.NPME_RecSumGrid_FullTrans_Func(Ng, a, dl, X, kw, R0)
    if funcType == "Laplace":
        k = oracle.kernel(oracle.LAPLACE, 6, 6.0)
    elif funcType == "Ralpha":
        k = oracle.kernel(oracle.RALPHA, 6, 6.0, alpha=0.5)
    else:
        k = oracle.kernel(oracle.HELMHOLTZ, 6, 6.0, k0=0.5 + 0.05j)
    fo = oracle.fgrid_fulltrans(k, Ng, a, dl, X, R0, part=oracle.FULL)
    np.testing.assert_allclose(f, fo, rtol=1e-13, atol=1e-13 * np.abs(fo).max())


def test_calctheta_fullF_reversesign_matches_oracle(oracle):
    import npme_amd as N
    rng = np.random.default_rng(31)
    Ng, Bn = (24, 20, 28), 6
    M = [n // 2 for n in Ng]
    Q = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    F = rng.standard_normal(Ng) + 1j * rng.standard_normal(Ng)
    lam = [oracle.lambda2(n, Bn) for n in Ng]
    for rev in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 1, 1)]:
        print("case theta reverse", rev, flush=True)
        T = N.NPME_RSTP_CalcTheta_FullF_ReverseSign(F, Q, lam, rev)
        To = oracle.theta_fullF_reverse(Ng, Bn, F, Q, rev)
        np.testing.assert_allclose(T, To, rtol=0, atol=1e-12 * np.abs(To).max(), err_msg=str(rev))


@pytest.mark.parametrize("rev", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)])
def test_boxpair_laplace_fp64(oracle, rev):
    import npme_amd as N
    print("case boxpair laplace fp64", rev, flush=True)
    cT, cS, qS = _boxes()
    bp = N.NPME_BoxPair(_kw(), cT, cS)
    R0, rMin = bp.separation()
    _, ct = oracle.box_dim_center(cT)
    _, cs = oracle.box_dim_center(cS)
    np.testing.assert_allclose(R0, ct - cs, rtol=0, atol=1e-12)
    assert rMin > 5.0
    V = bp.calcv(qS, rev)
    k = oracle.kernel(oracle.LAPLACE, 6, 6.0)
    Vo = oracle.boxpair(k, NG, BN, cT, cS, qS, dl=np.full(3, DEL), rev=rev).real
    mo = oracle.compare_v(Vo, V)
    assert mo["relV"] < 1e-10 and mo["reldV"] < 1e-10, mo
    me = oracle.compare_v(_exact(oracle, k, cT, _mirror(cT, cS, rev, oracle), qS), V)
    print("vs oracle", mo["relV"], "vs exact", me["relV"], me["reldV"])
    assert me["relV"] < 1e-6 and me["reldV"] < 5e-6, me


def test_boxpair_laplace_mixed(oracle):
    import npme_amd as N
    print("case boxpair laplace mixed", flush=True)
    cT, cS, qS = _boxes()
    bp = N.NPME_BoxPair(_kw(prec="mixed"), cT, cS)
    k = oracle.kernel(oracle.LAPLACE, 6, 6.0)
    for rev in [(0, 0, 0), (0, 1, 1)]:
        V = bp.calcv(qS, rev)
        Vo = oracle.boxpair(k, NG, BN, cT, cS, qS, dl=np.full(3, DEL), rev=rev).real
        mo = oracle.compare_v(Vo, V)
        me = oracle.compare_v(_exact(oracle, k, cT, _mirror(cT, cS, rev, oracle), qS), V)
        print(rev, "vs oracle", mo["relV"], mo["reldV"], "vs exact", me["relV"], me["reldV"])
        assert mo["relV"] < 1e-6 and mo["reldV"] < 5e-6, mo


@pytest.mark.parametrize("rev", [(0, 0, 0), (0, 0, 1), (1, 0, 1)])
def test_boxpair_helmholtz_fp64(oracle, rev):
    import npme_amd as N
    print("case boxpair helmholtz", rev, flush=True)
    cT, cS, qS = _boxes(complex_q=True)
    bp = N.NPME_BoxPair(_kw("Helmholtz", k0_r=0.5, k0_i=0.05), cT, cS)
    V = bp.calcv(qS, rev)
    k = oracle.kernel(oracle.HELMHOLTZ, 6, 6.0, k0=0.5 + 0.05j)
    Vo = oracle.boxpair(k, NG, BN, cT, cS, qS, dl=np.full(3, DEL), rev=rev)
    mo = oracle.compare_v(Vo, V)
    assert mo["relV"] < 1e-10 and mo["reldV"] < 1e-10, mo
    me = oracle.compare_v(_exact(oracle, k, cT, _mirror(cT, cS, rev, oracle), qS), V)
    print("vs exact", me["relV"], me["reldV"])
    assert me["relV"] < 1e-5, me


def test_boxpair_ralpha_and_user_kernel(oracle):
    """r^0.5 built in, and f = r as a user NPME_KfuncReal"""
    import npme_amd as N
    cT, cS, qS = _boxes()
    print("case boxpair ralpha", flush=True)
    V = N.NPME_BoxPair(_kw("Ralpha", alpha=0.5), cT, cS).calcv(qS)
    k = oracle.kernel(oracle.RALPHA, 6, 6.0, alpha=0.5)
    Vo = oracle.boxpair(k, NG, BN, cT, cS, qS, dl=np.full(3, DEL)).real
    assert oracle.compare_v(Vo, V)["relV"] < 1e-10

    class RF(N.NPME_KfuncReal):
        def Calc(self, n, *a):
            x, y, z = a[0], a[1], a[2]
            x[:] = np.sqrt(x * x + y * y + z * z)

    print("case boxpair user rFunc", flush=True)
    kw = N.NPME_KeywordInput(BnOrder=BN, precision="fp64", N1=NG[0], N2=NG[1], N3=NG[2],
                             del1=DEL, del2=DEL, del3=DEL, Rdir=6.0)
    V = N.NPME_BoxPair(kw, cT, cS, func=RF()).calcv(qS)
    ko = oracle.kernel_rfunc(6.0, 0.5)
    Vo = oracle.boxpair(ko, NG, BN, cT, cS, qS, dl=np.full(3, DEL)).real
    mo = oracle.compare_v(Vo, V)
    assert mo["relV"] < 1e-10 and mo["reldV"] < 1e-10, mo
    me = oracle.compare_v(_exact(oracle, ko, cT, cS, qS), V)
    assert me["relV"] < 1e-6, me


def test_boxpair_graph_replay_and_reversal_changes(oracle):
    """calls replay a captured graph from the second on; changing the
    reversal re-captures; every result equals the eager launch's"""
    import npme_amd as N
    cT, cS, qS = _boxes()
    print("case boxpair graph", flush=True)
    bp = N.NPME_BoxPair(_kw(prec="mixed"), cT, cS)
    seq = [(0, 0, 0), (0, 0, 0), (0, 0, 0), (1, 0, 1), (1, 0, 1), (1, 0, 1), (0, 0, 0)]
    got = [bp.calcv(qS, r) for r in seq]
    bp.set_calcv_graph(False)
    for r, V in zip(seq, got):
        np.testing.assert_array_equal(V, bp.calcv(qS, r))


def test_boxpair_refusals():
    import npme_amd as N
    cT, cS, qS = _boxes(shift=(9.0, 0.0, 0.0))
    print("case boxpair refusals", flush=True)
    with pytest.raises(N.NPMEError, match="separation"):
        N.NPME_BoxPair(_kw(), cT, cS)
    cT, cS, qS = _boxes()
    bp = N.NPME_BoxPair(_kw(), cT, cS)
    V = np.zeros((len(cT) + len(cS), 4))
    assert N.lib().npme_calcv_real(bp._h, V.ctypes.data, len(cT) + len(cS), qS.ctypes.data, None) != 0
    assert "box-pair" in N.npme_last_error()
    with pytest.raises((N.NPMEError, ValueError)):
        bp.calcv(qS[:-1])
