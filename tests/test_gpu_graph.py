"""CalcV replayed from a captured HIP graph (the default for repeated calls
with the same pointers on one set-up) against the same calls launching each
kernel: every replay is bit-identical to the eager result, with other eager
CalcV calls (other pointers, another handle) and synchronisations between
the replays; the graph is dropped by ResetCoord and by new pointers.  Real
and complex kernels, both precisions, neighbour list and cell rows."""
import numpy as np
import pytest

from conftest import uniform_box

pytestmark = pytest.mark.gpu


def _setup(n, L, prec, nl="auto", cx=False):
    import npme_amd as N
    coord, q = uniform_box(n, L)
    kw = dict(calcType="pme", funcType="Helmholtz" if cx else "Laplace", Rdir=6.0, BnOrder=6,
              tol=1e-6, precision=prec, neighborList=nl)
    if cx:
        kw.update(k0_r=0.7, k0_i=0.05, BnOrder=8)
        q = q + 1j * np.roll(q, 3)
    k = N.NPME_KeywordInput(**kw)
    arg = {"chargeComplex": q} if cx else {"chargeReal": q}
    assert N.NPME_AddMissingKeywords(k, n, coord, **arg), N.npme_last_error()
    setup = N.NPME_Interface_SetUpComplexKernel if cx else N.NPME_Interface_SetUpRealKernel
    h = setup(k, n, coord)
    assert h is not None, N.npme_last_error()
    return h, coord, q


@pytest.mark.parametrize("prec,nl,cx", [("mixed", "auto", False), ("fp64", "auto", False),
                                         ("mixed", "off", False), ("mixed", "auto", True)])
def test_graph_replays_equal_eager(prec, nl, cx):
    print("case graph replay", prec, nl, cx, flush=True)
    h, coord, q = _setup(6000, 30.0, prec, nl, cx)
    other, _, q2 = _setup(3000, 24.0, "mixed")
    h.set_calcv_graph(False)
    ref = h.calcv(q)
    ref2 = h.calcv(2.0 * q)
    h.set_calcv_graph(True)
    V = np.zeros_like(ref)
    for rep in range(6):
        # same buffers every call: the second call captures, later ones replay
        assert h.CalcV(V, len(q), q, coord)
        assert np.array_equal(V, ref), "replay %d differs" % rep
        # eager work between the replays: another handle, other charges
        other.calcv(q2)
        if rep == 3:
            assert np.array_equal(h.calcv(2.0 * q), ref2)
    # new geometry: the graph is dropped and the result follows the set-up
    coord2 = coord + 1.25
    assert h.ResetCoord(len(q), coord2)
    h.set_calcv_graph(False)
    ref3 = h.calcv(q)
    h.set_calcv_graph(True)
    for rep in range(3):
        assert h.CalcV(V, len(q), q, coord2)
        assert np.array_equal(V, ref3), "replay %d after ResetCoord differs" % rep
