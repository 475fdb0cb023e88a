"""Error paths on the GPU: a set-up that runs out of device memory part way
(the library's allocation-failure hook, npme_debug_fail_alloc, fails the
k-th allocation of the calling thread) returns NPME_ERR_NOMEM, leaves no
usable state behind, and does not disturb another handle working at the
same time: blocks that a failed set-up returns to the shared pool must not
still be in use by the kernels it queued, or the next handle that receives
them would see its buffers overwritten."""
import threading

import numpy as np
import pytest

from conftest import uniform_box

pytestmark = pytest.mark.gpu


def _kw(coord, q, **extra):
    import npme_amd as N
    k = N.NPME_KeywordInput(calcType="pme", funcType="Laplace", Rdir=6.0, BnOrder=6, tol=1e-6,
                            **extra)
    assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q), N.npme_last_error()
    return k


@pytest.mark.parametrize("mode", ["device", "host"])
def test_setup_out_of_memory_part_way(mode):
    import npme_amd as N
    coordA, qA = uniform_box(30000, 40.0, 51, 52)
    coordB, qB = uniform_box(20000, 35.0, 53, 54)
    kA, kB = _kw(coordA, qA, setup=mode), _kw(coordB, qB)
    hB = N.NPME_Interface_SetUpRealKernel(kB, len(qB), coordB)
    assert hB is not None, N.npme_last_error()
    Vref = hB.calcv(qB)
    # another thread evaluates B over and over while set-ups of A fail
    stop = threading.Event()
    bad = []

    def loop():
        while not stop.is_set():
            V = hB.calcv(qB)
            if not np.array_equal(V, Vref):
                bad.append(float(np.abs(V - Vref).max()))

    th = threading.Thread(target=loop)
    th.start()
    failed = 0
    try:
        for k in range(0, 200):
            print("case fail allocation", mode, k, flush=True)
            N.lib().npme_debug_fail_alloc(k)
            hA = N.NPME_Interface_SetUpRealKernel(kA, len(qA), coordA)
            N.lib().npme_debug_fail_alloc(-1)
            if hA is not None:
                break  # k is beyond the set-up's allocations
            failed += 1
            assert "alloc" in N.npme_last_error(), N.npme_last_error()
            # a fresh set-up of B's geometry receives the blocks A returned
            hC = N.NPME_Interface_SetUpRealKernel(kB, len(qB), coordB)
            assert hC is not None, N.npme_last_error()
            assert np.array_equal(hC.calcv(qB), Vref)
            del hC
    finally:
        N.lib().npme_debug_fail_alloc(-1)
        stop.set()
        th.join()
    assert failed >= 10, failed
    assert hA is not None
    assert not bad, bad[:5]
    # the handle that finally set up works
    assert np.all(np.isfinite(hA.calcv(qA)))


def test_reset_coord_out_of_memory_leaves_no_state():
    import npme_amd as N
    coord, q = uniform_box(15000, 30.0, 55, 56)
    h = N.NPME_Interface_SetUpRealKernel(_kw(coord, q), len(q), coord)
    assert h is not None
    V0 = h.calcv(q)
    N.lib().npme_debug_fail_alloc(5)
    ok = h.ResetCoord(len(q), coord)
    N.lib().npme_debug_fail_alloc(-1)
    assert not ok
    V = np.zeros((len(q), 4))
    assert not h.CalcV(V, len(q), q, coord) and "no valid set-up" in N.npme_last_error()
    assert h.ResetCoord(len(q), coord)
    assert np.array_equal(h.calcv(q), V0)
