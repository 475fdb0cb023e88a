"""npme_count_pairs_noalloc (no device allocation, partial counts in a
static device array) against npme_count_pairs, through the C ABI: two calls
in a row on one GPU, the second on a smaller set, each equal to the
allocating entry point's count; and the size limit."""
import numpy as np
import pytest

from conftest import uniform_box

pytestmark = pytest.mark.gpu


def _setup(n, L, seed_shift=0):
    import npme_amd as N
    coord, q = uniform_box(n, L)
    coord = coord + seed_shift
    k = N.NPME_KeywordInput(calcType="pme", funcType="Laplace", Rdir=6.0, BnOrder=6, tol=1e-6)
    assert N.NPME_AddMissingKeywords(k, n, coord, chargeReal=q), N.npme_last_error()
    h = N.NPME_Interface_SetUpRealKernel(k, n, coord)
    assert h is not None, N.npme_last_error()
    return h, coord


def _brute(coord, Rdir):
    d = coord[:, None, :] - coord[None, :, :]
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    return int((r2 < Rdir * Rdir).sum())


def test_count_pairs_noalloc_two_calls():
    print("case noalloc count, 40000 then 3000 charges", flush=True)
    big, _ = _setup(40000, 60.0)
    small, cs = _setup(3000, 25.0, seed_shift=1.5)
    a1 = big.count_pairs_noalloc()
    assert a1 == big.count_pairs() and a1 > 0
    a2 = small.count_pairs_noalloc()
    assert a2 == small.count_pairs() == _brute(cs, 6.0)
    # the large set again after the small one: no partial of the earlier call
    # survives into the result
    assert big.count_pairs_noalloc() == a1
