"""Pins the oracle's direct sum over the reference's pair set
(npo_direct_refpairs, LaplaceOrig's erfc kernel and the GenFunc rFunc user
kernel, whose short-range parts have no cutoff): a brute-force numpy
restatement of the
interaction-list rule of NPME_FindInteractionList_C (src/PartitionBox.cpp:
789-922) over the embedded B cells, summing erfc(beta r)/r over every point
pair of the kept cell pairs (src/PotentialLaplace.cpp:613-700); and the
rule's completeness: every pair closer than Rdir is kept, so the sum exceeds
the cutoff sum by the kept pairs beyond Rdir only."""
import numpy as np
import pytest
from scipy.special import erfc

from conftest import uniform_box


def _cells(coord, Rdir, nNeigh, nClust):
    csB = Rdir / nNeigh
    csA = nClust * csB
    mn = coord.min(axis=0)
    x1 = coord - mn
    a = np.floor(x1 / csA).astype(int)
    b = np.floor((x1 - a * csA) / csB).astype(int)
    return a * nClust + b


def _erfc_over_r(beta, r):
    e = erfc(beta * r) / r
    return e, -(beta * 2 / np.sqrt(np.pi) * np.exp(-(beta * r) ** 2) + e) / r ** 2


def _brute(coord, q, Rdir, beta, nNeigh, nClust, fg=None):
    g = _cells(coord, Rdir, nNeigh, nClust)
    n = len(q)
    d = coord[:, None, :] - coord[None, :, :]
    r2 = (d ** 2).sum(-1)
    Rr = nNeigh * (Rdir / nNeigh)
    close = r2 < Rr * Rr * 1.0001
    keys = [tuple(c) for c in g]
    cells = {}
    for i, k in enumerate(keys):
        cells.setdefault(k, []).append(i)
    contact = {}
    for ka, ia in cells.items():
        for kb, ib in cells.items():
            dn = np.abs(np.subtract(ka, kb))
            if ka == kb:
                contact[ka, kb] = True
            elif dn.max() <= nNeigh and (np.maximum(dn - 1, 0) ** 2).sum() <= nNeigh ** 2:
                contact[ka, kb] = bool(close[np.ix_(ia, ib)].any())
    keep = np.array([[i != j and contact.get((keys[i], keys[j]), False) for j in range(n)]
                     for i in range(n)])
    r = np.sqrt(np.where(keep, r2, 1.0))
    e, f1 = (fg or _erfc_over_r)(beta, r)
    e, f1 = np.where(keep, e, 0.0), np.where(keep, f1, 0.0)
    V = np.zeros((n, 4), np.result_type(e, q))
    V[:, 0] = e @ q
    for c in range(3):
        V[:, 1 + c] = (d[:, :, c] * f1) @ q
    return V, keep, r2


@pytest.mark.parametrize("nNeigh,nClust", [(2, 3), (3, 3), (1, 2)])
def test_refpairs_match_brute_force_rule(oracle, nNeigh, nClust):
    coord, q = uniform_box(600, (14.0, 11.0, 17.0))
    Rdir = 4.0
    beta = oracle.rdir2beta(Rdir, 1e-6)
    k = oracle.kernel_erf(Rdir, beta)
    V = oracle.direct_refpairs(k, coord, q, nNeigh, nClust)
    Vb, keep, r2 = _brute(coord, q, Rdir, beta, nNeigh, nClust)
    np.testing.assert_allclose(V, Vb, rtol=1e-12, atol=1e-12 * np.abs(Vb).max())
    # complete: every pair inside the cutoff is kept; some beyond it are too
    inside = (r2 < Rdir * Rdir) & ~np.eye(len(q), dtype=bool)
    assert keep[inside].all()
    assert (keep & (r2 > Rdir * Rdir)).any()
    # the cutoff sum plus the kept tail beyond Rdir
    Vc = oracle.direct_sr(k, coord, q)
    assert np.abs(V - Vc).max() > 0
    assert np.abs(V - Vc).max() < 1e-6 * np.abs(V).max()


@pytest.mark.parametrize("c", [1.0, 0.6 + 0.8j])
def test_refpairs_rfunc_match_brute_force_rule(oracle, c):
    """the user kernel c r with the split c r erf + c r erfc (rFunc of
    app/userDefineFunc.cpp): GenFunc calls funcSR on every kept pair"""
    coord, q = uniform_box(500, (12.0, 13.0, 10.0))
    cx = isinstance(c, complex)
    if cx:
        q = q + 1j * np.random.default_rng(3).standard_normal(len(q))
    Rdir, beta = 4.0, 0.5
    k = oracle.kernel_rfunc(Rdir, beta, c=c, complex_kernel=cx)

    def fg(b, r):
        return (c * r * erfc(b * r),
                c * (erfc(b * r) / r - 2 / np.sqrt(np.pi) * b * np.exp(-(b * r) ** 2)))

    V = oracle.direct_refpairs(k, coord, q, 2, 3)
    Vb, keep, r2 = _brute(coord, q, Rdir, beta, 2, 3, fg)
    np.testing.assert_allclose(V, Vb, rtol=1e-12, atol=1e-12 * np.abs(Vb).max())
    assert (keep & (r2 > Rdir * Rdir)).any()
