"""GPU tests of the serving loop (npme_server_*): requests of mixed kernels,
sizes and tolerances run concurrently give the same V as each request run
alone (bit-identical: the kernels are deterministic and requests share no
state), meet their requested tolerance against the exact sum, and respect
the device-memory admission budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = [("Laplace", {}), ("Ralpha", {"alpha": 0.5}), ("Ralpha", {"alpha": 1.5}),
         ("Helmholtz", {"k0": 0.3})]


def _requests(sizes, seed=31):
    rng = np.random.default_rng(seed)
    reqs = []
    for i, n in enumerate(sizes):
        ft, kw = KINDS[i % len(KINDS)]
        tol = 1e-6 if (i // len(KINDS)) % 2 == 0 else 1e-4
        L = (n / 0.1) ** (1.0 / 3.0)
        coord = rng.uniform(0.0, L, (n, 3))
        if ft == "Helmholtz":
            q = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        else:
            q = rng.uniform(-1.0, 1.0, n)
            q -= q.mean()
        reqs.append((ft, coord, q, tol, kw))
    return reqs


def _run(server, reqs):
    ticks = [server.submit(ft, c, q, tol=tol, Rdir=8.0, **kw) for ft, c, q, tol, kw in reqs]
    return [server.wait(t) for t in ticks]


def test_server_mix_equals_alone():
    import npme_amd as N
    reqs = _requests([500, 900, 2000, 3500, 6000, 12000, 700, 2500, 5000, 1500, 8000, 30000])
    print("case server mix: %d requests on 4 slots" % len(reqs), flush=True)
    S = N.NPME_Server(nSlots=4)
    mixed = _run(S, reqs)
    S.close()
    # the slots did run concurrently
    ivals = sorted((inf["tStart"], inf["tEnd"]) for _, inf in mixed)
    overlap = any(ivals[i + 1][0] < ivals[i][1] for i in range(len(ivals) - 1))
    assert overlap
    print("case server alone", flush=True)
    S1 = N.NPME_Server(nSlots=1)
    for (ft, c, q, tol, kw), (Vm, _) in zip(reqs, mixed):
        (Va, _), = _run(S1, [(ft, c, q, tol, kw)])
        assert np.array_equal(Va, Vm), ft
    S1.close()


def test_server_results_meet_tolerance(oracle):
    import npme_amd as N
    reqs = _requests([1200, 1500, 1800, 2000, 1300, 1600, 1900, 2200])
    print("case server tolerance", flush=True)
    S = N.NPME_Server(nSlots=4)
    out = _run(S, reqs)
    S.close()
    for (ft, c, q, tol, kw), (V, inf) in zip(reqs, out):
        if ft == "Laplace":
            k = oracle.kernel(oracle.LAPLACE, 6, 8.0)
        elif ft == "Ralpha":
            k = oracle.kernel(oracle.RALPHA, 6, 8.0, alpha=kw["alpha"])
        else:
            k = oracle.kernel(oracle.HELMHOLTZ, 8, 8.0, k0=kw["k0"])
        m = oracle.compare_v(oracle.exact(k, c, q), V)
        print(ft, kw, tol, inf["grid"], m["relV"])
        assert m["relV"] <= 2.0 * tol


def test_server_memory_budget_serialises():
    """With a budget below one request's estimate, requests run one at a
    time (a request always runs when nothing else does)."""
    import npme_amd as N
    reqs = _requests([800, 1000, 1200, 1400])
    print("case server budget", flush=True)
    S = N.NPME_Server(nSlots=4, memGB=0.01)
    out = _run(S, reqs)
    S.close()
    ivals = sorted((inf["tStart"], inf["tEnd"]) for _, inf in out)
    assert all(ivals[i + 1][0] >= ivals[i][1] for i in range(len(ivals) - 1))


def test_server_rejects_bad_requests():
    import npme_amd as N
    S = N.NPME_Server(nSlots=2)
    c = np.random.default_rng(1).uniform(0, 10, (100, 3))
    with pytest.raises(N.NPMEError):
        S.submit("Yukawa", c, np.ones(100))
    with pytest.raises(N.NPMEError):
        S.submit("Laplace", c[:1], np.ones(1))
    # a request that fails at set-up reports its error at wait()
    t = S.submit("Laplace", c, np.ones(100), tol=1e-6, Rdir=8.0, BnOrder=7)
    with pytest.raises(N.NPMEError):
        S.wait(t)
    S.close()


def test_server_pinned_request_buffers():
    """Requests built in page-locked buffers (pinned_empty) give the same V
    as from ordinary arrays, including views into a larger pinned buffer
    reused for a second request, and report their phase times."""
    import npme_amd as N
    reqs = _requests([3000, 20000, 5000, 40000], seed=37)
    print("case server pinned buffers: %d requests" % len(reqs), flush=True)
    S = N.NPME_Server(nSlots=2)
    try:
        ref = _run(S, reqs)
        nmax = max(len(q) for _, _, q, _, _ in reqs)
        cbuf, qbuf = N.pinned_empty((nmax, 3)), N.pinned_empty(nmax, np.complex128)
        vbuf = N.pinned_empty(8 * nmax)
        for (ft, c, q, tol, kw), (Vr, _) in zip(reqs, ref):
            n = len(q)
            coord = cbuf[:n]
            coord[:] = c
            if ft == "Helmholtz":
                qp = qbuf[:n]
                V = vbuf[:8 * n].view(np.complex128).reshape(n, 4)
            else:
                qp = qbuf.view(np.float64)[:n]
                V = vbuf[:4 * n].reshape(n, 4)
            qp[:] = q
            V.fill(np.nan)
            t = S.submit(ft, coord, qp, tol=tol, Rdir=8.0, V=V, **kw)
            Vp, info = S.wait(t)
            assert Vp is V or np.shares_memory(Vp, V)
            assert np.array_equal(V, Vr), ft
            assert set(info["ms"]) == {"keywords", "setup", "copy_in", "calcv", "copy_out"}
            assert all(v >= 0 for v in info["ms"].values())
    finally:
        S.close()


def test_server_request_equals_library_call():
    """A served request (positions staged on the device, grid chosen from
    their device bounds) gives bit for bit the V of the library's own
    NPME_AddMissingKeywords + SetUp + CalcV with the same keywords."""
    import npme_amd as N
    rng = np.random.default_rng(41)
    S = N.NPME_Server(nSlots=2)
    try:
        for n, ft, kw, tol in [(7000, "Laplace", {}, 1e-6), (15000, "Helmholtz", {"k0": 0.3}, 1e-4)]:
            print("case server vs library", ft, n, flush=True)
            L = (n / 0.1) ** (1.0 / 3.0)
            coord = rng.uniform(0.0, L, (n, 3))
            if ft == "Helmholtz":
                q = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            else:
                q = rng.uniform(-1.0, 1.0, n)
            V, info = S.wait(S.submit(ft, coord, q, tol=tol, Rdir=8.0, **kw))
            k = N.NPME_KeywordInput(calcType="pme", funcType=ft, Rdir=8.0, tol=tol, precision="mixed")
            if ft == "Helmholtz":
                k.AddKeyword("k0_r", 0.3)
                k.AddKeyword("k0_i", 0.0)
                assert N.NPME_AddMissingKeywords(k, n, coord, chargeComplex=q)
                h = N.NPME_Interface_SetUpComplexKernel(k, n, coord)
            else:
                assert N.NPME_AddMissingKeywords(k, n, coord, chargeReal=q)
                h = N.NPME_Interface_SetUpRealKernel(k, n, coord)
            assert h is not None, N.npme_last_error()
            assert tuple(info["grid"]) == (h.Get_N1(), h.Get_N2(), h.Get_N3())
            assert np.array_equal(V, h.calcv(q))
    finally:
        S.close()


def test_server_reservation_follows_footprint():
    """Admission by each request's own footprint: once its grid is chosen a
    request reserves the grids, per-charge arrays and neighbour list at its
    density, and after set-up what the set-up holds.  The reservation is an
    upper estimate of the resident bytes (within a factor 4 for uniform
    sets); a dense clustered request reserves more per charge than a
    uniform one of the same size; under a budget smaller than two requests
    every request still completes with the result it has alone."""
    import npme_amd as N
    rng = np.random.default_rng(41)
    reqs = _requests([4000, 20000, 60000, 9000])
    # the same charge count packed 8x denser (a bigger neighbour list)
    n = 20000
    L = (n / 0.8) ** (1.0 / 3.0)
    dense = ("Laplace", rng.uniform(0.0, L, (n, 3)), rng.uniform(-1, 1, n), 1e-6, {})
    print("case server footprint", flush=True)
    S = N.NPME_Server(nSlots=2)
    out = _run(S, reqs + [dense])
    S.close()
    for (ft, c, q, tol, kw), (V, inf) in zip(reqs + [dense], out):
        print(ft, len(q), "reserved %.4f GB resident %.4f GB" % (inf["reservedGB"], inf["deviceGB"]))
        assert inf["deviceGB"] <= inf["reservedGB"] * 1.25 + 0.01
        assert inf["reservedGB"] <= 4 * inf["deviceGB"] + 0.02
    per = [inf["reservedGB"] / len(q) for (_, _, q, _, _), (_, inf) in zip(reqs + [dense], out)]
    assert per[-1] > 2 * per[1]  # dense 20k against uniform 20k (same kernel and tol)
    # a budget below two requests' footprints: requests wait for memory,
    # none deadlocks, and every result equals the unconstrained one
    tight = max(inf["reservedGB"] for _, inf in out) * 1.5
    S2 = N.NPME_Server(nSlots=3, memGB=tight)
    out2 = _run(S2, reqs + [dense])
    S2.close()
    for (V, _), (V2, _) in zip(out, out2):
        assert np.array_equal(V, V2)


def test_server_batches_equal_alone():
    """Small requests queued together run as batches (their set-ups in turn,
    their CalcVs as one graph launch): with one slot and a burst of 20 small
    requests of every kernel, some run in batches of several, and every
    result is bit-identical to the same request run with batching off."""
    import npme_amd as N
    sizes = [600, 900, 1500, 2500, 4000, 700, 1200, 3000, 800, 2000,
             1000, 1800, 650, 2200, 3500, 900, 1100, 2800, 750, 1600]
    reqs = _requests(sizes, seed=53)
    print("case server batches: %d requests on 1 slot" % len(reqs), flush=True)
    S = N.NPME_Server(nSlots=1, maxBatch=8, batchMaxN=32768)
    out = _run(S, reqs)
    S.close()
    sizes_seen = [inf["batch"] for _, inf in out]
    print("batch sizes", sizes_seen)
    assert max(sizes_seen) > 1
    S1 = N.NPME_Server(nSlots=1, maxBatch=1)
    alone = _run(S1, reqs)
    S1.close()
    assert all(inf["batch"] == 1 for _, inf in alone)
    for (V, _), (Va, _), (ft, _, _, _, _) in zip(out, alone, reqs):
        assert np.array_equal(V, Va), ft


def test_server_batch_raise_over_budget():
    """A batch whose members' footprints exceed the budget once their grids
    are known runs to the end: the members are set up in turn on one slot,
    so a member waiting to raise its reservation is the only work waiting
    and goes on (as a request that runs alone does).  A large request holds
    the one slot while two dense 20k requests (list-heavy footprints above
    their provisional estimates) queue behind it and then run as a batch of
    two under a 0.25 GB budget; their results equal the same requests run
    alone."""
    import npme_amd as N
    rng = np.random.default_rng(71)
    n0 = 200000
    L0 = (n0 / 0.1) ** (1.0 / 3.0)
    q0 = rng.uniform(-1.0, 1.0, n0)
    blocker = ("Laplace", rng.uniform(0.0, L0, (n0, 3)), q0 - q0.mean(), 1e-4, {})
    dense = []
    for _ in range(2):
        n = 20000
        L = (n / 1.0) ** (1.0 / 3.0)
        q = rng.uniform(-1.0, 1.0, n)
        dense.append(("Laplace", rng.uniform(0.0, L, (n, 3)), q - q.mean(), 1e-6, {}))
    print("case server batch raise over a 0.25 GB budget", flush=True)
    S = N.NPME_Server(nSlots=1, memGB=0.25, maxBatch=8, batchMaxN=32768)
    out = _run(S, [blocker] + dense)
    S.close()
    print("batch sizes", [inf["batch"] for _, inf in out],
          "reserved GB", [round(inf["reservedGB"], 3) for _, inf in out])
    assert [inf["batch"] for _, inf in out[1:]] == [2, 2]
    assert all(inf["reservedGB"] > 0.125 for _, inf in out[1:])  # raised past half the budget
    S1 = N.NPME_Server(nSlots=1, maxBatch=1)
    alone = _run(S1, dense)
    S1.close()
    for (V, _), (Va, _) in zip(out[1:], alone):
        assert np.array_equal(V, Va)
