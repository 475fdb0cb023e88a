#!/usr/bin/env python3
"""Benchmark: charges/sec of the NPME 1/r DM potential evaluation (CalcV) at
N = 1,000,000 on one MI355X.

Workload: 1M point charges uniform in a 215.443^3 cube (density 0.1), q ~
U[-1,1] shifted to zero net charge (SplitMix64 seeds 1001/1002), Laplace
kernel with derivative-match splitting, B-spline order 6, Rdir = 8, grid
chosen by the RSEM estimate for tol = 1e-6 (512^3, Nder = 6).  A step is one
CalcV of the mixed-precision served form with the charges already resident
in HBM; set-up (cell lists, kernel grid) happens once before timing.

Also reported: the plain FP64 GPU path, the relative RMS error against the
exact FP64 sum on 4,000 random targets, per-stage device times, the
roofline of the dominant kernel and the CPU baseline (the FP64 restatement
of the reference algorithm on the host cores).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8000.0          # GB/s, MI355X HBM3E
FP32_PEAK = 157.3          # TFLOP/s vector FP32
STAGES = ["gather", "direct", "spread", "fftA_r2c_z", "fftB_y", "fftC_x_conv", "fftD_y",
          "fftE_c2r_z", "interp_combine"]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=1000000)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--no-accuracy", action="store_true")
    p.add_argument("--stream-requests", type=int, default=5000,
                   help="requests of the mixed-size serving stream (0 skips it)")
    p.add_argument("--stream-rate", type=float, default=200.0, help="Poisson arrival rate, 1/s")
    p.add_argument("--slots", type=int, default=4, help="concurrent request slots of the server")
    p.add_argument("--pageable-requests", action="store_true",
                   help="serving client builds requests in pageable instead of page-locked memory")
    p.add_argument("--serving-only", action="store_true",
                   help="run only the serving stream and print its latency JSON (the child "
                        "process of a full run)")
    p.add_argument("--crash-log", default="", help=argparse.SUPPRESS)
    p.add_argument("--stages-only", action="store_true",
                   help="set up the N=1M workload and run --steps CalcV calls with the stages "
                        "serialised on one stream and timed by HIP events (the timing the "
                        "roofline uses); prints the stage times, for a profiler's kernel trace")
    p.add_argument("--calib-copy", action="store_true",
                   help="after the steps, copy a 256 MiB device buffer 3 times (a plain copy "
                        "kernel of known size, to calibrate profiler byte counters)")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(coord, q, prec):
    import npme_amd as N
    k = N.NPME_KeywordInput()
    for a, b in dict(calcType="pme", funcType="Laplace", EwaldSplit="DerivMatch", Rdir=8.0,
                     BnOrder=6, tol=1e-6, precision=prec).items():
        k.AddKeyword(a, b)
    if not N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q):
        raise RuntimeError(N.npme_last_error())
    t0 = time.perf_counter()
    h = N.NPME_Interface_SetUpRealKernel(k, len(q), coord)
    if h is None:
        raise RuntimeError(N.npme_last_error())
    return h, time.perf_counter() - t0


def settle(h, dV, dq, seconds=0.25):
    """Untimed CalcV calls for about `seconds`, before any timing: the GPU's
    clocks and caches reach their steady state (a few warm-up steps of ~1 ms
    each leave the first timed block several percent slow)."""
    import torch
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            h.calcv_device(dV.data_ptr(), dq.data_ptr())
        torch.cuda.synchronize()


def time_steps(h, dV, dq, steps, warmup):
    import torch
    for _ in range(warmup):
        h.calcv_device(dV.data_ptr(), dq.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        h.calcv_device(dV.data_ptr(), dq.data_ptr())
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def stage_profile(h, dV, dq, reps=5):
    h.set_profiling(True)
    acc = None
    for _ in range(reps):
        h.calcv_device(dV.data_ptr(), dq.data_ptr())
        st = h.stage_times()
        acc = st if acc is None else acc + st
    h.set_profiling(False)
    return acc / reps


def accuracy(h, q, V, n):
    idx = np.sort(np.random.default_rng(1003).choice(n, min(4000, n), replace=False))
    Ve = h.calcv_exact_subset(q, idx)
    d = V[idx] - Ve
    magV = np.mean(np.abs(Ve[:, 0]))
    return {"targets": len(idx),
            "relV": float(np.sqrt(np.mean(d[:, 0] ** 2)) / magV),
            "reldV": float(np.sqrt(np.mean(d[:, 1:] ** 2)) / np.mean(np.abs(Ve[:, :3]))),
            "max_absV_over_meanV": float(np.max(np.abs(d[:, 0])) / magV)}


class StallWatch:
    """Calls on_stall once, from a thread of its own, when mark() has not
    been called for stall_s seconds between start() and stop()."""

    def __init__(self, stall_s, on_stall, poll_s=5.0):
        import threading
        self.stall_s, self.on_stall, self.poll_s = stall_s, on_stall, poll_s
        self.last = time.monotonic()
        self.count = 0
        self._stop = threading.Event()
        self._th = threading.Thread(target=self._run, daemon=True)

    def mark(self):
        self.last = time.monotonic()
        self.count += 1

    def start(self):
        self.last = time.monotonic()
        self._th.start()
        return self

    def stop(self):
        self._stop.set()
        self._th.join()

    def _run(self):
        while not self._stop.wait(self.poll_s):
            if time.monotonic() - self.last > self.stall_s:
                self.on_stall()
                return


def serving_stream(nreq, rate, slots, seed=2001, pinned=True, diag=None, stall_s=60.0):
    """Mixed-size request stream through the serving loop: Poisson arrivals
    at `rate`; N log-uniform in [500, 1e6]; positions uniform in a cube at
    density 0.1; charges U[-1,1] (zero net) or complex N(0,1); kernel 70%
    1/r, 20% r^alpha (alpha 0.5 or 1.5), 10% Helmholtz k0 = 0.3; tol 1e-4 or
    1e-6 with equal probability; Rdir 8.  Latency = completion minus the
    scheduled arrival, so a late submission counts against the server.

    The client builds each request in buffers from a pool of `npool`
    max-size buffer sets, reused as requests complete: page-locked host
    memory (N.pinned_empty) by default, as a serving front end registers
    its request buffers, or ordinary pageable memory.

    A stream in which no request completes for stall_s seconds is ended:
    the tracebacks of every thread go to diag (a file) and the process
    exits with status 3, so a stalled stream is reported, not waited for."""
    import queue
    import threading
    import npme_amd as N
    rng = np.random.default_rng(seed)
    nmax = 1000000
    base = rng.uniform(0.0, 1.0, (nmax, 3))
    baseq = rng.uniform(-1.0, 1.0, nmax)
    baseqc = rng.standard_normal(nmax) + 1j * rng.standard_normal(nmax)
    sizes = np.exp(rng.uniform(np.log(500), np.log(nmax), nreq)).astype(np.int64)
    u = rng.uniform(0, 1, nreq)
    kinds = np.where(u < 0.7, 0, np.where(u < 0.9, 1, 2))
    alphas = np.where(rng.uniform(0, 1, nreq) < 0.5, 0.5, 1.5)
    tols = np.where(rng.uniform(0, 1, nreq) < 0.5, 1e-4, 1e-6)
    arrive = np.cumsum(rng.exponential(1.0 / rate, nreq)) + 0.5
    # buffer sets (coordinates, real charges, V as 8 doubles per charge),
    # allocated before the server's clock starts
    npool = 24
    alloc = N.pinned_empty if pinned else (lambda shape, dtype=np.float64: np.zeros(shape, dtype))
    pool = [(alloc((nmax, 3)), alloc(nmax), alloc(8 * nmax)) for _ in range(npool)]
    for c, qr, v in pool:  # pages resident before the stream
        c.fill(0.0)
        qr.fill(0.0)
        v.fill(0.0)
    qcomplex = alloc(nmax, np.complex128)
    qcomplex[:] = baseqc
    free = queue.Queue()
    for k in range(npool):
        free.put(k)
    slot_of = [0] * nreq
    # the client's garbage collector is paused for the stream, as a latency-
    # sensitive front end does: a full collection stalls the submitting thread
    import gc
    gc.collect()
    gc.disable()
    S = N.NPME_Server(nSlots=slots)
    tickets = [None] * nreq
    infos = [None] * nreq
    errors = []
    ready = threading.Semaphore(0)

    def consume():
        for i in range(nreq):
            ready.acquire()
            try:
                _, infos[i] = S.wait(tickets[i])
            except Exception as e:  # recorded, reported in the JSON
                errors.append(str(e))
            free.put(slot_of[i])
            watch.mark()
            if (i + 1) % 1000 == 0:
                log("stream: %d/%d requests done" % (i + 1, nreq))

    def stalled():
        import faulthandler
        msg = "serving stream: no request completed for %.0f s (%d of %d done)" % (
            stall_s, watch.count, nreq)
        log(msg)
        if diag is not None:
            diag.write(msg + "\n")
            diag.flush()
            faulthandler.dump_traceback(file=diag, all_threads=True)
        os._exit(3)

    watch = StallWatch(stall_s, stalled).start()
    th = threading.Thread(target=consume)
    th.start()

    # the client prepares each request's arrays ahead of its arrival in a
    # thread of its own, so building a large request does not delay the
    # submission of the next ones
    prep = queue.Queue(maxsize=16)

    def produce():
        for i in range(nreq):
            n = int(sizes[i])
            L = (n / 0.1) ** (1.0 / 3.0)
            k = free.get()
            slot_of[i] = k
            c, qr, v = pool[k]
            coord = c[:n]
            np.multiply(base[:n], L, out=coord)
            if kinds[i] == 2:
                q = qcomplex[:n]
                V = v[:8 * n].view(np.complex128).reshape(n, 4)
            else:
                q = qr[:n]
                np.subtract(baseq[:n], baseq[:n].mean(), out=q)
                V = v[:4 * n].reshape(n, 4)
            prep.put((coord, q, V))

    pt = threading.Thread(target=produce)
    pt.start()
    lags = np.zeros(nreq)
    for i in range(nreq):
        coord, q, V = prep.get()
        t = S.clock()
        if t < arrive[i]:
            time.sleep(arrive[i] - t)
        lags[i] = S.clock() - arrive[i]
        if kinds[i] == 2:
            tickets[i] = S.submit("Helmholtz", coord, q, tol=tols[i], Rdir=8.0, k0=0.3,
                                  tArrival=arrive[i], V=V)
        else:
            ft = "Laplace" if kinds[i] == 0 else "Ralpha"
            tickets[i] = S.submit(ft, coord, q, tol=tols[i], Rdir=8.0, alpha=alphas[i],
                                  tArrival=arrive[i], V=V)
        ready.release()
    pt.join()
    th.join()
    watch.stop()
    S.close()
    gc.enable()
    ok = [i for i in range(nreq) if infos[i] is not None]
    lat = np.array([infos[i]["tEnd"] - infos[i]["tArrival"] for i in ok])
    t_first = arrive[0]
    t_last = max(infos[i]["tEnd"] for i in ok)
    charges = int(sum(sizes[i] for i in ok))
    by = {}
    for k, name in enumerate(["laplace", "ralpha", "helmholtz"]):
        sel = [j for j, i in enumerate(ok) if kinds[i] == k]
        if sel:
            by[name] = {"requests": len(sel), "p50_ms": float(np.percentile(lat[sel], 50) * 1e3),
                        "p99_ms": float(np.percentile(lat[sel], 99) * 1e3)}
    big = [j for j, i in enumerate(ok) if sizes[i] >= 500000]

    def phases(sel):
        # mean wall time per phase (ms) of the selected requests; "wait" is
        # the time from arrival to a slot taking the request
        if not sel:
            return None
        out = {"requests": len(sel),
               "wait": float(np.mean([(infos[ok[j]]["tStart"] - infos[ok[j]]["tArrival"]) * 1e3
                                      for j in sel]))}
        for k in ("keywords", "setup", "copy_in", "calcv", "copy_out"):
            out[k] = float(np.mean([infos[ok[j]]["ms"][k] for j in sel]))
        return out
    tail = [j for j in range(len(ok)) if lat[j] >= np.percentile(lat, 99)]
    small = [j for j, i in enumerate(ok) if sizes[i] < 30000]
    return {"requests": nreq, "completed": len(ok), "errors": errors[:5],
            "rate_per_s": rate, "slots": slots,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p90_ms": float(np.percentile(lat, 90) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "max_ms": float(lat.max() * 1e3),
            "mean_ms": float(lat.mean() * 1e3),
            "p99_ms_N_ge_500k": float(np.percentile(lat[big], 99) * 1e3) if big else None,
            "charges_per_s": charges / (t_last - t_first), "total_charges": charges,
            "max_submit_lag_ms": float(lags.max() * 1e3),
            "p99_submit_lag_ms": float(np.percentile(lags, 99) * 1e3),
            "max_submit_lag_request": int(np.argmax(lags)), "by_kernel": by,
            "phase_ms_N_ge_500k": phases(big), "phase_ms_slowest_1pct": phases(tail),
            "phase_ms_N_lt_30k": phases(small),
            "workload": "Poisson %g req/s, N log-uniform [500, 1e6], density 0.1, 70%% 1/r / 20%% "
                        "r^alpha (0.5|1.5) / 10%% Helmholtz k0=0.3, tol 1e-4|1e-6, Rdir 8, mixed "
                        "precision, seed %d; latency includes set-up (RSEM grid, cell sort, lists, "
                        "kernel grid), host<->device copies and CalcV; request buffers %s"
                        % (rate, seed, "page-locked (pool of 24)" if pinned else "pageable")}


def serving_stream_child(a):
    """The serving stream in a child process: a crash there (the process
    ends on a signal) leaves the rest of the bench line intact and is
    reported in it, with the child's fault-handler traceback."""
    import subprocess
    import tempfile
    fd, crash = tempfile.mkstemp(prefix="npme_stream_", suffix=".log")
    os.close(fd)
    cmd = [sys.executable, os.path.abspath(__file__), "--serving-only", "--crash-log", crash,
           "--stream-requests", str(a.stream_requests), "--stream-rate", str(a.stream_rate),
           "--slots", str(a.slots)]
    if a.pageable_requests:
        cmd.append("--pageable-requests")
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, timeout=900, text=True)
        lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
        if r.returncode == 0 and lines:
            return json.loads(lines[-1])["latency"]
        tail = open(crash).read()[-4000:]
        log("serving stream process failed with status %d" % r.returncode)
        sys.stderr.write(tail)
        err = ("serving stream stalled: no request completed for 60 s" if r.returncode == 3 else
               "serving stream process exited with status %d" % r.returncode)
        return {"error": err, "traceback": tail}
    except subprocess.TimeoutExpired:
        return {"error": "serving stream process exceeded 900 s"}
    finally:
        os.unlink(crash)


def main():
    import faulthandler
    a = parse()
    if a.serving_only:
        diag = open(a.crash_log, "w") if a.crash_log else sys.stderr
        faulthandler.enable(file=diag, all_threads=True)
        lat = serving_stream(a.stream_requests, a.stream_rate, a.slots, pinned=not a.pageable_requests,
                             diag=diag)
        print(json.dumps({"latency": lat}))
        return 0
    faulthandler.enable(all_threads=True)
    if a.gpus != 1:
        print("bench.py measures one MI355X: --gpus %d is not supported (only --gpus 1)" % a.gpus,
              file=sys.stderr)
        return 2
    import torch
    from npme_amd.inputs import metric_input
    n = a.n
    coord, q = metric_input(n)
    dev = torch.device("cuda:0")
    dq = torch.from_numpy(q).to(dev)
    dV = torch.empty((n, 4), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # the library reads dq from its own streams

    h, t_setup = setup(coord, q, "mixed")
    Ngrid = (h.Get_N1(), h.Get_N2(), h.Get_N3())
    Nder = h.Get_Nder()
    log("grid", Ngrid, "Nder", Nder, "setup %.3f s" % t_setup)
    settle(h, dV, dq)
    if a.stages_only:
        stage_profile(h, dV, dq, max(1, a.warmup))
        stages = stage_profile(h, dV, dq, a.steps)
        print(json.dumps({"stage_ms_serialised": dict(zip(STAGES, np.round(stages, 4).tolist())),
                          "ordered_pairs_in_cutoff": h.count_pairs(), "steps": a.steps}))
        return 0
    ms = time_steps(h, dV, dq, a.steps, a.warmup)
    # the same calls launching every kernel each time (no graph replay)
    h.set_calcv_graph(False)
    ms_nograph = time_steps(h, dV, dq, a.steps, a.warmup)
    h.set_calcv_graph(True)
    log("CalcV %.3f ms/step with graph replay, %.3f ms/step launching each call" % (ms, ms_nograph))
    if a.calib_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 3 x %d bytes read and written" % (src.numel() * 4))
        del src, dst
    stages = stage_profile(h, dV, dq)
    # the host-buffer entry point (npme_calcv_real: charges in and V out over
    # PCIe, pageable numpy arrays); reported beside value, never as it
    Vh = np.empty((n, 4))
    h.CalcV(Vh, n, q, coord)
    th = []
    for _ in range(5):
        t0 = time.perf_counter()
        h.CalcV(Vh, n, q, coord)
        th.append((time.perf_counter() - t0) * 1e3)
    host_ms = float(np.median(th))
    log("host buffers: %.3f ms/step" % host_ms)
    P = h.count_pairs()
    V = dV.cpu().numpy()
    acc = None if a.no_accuracy else accuracy(h, q, V, n)
    log("mixed: %.3f ms/step" % ms, acc)

    fp64 = None
    if not a.no_fp64:
        del h
        h64, t64 = setup(coord, q, "fp64")
        settle(h64, dV, dq)
        ms64 = time_steps(h64, dV, dq, max(3, a.steps // 4), 1)
        st64 = stage_profile(h64, dV, dq, 3)
        acc64 = None if a.no_accuracy else accuracy(h64, q, dV.cpu().numpy(), n)
        fp64 = {"value": n / (ms64 / 1e3), "ms_per_step": ms64, "setup_s": t64,
                "accuracy": acc64, "stage_ms": dict(zip(STAGES, np.round(st64, 4).tolist()))}
        log("fp64: %.3f ms/step" % ms64, acc64)
        del h64

    # roofline of each stage from its algorithmic work per launch
    M1, M2, M3 = (x // 2 for x in Ngrid)
    K3 = Ngrid[2] // 2 + 1
    Mt = M1 * M2 * M3
    Fc = (Ngrid[0] // 2 + 1) * (Ngrid[1] // 2 + 1) * K3
    algo = {
        "gather": ("hbm", n * (4 + 8 + 4 + 4)),
        "direct": ("fp32", P * (20 + 4 * Nder)),
        "spread": ("hbm", n * 16 + 4 * Mt),
        "fftA_r2c_z": ("hbm", 4 * Mt + 8 * M1 * M2 * K3),
        "fftB_y": ("hbm", 8 * M1 * M2 * K3 + 8 * M1 * Ngrid[1] * K3),
        "fftC_x_conv": ("hbm", 2 * 8 * M1 * Ngrid[1] * K3 + 4 * Fc),
        "fftD_y": ("hbm", 8 * M1 * Ngrid[1] * K3 + 8 * M1 * M2 * K3),
        "fftE_c2r_z": ("hbm", 8 * M1 * M2 * K3 + 4 * Mt),
        "interp_combine": ("hbm", n * (16 + 32 + 8 + 4 + 32) + 4 * Mt),
    }
    stage_info = {}
    for name, t in zip(STAGES, stages):
        kind, work = algo[name]
        if kind == "hbm":
            ach = work / (t * 1e-3) / 1e9
            stage_info[name] = {"ms": round(float(t), 4), "GB/s": round(ach, 1),
                                "frac": round(ach / HBM_PEAK, 4), "bytes": work}
        else:
            ach = work / (t * 1e-3) / 1e12
            stage_info[name] = {"ms": round(float(t), 4), "TFLOP/s": round(ach, 3),
                                "frac": round(ach / FP32_PEAK, 4), "flops": work}
    dom = STAGES[int(np.argmax(stages))]
    kind, work = algo[dom]
    t_dom = float(stages[int(np.argmax(stages))])
    if kind == "hbm":
        ach = work / (t_dom * 1e-3) / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": None}
    else:
        ach = work / (t_dom * 1e-3) / 1e12
        roof = {"kernel": dom, "bound": "valu-fp32", "achieved": ach, "peak": FP32_PEAK,
                "unit": "TFLOP/s", "frac": ach / FP32_PEAK, "traffic": None,
                "note": "direct sum has no MFMA form; FP32 VALU peak"}
    roof["timing"] = ("average launch time from HIP events on the library's stream with the "
                      "stages serialised (npme_set_profiling); the timed CalcV runs the direct "
                      "sum concurrently with the reciprocal sum")
    tf = os.path.join(ROOT, "profiles", "traffic_%s_1M.txt" % dom)
    if os.path.exists(tf) and n == 1000000:
        for line in open(tf):
            if line.startswith("traffic_bytes_per_launch"):
                roof["traffic"] = float(line.split()[1])

    latency = None
    if a.stream_requests > 0:
        log("serving stream: %d requests at %g/s on %d slots" % (a.stream_requests, a.stream_rate, a.slots))
        latency = serving_stream_child(a)
        if "p50_ms" in latency:
            log("stream: p50 %.2f ms p99 %.2f ms" % (latency["p50_ms"], latency["p99_ms"]))

    cpu = None
    if not a.no_cpu_baseline:
        from oracle import oracle as O
        ko = O.kernel(O.LAPLACE, Nder, 8.0)
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log("cpu baseline: oracle CalcV at N=%d on %d threads" % (n, nthreads))
        t0 = time.perf_counter()
        O.calcv(ko, Ngrid, 6, coord, q)
        tw = time.perf_counter() - t0
        tt = O.last_calcv_times()
        tc = tt["direct"] + tt["recip"]
        cpu = {"value": n / tc, "unit": "charges/s", "cores": nthreads, "kind": "port",
               "sample": "full workload (N=%d, grid %s), compute only: direct %.2f s + "
                         "reciprocal %.2f s; kernel-grid set-up %.2f s excluded; wall %.1f s"
                         % (n, "x".join(map(str, Ngrid)), tt["direct"], tt["recip"], tt["setup"], tw)}

    out = {
        "metric": "charges/sec for 1/r DM at N=1M, rel err 1e-6; p99 latency on mixed-size stream",
        "value": n / (ms / 1e3),
        "unit": "charges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (mixed: FP32 spreading/FFT/direct pairs, FP64 accumulation and set-up)",
        "data": "synthetic",
        "config": {"workload": "Laplace 1/r DM CalcV, N=%d uniform cube density 0.1, Bn=6, "
                               "Rdir=8, tol=1e-6, grid %s, Nder %d" % (n, "x".join(map(str, Ngrid)), Nder),
                   "N": n, "grid": list(Ngrid), "Nder": Nder, "BnOrder": 6, "Rdir": 8.0,
                   "precision": "mixed", "parallelism": "single GPU"},
        "accuracy_vs_exact": acc,
        "setup_s": t_setup,
        "host_buffers": {"ms_per_step": host_ms, "charges_per_s": n / (host_ms / 1e3),
                         "note": "npme_calcv_real from pageable host arrays: 8 MB charges in, "
                                 "32 MB V out over PCIe, median of 5"},
        "launch": {"graph_replay_ms_per_step": ms, "kernel_launch_ms_per_step": ms_nograph,
                   "note": "value uses the default mode: the CalcV launches captured once as a "
                           "HIP graph and replayed"},
        "ordered_pairs_in_cutoff": P,
        "stage_ms": dict(zip(STAGES, np.round(stages, 4).tolist())),
        "stages": stage_info,
        "fp64_gpu_path": fp64,
        "latency": latency,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
