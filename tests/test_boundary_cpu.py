"""CPU-side checks of the C-ABI library: it loads, exports every entry
point declared in include/npme_amd.h, and its host-only logic (keyword
defaults, validation and the RSEM grid choice of NPME_AddMissingKeywords)
agrees with the oracle.  No device calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, uniform_box

HEADER = os.path.join(ROOT, "include", "npme_amd.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(npme_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import npme_amd
    L = ctypes.CDLL(npme_amd.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def _kw(**extra):
    import npme_amd as N
    k = N.NPME_KeywordInput()
    base = dict(calcType="pme", funcType="Laplace", Rdir=8.0, BnOrder=6, tol=1e-6)
    base.update(extra)
    for a, b in base.items():
        k.AddKeyword(a, b)
    return k


def test_add_missing_keywords_defaults_and_rsem(oracle):
    import npme_amd as N
    coord, q = uniform_box(10000, 40.0)
    k = _kw()
    assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q)
    d = k.as_dict()
    # defaults of src/Interface.cpp:53-517 and src/Constant.h:48-63
    assert d["EwaldSplit"] == "DerivMatch" and d["useCompactF"] == "1"
    assert float(d["FFTmemGB"]) == 10.0 and d["nNeigh"] == "2" and d["nCellClust1D"] == "3"
    assert all(float(d[x]) == 4.0 for x in ("a1", "a2", "a3"))
    assert all(float(d[x]) == 8.0 for x in ("del1", "del2", "del3"))
    grid = tuple(int(d[x]) for x in ("N1", "N2", "N3"))
    blk = tuple(int(d[x]) for x in ("n1", "n2", "n3"))
    assert (grid, blk, int(d["nDeriv"])) == oracle.rsem_static(q, coord, 6, -1.0, 1e-6, 8.0)
    # default BnOrder is 8
    k8 = N.NPME_KeywordInput()
    for a, b in dict(calcType="pme", funcType="Laplace", Rdir=8.0).items():
        k8.AddKeyword(a, b)
    assert N.NPME_AddMissingKeywords(k8, len(q), coord, chargeReal=q)
    assert k8.GetValue("BnOrder") == "8"


def test_add_missing_keywords_explicit_grid_nderiv_default():
    import npme_amd as N
    coord, q = uniform_box(500, 10.0)
    for bn, nd in ((4, 3), (6, 6), (8, 8)):
        k = _kw(BnOrder=bn, N1=64, N2=64, N3=64, n1=16, n2=16, n3=16)
        assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q)
        assert k.GetValue("nDeriv") == str(nd)


@pytest.mark.parametrize("Bn", [5, 7])
def test_setup_rejects_odd_spline_order(Bn):
    """An odd BnOrder is refused before any device work (its |b(m)|^2
    vanishes at m = N/2), so this runs without a GPU."""
    import npme_amd as N
    coord, q = uniform_box(500, 10.0)
    k = _kw(BnOrder=Bn, N1=64, N2=64, N3=64, n1=16, n2=16, n3=16, nDeriv=6)
    assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q)
    assert N.NPME_Interface_SetUpRealKernel(k, len(q), coord) is None
    assert "even" in N.npme_last_error()


@pytest.mark.parametrize("Bn", [14, 16])
def test_setup_rejects_mixed_above_order_12(Bn):
    """The mixed form is provided up to spline order 12; higher orders are
    refused with NPME_ERR_UNSUPPORTED before any device work unless
    precision fp64 is asked for."""
    import npme_amd as N
    coord, q = uniform_box(500, 10.0)
    k = _kw(BnOrder=Bn, N1=80, N2=80, N3=80, n1=8, n2=8, n3=8, nDeriv=6, precision="mixed")
    assert N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q)
    assert N.NPME_Interface_SetUpRealKernel(k, len(q), coord) is None
    assert "fp64" in N.npme_last_error()


@pytest.mark.parametrize("bad", [dict(calcType="fast"), dict(funcType="Yukawa"),
                                 dict(EwaldSplit="Other"), dict(funcType="Ralpha"),
                                 dict(funcType="Helmholtz", k0_r=1.0, k0_i=0.0)])
def test_add_missing_keywords_rejects(bad):
    import npme_amd as N
    coord, q = uniform_box(500, 10.0)
    k = _kw(**bad)
    assert not N.NPME_AddMissingKeywords(k, len(q), coord, chargeReal=q)
    assert N.npme_last_error()


def test_add_missing_keywords_helmholtz(oracle):
    import npme_amd as N
    coord, _ = uniform_box(4000, 30.0)
    rng = np.random.default_rng(3)
    qc = rng.standard_normal(4000) + 1j * rng.standard_normal(4000)
    k = _kw(funcType="Helmholtz", k0_r=0.5, k0_i=0.0, Rdir=10.0, BnOrder=8)
    assert N.NPME_AddMissingKeywords(k, len(qc), coord, chargeComplex=qc)
    d = k.as_dict()
    got = (tuple(int(d[x]) for x in ("N1", "N2", "N3")), tuple(int(d[x]) for x in ("n1", "n2", "n3")),
           int(d["nDeriv"]))
    assert got == oracle.rsem_helmholtz(qc, coord, 8, 0.5, 1e-6, 10.0)
    # negative imaginary wave number is refused (src/Interface.cpp:160-165)
    k2 = _kw(funcType="Helmholtz", k0_r=0.5, k0_i=-0.1, Rdir=10.0, BnOrder=8)
    assert not N.NPME_AddMissingKeywords(k2, len(qc), coord, chargeComplex=qc)


def test_keyword_file_roundtrip(tmp_path):
    import npme_amd as N
    p = tmp_path / "kw.txt"
    p.write_text("@Keyword\n  BnOrder  6  (B-spline order)\n  calcType pme_exact\n"
                 "  funcType Laplace\n  Rdir 4.0   (direct space cutoff)\n")
    k = N.NPME_KeywordInput(str(p))
    assert k.GetValue("BnOrder") == "6" and k.GetValue("Rdir") == "4.0"
    assert k.GetValue("calcType") == "pme_exact" and k.GetValue("missing") is None


def test_cpp_dropin_header_host_calls(oracle, tmp_path):
    """include/npme_amd.hpp compiles as the reference's NPME_Library
    interface, links against the library and returns the same RSEM grid as
    the oracle (host calls only)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "npme_amd")
    exe = str(tmp_path / "wrapper_host")
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "wrapper_host.cpp"), "-L", libdir, "-lnpme_amd",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    coord, q = uniform_box(10000, 40.0)
    inp = tmp_path / "charges.txt"
    np.savetxt(inp, np.column_stack([coord, q]), fmt="%.17g")
    out = subprocess.run([exe, str(inp)], check=True, capture_output=True, text=True).stdout.split()
    N, nb, nd = oracle.rsem_static(q, coord, 6, -1.0, 1e-6, 8.0)
    assert [int(v) for v in out] == list(N) + list(nb) + [nd]


def test_cpp_header_files_and_descriptors(tmp_path):
    """The src/ReadPrint.h / SupportFunctions.h helpers and the built-in
    kernel classes of the C++ header (host Calc: LR + SR = the kernel,
    NPME_KernelFuncCheck, NPME_Interface_SelectKernelPtr), compiled with
    -Wall -Wextra -Werror and run on the host (tests/cpp/readprint_host.cpp)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "npme_amd")
    exe = str(tmp_path / "readprint_host")
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "readprint_host.cpp"), "-L", libdir, "-lnpme_amd",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)


def test_server_request_board_threads(tmp_path):
    """The serving loop's ticket bookkeeping (npme_amd/csrc/request_board.h)
    under threads, on the host: 20,000 requests submitted while four
    threads wait for them and six workers finish them in random order; a
    ticket is waited for once, a second waiter is refused, an unknown
    ticket is refused (tests/cpp/request_board_host.cpp)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "request_board")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-pthread", "-I",
                    os.path.join(root, "npme_amd", "csrc"),
                    os.path.join(root, "tests", "cpp", "request_board_host.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_user_define_func_tool_host():
    """test/06_userDefineFunc: the userDefineFunc tool (app/userDefineFunc.cpp
    against the C++ header) evaluates its plugin rFunc on 8 random points in
    both Calc forms and passes NPME_KernelFuncCheck (analytic against
    numerical derivatives) on the host."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "npme_amd", "bin", "userDefineFunc")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    xyz = [list(map(float, l.split("=")[1].split())) for l in lines if l.startswith("xyz =")]
    rs = [float(l.split("=")[1]) for l in lines if l.startswith("r =") and "drdx" not in l]
    grads = [l for l in lines if "drdx" in l]
    assert len(xyz) == 8 and len(rs) == 8 and len(grads) == 8
    for p, rv in zip(xyz, rs):
        assert abs(rv - np.sqrt(sum(c * c for c in p))) < 1e-4
    assert "NPME_KernelFuncCheck passed" in r.stdout


def test_bench_stall_watch():
    """The serving stream's watchdog fires once when progress stops for
    longer than its limit, and never while progress is marked."""
    import time
    import bench
    fired = []
    w = bench.StallWatch(0.3, lambda: fired.append(time.monotonic()), poll_s=0.02).start()
    for _ in range(20):  # progress every 0.05 s for 1 s
        time.sleep(0.05)
        w.mark()
    assert fired == [] and w.count == 20
    time.sleep(0.6)
    w.stop()
    assert len(fired) == 1
