"""GPU runs of the command-line tools on the reference's own test inputs
(tests/golden/cli: the keyword files of test/01, 02, 04, 05 and the
makeRandomBox inputs, with nCharge lowered to 4,000 so the exact sum stays
short): makeRandomBox writes the box, npme / RalphaDM / helmholtzDM run
calcType pme_exact, and compareV reads the two potential files.  Checked:
the relative errors in the log meet the requested tolerance, compareV
reports the same errors from the files, and the pme file holds the values
of the library's CalcV for the same keywords."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "npme_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden", "cli")


def _run(args, cwd, timeout=300):
    print("run", " ".join(os.path.basename(a) for a in args), flush=True)
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _kw_file(src, dst, **repl):
    out = []
    for line in open(src).read().split("\n"):
        w = line.split()
        if len(w) >= 2 and w[0] in repl:
            line = line.replace(w[1], str(repl.pop(w[0])), 1)
        out.append(line)
    out += ["  %s %s" % kv for kv in repl.items()]
    open(dst, "w").write("\n".join(out) + "\n")


def _box(tmp, kind):
    kw = os.path.join(tmp, "box%s.txt" % kind)
    _kw_file(os.path.join(GOLD, "input_makeRandomBox%s.txt" % kind), kw, nCharge=4000,
             crdFile="crd%s.txt" % kind)
    _run([os.path.join(BIN, "makeRandomBox"), kw], tmp)
    return "crd%s.txt" % kind


CASES = [
    # tool, keyword fixture, charge kind, relV bound
    ("npme", "laplaceDM_04.txt", "Real", 5e-6),
    ("laplaceDM", "laplaceDM_04.txt", "Real", 5e-6),
    ("RalphaDM", "RalphaDM_05.txt", "Real", 5e-6),
    ("helmholtzDM", "helmholtzDM_04.txt", "Complex", 5e-6),
    ("npme", "helmholtzDM_04.txt", "Complex", 5e-6),
    # erf/erfc split on the grid RSEM picks for the derivative match: less
    # accurate than DM there, as in the reference
    ("npme", "laplaceOrig_06.txt", "Real", 2e-5),
]


@pytest.mark.parametrize("tool,kwfile,kind,bound", CASES)
def test_calc_tool_pme_exact(tmp_path, tool, kwfile, kind, bound):
    import npme_amd as N
    tmp = str(tmp_path)
    print("case cli", tool, kwfile, flush=True)
    crd = _box(tmp, kind)
    kw = os.path.join(tmp, kwfile)
    _kw_file(os.path.join(GOLD, kwfile), kw, calcType="pme_exact")
    _run([os.path.join(BIN, tool), crd, kwfile], tmp)
    stem = crd.split(".")[0] + "_" + kwfile.split(".")[0]
    log = open(os.path.join(tmp, stem + ".log")).read()
    m = re.search(r"errorV = (\S+) errordVdr = (\S+) \(Relative Error\)", log)
    assert m, log[-2000:]
    relV, reldV = float(m.group(1)), float(m.group(2))
    print(tool, kwfile, "relV %.2e reldV %.2e" % (relV, reldV))
    assert relV <= bound
    assert re.search(r"time = \S+ \(direct sum\)", log) and re.search(r"\(rec sum\)", log)
    assert "time_exact" in log and "Total_Run_Time" in log
    out = _run([os.path.join(BIN, "compareV"), stem + "_V_pme.output", stem + "_V_exact.output"],
               tmp)
    c = re.search(r"error_V = (\S+) error_dVrdr = (\S+) \(rel\. error\)", out)
    # the files hold 11 significant digits
    assert float(c.group(1)) == pytest.approx(relV, rel=0.02, abs=1e-9)
    # the pme file is the library's CalcV for the same keywords
    coord, q = N.NPME_ReadSingleBoxChargeCoord(os.path.join(tmp, crd))
    k = N.NPME_KeywordInput(kw)
    if tool in ("laplaceDM", "RalphaDM", "helmholtzDM"):
        k.AddKeyword("funcType", {"laplaceDM": "Laplace", "RalphaDM": "Ralpha",
                                  "helmholtzDM": "Helmholtz"}[tool])
    cx = kind == "Complex"
    assert N.NPME_AddMissingKeywords(k, len(q), coord, **({"chargeComplex": q} if cx else
                                                          {"chargeReal": q}))
    setup = N.NPME_Interface_SetUpComplexKernel if cx else N.NPME_Interface_SetUpRealKernel
    h = setup(k, len(q), coord)
    assert h is not None, N.npme_last_error()
    V = h.calcv(q)
    Vf = N.NPME_ReadSingleBox_V(os.path.join(tmp, stem + "_V_pme.output"))
    assert np.all(np.abs(Vf - V) <= 1e-10 * np.abs(V) + 1e-300)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_user_define_func_calcv(tmp_path, prec):
    """userDefineFunc (app/userDefineFunc.cpp, test/06): the C++ plugin
    rFunc (f = r) and its erf/erfc parts through NPME_InterfaceReal_GenFunc
    on the GPU, 2,000 charges on a 96^3 grid, against CalcV_exact (the
    plugin's Calc on the host).  The oracle reaches 4.6e-8 on the same
    configuration."""
    out = _run([os.path.join(BIN, "userDefineFunc"), "calcv", "2000", prec], str(tmp_path))
    assert "NPME_KernelFuncCheck passed" in out
    m = re.search(r"relV (\S+) reldV (\S+)", out)
    assert m, out
    relV, reldV = float(m.group(1)), float(m.group(2))
    print(prec, "relV", relV, "reldV", reldV)
    assert relV < 1e-6 and reldV < 1e-5
