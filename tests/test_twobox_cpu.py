"""The host side of the two-box grids through the C ABI (no GPU): the
translated kernel grid of a user kernel (npme_recsumgrid_fulltrans_func_user,
NPME_RecSumGrid_FullTrans_Func of src/RecSumGrid.cpp:663-794 with an
NPME_KfuncReal / NPME_KfuncComplex) against the oracle's restatement, and
the argument checks of the box-pair and grid entry points."""
import numpy as np
import pytest


def _rfunc(c=None):
    import npme_amd as N
    base = N.NPME_KfuncComplex if c is not None else N.NPME_KfuncReal

    class RF(base):
        """f = r (or c r): rFunc of app/userDefineFunc.cpp:180-215"""

        def Calc(self, n, *a):
            if c is None and len(a) == 3:
                x, y, z = a
                x[:] = np.sqrt(x * x + y * y + z * z)
            elif c is not None and len(a) == 4:
                xr, fi, y, z = a
                r = np.sqrt(xr * xr + y * y + z * z)
                xr[:], fi[:] = c.real * r, c.imag * r
            else:
                raise AssertionError("the grid uses the value form only")

    return RF()


@pytest.mark.parametrize("c", [None, 0.7 - 0.4j])
def test_fulltrans_user_kernel_matches_oracle(oracle, c):
    import npme_amd as N
    Ng, a, dl, X = (20, 16, 24), (4.0, 4.0, 4.0), (5.0, 5.0, 5.0), (6.0, 5.0, 7.0)
    R0 = (26.0, -3.0, 1.5)
    f = N.NPME_RecSumGrid_FullTrans_Func(Ng, a, dl, X, _rfunc(c), R0)
    k = oracle.kernel_rfunc(8.0, 0.5, c=1.0 if c is None else c, complex_kernel=c is not None)
    fo = oracle.fgrid_fulltrans(k, Ng, a, dl, X, R0, part=oracle.FULL)
    np.testing.assert_allclose(f, fo, rtol=1e-14, atol=1e-14 * np.abs(fo).max())
    if c is None:
        assert np.abs(f.imag).max() == 0.0


def test_fulltrans_refuses_close_separation_and_odd_grid():
    import npme_amd as N
    with pytest.raises(N.NPMEError, match="separation"):
        N.NPME_RecSumGrid_FullTrans_Func((20, 16, 24), (4.0,) * 3, (5.0,) * 3, (6.0, 5.0, 7.0),
                                         _rfunc(), (10.0, 0.0, 0.0))
    with pytest.raises(N.NPMEError, match="even"):
        N.NPME_RecSumGrid_FullTrans_Func((21, 16, 24), (4.0,) * 3, (5.0,) * 3, (6.0, 5.0, 7.0),
                                         _rfunc(), (40.0, 0.0, 0.0))


def test_boxpair_keywords_need_the_grid():
    import npme_amd as N
    k = N.NPME_KeywordInput(funcType="Laplace", BnOrder=8, precision="fp64")
    cT = np.random.default_rng(1).random((10, 3))
    with pytest.raises(N.NPMEError):
        N.NPME_BoxPair(k, cT, cT + 30.0)
    assert "N1" in N.npme_last_error()


def test_cpp_fulltrans_reference_signature(oracle, tmp_path):
    """A C++ caller with the reference's NPME_RecSumGrid_FullTrans_Func
    signature (tests/cpp/twobox_host.cpp, -Wall -Wextra -Werror) gets the
    oracle's grid for a user kernel, and false for a separation inside the
    grid box."""
    import os
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.join(root, "npme_amd")
    exe = str(tmp_path / "twobox_host")
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "twobox_host.cpp"), "-L", libdir, "-lnpme_amd",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    v = np.array([[float(t) for t in line.split()] for line in r.stdout.strip().splitlines()])
    f = (v[:, 0] + 1j * v[:, 1]).reshape(12, 8, 10)
    k = oracle.kernel_rfunc(8.0, 0.5)
    fo = oracle.fgrid_fulltrans(k, (12, 8, 10), (4.0,) * 3, (3.0,) * 3, (4.0, 3.5, 4.5), (20.0, -2.0, 1.0))
    np.testing.assert_allclose(f, fo, rtol=1e-14, atol=0)
