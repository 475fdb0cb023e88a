"""User-defined kernels through NPME_InterfaceReal_GenFunc and
NPME_InterfaceComplex_GenFunc (src/Interface.h:291-296, :563-650): the
reference's example kernel rFunc of app/userDefineFunc.cpp (f = r, with the
split r erf(beta r) + r erfc(beta r) its header describes), and a complex
version f = c r, as Python plugins (NPME_KfuncReal / NPME_KfuncComplex
subclasses whose Calc works in place on the library's arrays).

Checked through the C ABI against the oracle's restatement of the same
GenFunc path (the rFunc kernels evaluated directly, no table, and the
reference's direct-sum pair set: funcSR on every point pair of the
interacting cell pairs, src/PotentialGenFunc.cpp:523-645) on the same 96^3
grid, and against the exact sum:
  fp64 GPU path vs oracle CalcV ............. 1e-10 (V and grad): the device
                                               short-range table's error is
                                               below 1e-12
  mixed GPU path vs oracle CalcV ............ 1e-6 (V), 1e-5 (grad)
  either path vs the exact sum .............. 1e-6 (the oracle's own error
                                               on this grid is 4.6e-8)
  CalcV_exact (the user's full kernel on the
  host) vs the oracle's exact sum ........... 1e-12"""
import numpy as np
import pytest
from scipy.special import erf, erfc

from conftest import uniform_box

pytestmark = pytest.mark.gpu

BETA, RDIR, NG = 0.5, 8.0, 96


def _r(x, y, z):
    return np.sqrt(np.abs(x * x + y * y + z * z))


def _parts(kind, r):
    """f(r) and f'(r) of the full kernel r, its long-range part r erf(beta r)
    or its short-range part r erfc(beta r)."""
    b = BETA
    g = 2.0 / np.sqrt(np.pi) * b * np.exp(-(b * r) ** 2)
    if kind == "full":
        return r, np.ones_like(r)
    if kind == "lr":
        return r * erf(b * r), erf(b * r) + r * g
    return r * erfc(b * r), erfc(b * r) - r * g


def _unit(x, r):
    # the direction cosine x / r (0 at r = 0, where the derivative of a
    # smooth radial function vanishes)
    with np.errstate(invalid="ignore", divide="ignore"):
        u = np.where(r > 0, x / np.where(r > 0, r, 1.0), 0.0)
    return u


def _real_kernel(kind):
    import npme_amd as N

    class RFunc(N.NPME_KfuncReal):
        """rFunc of app/userDefineFunc.cpp:45-215 and its split parts."""

        def Calc(self, n, *a):
            if len(a) == 3:  # Calc(N, x_f0, y, z)
                x_f0, y, z = a
                x_f0[:] = _parts(kind, _r(x_f0, y, z))[0]
            else:  # Calc(N, f0, x_fX, y_fY, z_fZ)
                f0, x, y, z = a
                r = _r(x, y, z)
                f, d = _parts(kind, r)
                f0[:] = f
                ux, uy, uz = _unit(x, r), _unit(y, r), _unit(z, r)
                x[:], y[:], z[:] = d * ux, d * uy, d * uz

    return RFunc()


def _complex_kernel(kind, c):
    import npme_amd as N

    class CFunc(N.NPME_KfuncComplex):
        """f = c r (and c times its split parts) as a complex kernel."""

        def Calc(self, n, *a):
            if len(a) == 4:  # Calc(N, x_f0_r, f0_i, y, z)
                x_f0_r, f0_i, y, z = a
                f = _parts(kind, _r(x_f0_r, y, z))[0]
                x_f0_r[:], f0_i[:] = c.real * f, c.imag * f
            else:  # Calc(N, f0_r, f0_i, x_fX_r, fX_i, y_fY_r, fY_i, z_fZ_r, fZ_i)
                f0r, f0i, x, xi, y, yi, z, zi = a
                r = _r(x, y, z)
                f, d = _parts(kind, r)
                f0r[:], f0i[:] = c.real * f, c.imag * f
                gx, gy, gz = d * _unit(x, r), d * _unit(y, r), d * _unit(z, r)
                x[:], xi[:] = c.real * gx, c.imag * gx
                y[:], yi[:] = c.real * gy, c.imag * gy
                z[:], zi[:] = c.real * gz, c.imag * gz

    return CFunc()


def _keywords(prec, n, coord, q, cx):
    import npme_amd as N
    k = N.NPME_KeywordInput(calcType="pme", funcType="Helmholtz" if cx else "Laplace",
                            Rdir=RDIR, BnOrder=8, tol=1e-6, precision=prec, N1=NG, N2=NG,
                            N3=NG, n1=8, n2=8, n3=8, nDeriv=6, k0_r=0.0, k0_i=0.0)
    if cx:
        assert N.NPME_AddMissingKeywords(k, n, coord, chargeComplex=q), N.npme_last_error()
    else:
        assert N.NPME_AddMissingKeywords(k, n, coord, chargeReal=q), N.npme_last_error()
    return k


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_genfunc_real_rfunc(oracle, prec):
    import npme_amd as N
    n = 2000
    coord, q = uniform_box(n, 20.0)
    print("case genfunc real rFunc", prec, flush=True)
    k = _keywords(prec, n, coord, q, False)
    h = N.NPME_InterfaceReal_GenFunc()
    assert h.SetUp(k, n, coord, _real_kernel("full"), _real_kernel("lr"), _real_kernel("sr")), \
        N.npme_last_error()
    ef, eg = h.table_error()
    print("table error", ef, eg, h.Print())
    assert ef < 1e-12 and eg < 1e-11
    V = h.calcv(q)
    ko = oracle.kernel_rfunc(RDIR, BETA)
    mo = oracle.compare_v(oracle.calcv_ref(ko, (NG, NG, NG), 8, coord, q, nblk=(8, 8, 8)), V)
    Ve = oracle.exact(ko, coord, q)
    me = oracle.compare_v(Ve, V)
    print(prec, "vs oracle", mo, "vs exact", me)
    if prec == "fp64":
        assert mo["relV"] < 1e-10 and mo["reldV"] < 1e-10
    else:
        assert mo["relV"] < 1e-6 and mo["reldV"] < 1e-5
    assert me["relV"] < 1e-6
    # CalcV_exact evaluates the user's full kernel on the host
    Vx = np.zeros_like(V)
    assert h.CalcV_exact(Vx, n, q, coord)
    assert oracle.compare_v(Ve, Vx)["relV"] < 1e-12
    idx = np.array([0, 5, 1999])
    np.testing.assert_allclose(h.calcv_exact_subset(q, idx), Vx[idx], rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_genfunc_complex_rfunc(oracle, prec):
    import npme_amd as N
    n = 2000
    coord, _ = uniform_box(n, 20.0)
    rng = np.random.default_rng(21)
    qc = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    c = 1.0 + 0.5j
    print("case genfunc complex c r", prec, flush=True)
    k = _keywords(prec, n, coord, qc, True)
    h = N.NPME_InterfaceComplex_GenFunc()
    assert h.SetUp(k, n, coord, _complex_kernel("full", c), _complex_kernel("lr", c),
                   _complex_kernel("sr", c)), N.npme_last_error()
    V = h.calcv(qc)
    ko = oracle.kernel_rfunc(RDIR, BETA, c=c, complex_kernel=True)
    mo = oracle.compare_v(oracle.calcv_ref(ko, (NG, NG, NG), 8, coord, qc, nblk=(8, 8, 8)), V)
    Ve = oracle.exact(ko, coord, qc)
    me = oracle.compare_v(Ve, V)
    print(prec, "vs oracle", mo, "vs exact", me)
    if prec == "fp64":
        assert mo["relV"] < 1e-10 and mo["reldV"] < 1e-10
    else:
        assert mo["relV"] < 1e-6 and mo["reldV"] < 1e-5
    assert me["relV"] < 1e-6
    Vx = np.zeros_like(V)
    assert h.CalcV_exact(Vx, n, qc, coord)
    assert oracle.compare_v(Ve, Vx)["relV"] < 1e-12


def test_genfunc_reset_coord_and_builtin_agree(oracle):
    """ResetCoord re-samples the user's long-range part for the new box;
    the result equals a fresh set-up, and the built-in r^alpha path of
    the same class is unchanged."""
    import npme_amd as N
    n = 1500
    coord, q = uniform_box(n, 18.0)
    k = _keywords("fp64", n, coord, q, False)
    fs = (_real_kernel("full"), _real_kernel("lr"), _real_kernel("sr"))
    h = N.NPME_InterfaceReal_GenFunc()
    assert h.SetUp(k, n, coord, *fs)
    coord2 = coord * 1.05 + 2.0
    assert h.ResetCoord(n, coord2)
    V2 = h.calcv(q)
    h2 = N.NPME_InterfaceReal_GenFunc()
    assert h2.SetUp(k, n, coord2, *fs)
    np.testing.assert_array_equal(V2, h2.calcv(q))


def test_genfunc_rejects_non_radial_short_range():
    import npme_amd as N

    class Aniso(N.NPME_KfuncReal):
        def Calc(self, n, *a):
            if len(a) == 3:
                a[0][:] = a[0] * 2.0 + a[1]
            else:
                f0, x, y, z = a
                f0[:] = x * 2.0 + y
                x[:], y[:], z[:] = 2.0, 1.0, 0.0

    n = 500
    coord, q = uniform_box(n, 12.0)
    k = _keywords("fp64", n, coord, q, False)
    h = N.NPME_InterfaceReal_GenFunc()
    assert not h.SetUp(k, n, coord, _real_kernel("full"), _real_kernel("lr"), Aniso())
    assert "not radial" in N.npme_last_error()


def test_genfunc_calc_exception_is_raised():
    import npme_amd as N

    class Broken(N.NPME_KfuncReal):
        def Calc(self, n, *a):
            raise ValueError("user kernel failed")

    n = 500
    coord, q = uniform_box(n, 12.0)
    k = _keywords("fp64", n, coord, q, False)
    h = N.NPME_InterfaceReal_GenFunc()
    with pytest.raises(ValueError, match="user kernel failed"):
        h.SetUp(k, n, coord, _real_kernel("full"), Broken(), _real_kernel("sr"))
