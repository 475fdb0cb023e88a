"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/libnpme_oracle.so (built by oracle/Makefile).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as
the checker or as the timed CPU baseline; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnpme_oracle.so")

LAPLACE, RALPHA, HELMHOLTZ, LAPLACE_ERF, RFUNC, RFUNC_CX = 0, 1, 2, 3, 4, 5
_COMPLEX_KINDS = (HELMHOLTZ, RFUNC_CX)
FULL, LR, SR = 0, 1, 2


class Kernel(C.Structure):
    _fields_ = [("kind", C.c_int), ("alpha", C.c_double), ("k0r", C.c_double),
                ("k0i", C.c_double), ("Nder", C.c_int), ("Rdir", C.c_double),
                ("a", C.c_double * 62), ("b", C.c_double * 62), ("beta", C.c_double)]

    def coeffs(self):
        n = self.Nder + 1
        a = np.array(self.a[:2 * n]).reshape(n, 2)
        b = np.array(self.b[:2 * n]).reshape(n, 2)
        if self.kind in _COMPLEX_KINDS:
            return a[:, 0] + 1j * a[:, 1], b[:, 0] + 1j * b[:, 1]
        return a[:, 0].copy(), b[:, 0].copy()


class RecSumParm(C.Structure):
    _fields_ = [("N", C.c_long * 3), ("nblk", C.c_long * 3), ("Bn", C.c_long),
                ("del_", C.c_double * 3), ("a", C.c_double * 3)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, p, l, i = C.c_double, C.c_void_p, C.c_long, C.c_int
        sz = C.c_size_t
        L.npo_bspline.restype = d
        L.npo_bspline.argtypes = [i, d]
        L.npo_bspline_dw.restype = d
        L.npo_bspline_dw.argtypes = [i, d, C.POINTER(d)]
        L.npo_kernel_init.argtypes = [C.POINTER(Kernel), i, d, d, d, i, d]
        L.npo_rdir2beta.argtypes = [d, d]
        L.npo_rdir2beta.restype = d
        L.npo_kernel_init_erf.argtypes = [C.POINTER(Kernel), d, d]
        L.npo_kernel_init_rfunc.argtypes = [C.POINTER(Kernel), C.c_int, d, d, d, d]
        L.npo_kernel_eval.argtypes = [C.POINTER(Kernel), i, d, p, p]
        L.npo_optimal_grid_size.restype = l
        L.npo_optimal_grid_size.argtypes = [l]
        L.npo_fft_size_block.argtypes = [l, l, C.POINTER(l), C.POINTER(l)]
        L.npo_box_dim_center.argtypes = [sz, p, p, p]
        L.npo_pme_corr_box.argtypes = [p, p, p, p, l, p, p]
        L.npo_rsem_static.argtypes = [sz, p, p, l, d, d, d, i, p, p, C.POINTER(i)]
        L.npo_rsem_static_pred_error.argtypes = [sz, p, p, l, d, i, d, p, C.POINTER(d), C.POINTER(i)]
        L.npo_rsem_helmholtz.argtypes = [sz, p, p, l, d, d, d, d, i, p, p, C.POINTER(i)]
        L.npo_cell_index.argtypes = [sz, p, d, l, l, p, p, p]
        L.npo_fft1d.argtypes = [l, p, i]
        L.npo_fft3d.argtypes = [l, l, l, p, i]
        L.npo_fgrid_compact.argtypes = [C.POINTER(Kernel), p, p, p, p, p]
        L.npo_lambda2.argtypes = [l, l, p]
        L.npo_spread_basic.argtypes = [sz, p, i, p, p, p, l, p, p]
        L.npo_spline_start.argtypes = [sz, p, p, p, l, p, p]
        L.npo_theta_8parity.argtypes = [p, l, p, p, p]
        L.npo_theta_fullF_reverse.argtypes = [p, l, p, p, p, p]
        L.npo_fgrid_fulltrans.argtypes = [C.POINTER(Kernel), i, p, p, p, p, p, p]
        L.npo_min_distance_r0.restype = d
        L.npo_min_distance_r0.argtypes = [p, p]
        L.npo_interp_basic.argtypes = [sz, p, p, p, p, l, p, p]
        L.npo_direct_sr.argtypes = [C.POINTER(Kernel), sz, p, p, p]
        L.npo_exact.argtypes = [C.POINTER(Kernel), sz, p, p, sz, p, p]
        L.npo_direct_refpairs.argtypes = [C.POINTER(Kernel), sz, p, p, l, l, p]
        L.npo_count_pairs.restype = d
        L.npo_count_pairs.argtypes = [sz, p, d]
        L.npo_calcv.argtypes = [C.POINTER(Kernel), C.POINTER(RecSumParm), sz, p, p, p]
        L.npo_compare_v.argtypes = [sz, i, p, p, p]
        L.npo_last_calcv_times.argtypes = [p]
        L.npo_splitmix_uniform.argtypes = [C.c_ulonglong, sz, p]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128_as_f64(q):
    q = np.ascontiguousarray(q, dtype=np.complex128)
    return q.view(np.float64)


# ---------------------------------------------------------------- basics
def bspline(n, w):
    return lib().npo_bspline(int(n), float(w))


def bspline_dw(n, w):
    d = C.c_double()
    v = lib().npo_bspline_dw(int(n), float(w), C.byref(d))
    return v, d.value


def kernel(kind, Nder, Rdir, alpha=-1.0, k0=0.0):
    k = Kernel()
    k0 = complex(k0)
    rc = lib().npo_kernel_init(C.byref(k), int(kind), float(alpha), k0.real, k0.imag,
                               int(Nder), float(Rdir))
    if rc != 0:
        raise ValueError("bad kernel parameters")
    return k


def rdir2beta(Rdir, tol):
    return lib().npo_rdir2beta(float(Rdir), float(tol))


def kernel_erf(Rdir, beta):
    """1/r with the Ewald split (EwaldSplit = LaplaceOrig)"""
    k = Kernel()
    if lib().npo_kernel_init_erf(C.byref(k), float(Rdir), float(beta)) != 0:
        raise ValueError("bad kernel parameters")
    return k


def kernel_rfunc(Rdir, beta, c=1.0, complex_kernel=False):
    """The rFunc user kernel of app/userDefineFunc.cpp: f = c r, split as
    c r erf(beta r) + c r erfc(beta r)"""
    k = Kernel()
    c = complex(c)
    if lib().npo_kernel_init_rfunc(C.byref(k), int(bool(complex_kernel)), c.real, c.imag,
                                   float(Rdir), float(beta)) != 0:
        raise ValueError("bad kernel parameters")
    return k


def kernel_eval(k, part, r2):
    f0 = np.zeros(2)
    f1 = np.zeros(2)
    lib().npo_kernel_eval(C.byref(k), int(part), float(r2), _ptr(f0), _ptr(f1))
    if k.kind in _COMPLEX_KINDS:
        return complex(f0[0], f0[1]), complex(f1[0], f1[1])
    return f0[0], f1[0]


def optimal_grid_size(n):
    return lib().npo_optimal_grid_size(int(n))


def fft_size_block(Nideal, nideal=20):
    N, n = C.c_long(), C.c_long()
    lib().npo_fft_size_block(int(Nideal), int(nideal), C.byref(N), C.byref(n))
    return N.value, n.value


def box_dim_center(coord):
    coord = _f64(coord)
    X0, R0 = np.zeros(3), np.zeros(3)
    lib().npo_box_dim_center(len(coord), _ptr(coord), _ptr(X0), _ptr(R0))
    return X0, R0


def pme_corr_box(X0, R0, N, dl, Bn):
    X0, R0, dl = _f64(X0), _f64(R0), _f64(dl)
    Nn = np.ascontiguousarray(N, dtype=np.int64)
    X, R = np.zeros(3), np.zeros(3)
    lib().npo_pme_corr_box(_ptr(X0), _ptr(R0), _ptr(Nn), _ptr(dl), int(Bn), _ptr(X), _ptr(R))
    return X, R


def rsem_static(q, coord, Bn, alpha, tol, Rdir, nDeriv=-1):
    q, coord = _f64(q), _f64(coord)
    N = np.zeros(3, np.int64)
    nb = np.zeros(3, np.int64)
    nd = C.c_int()
    rc = lib().npo_rsem_static(len(q), _ptr(q), _ptr(coord), int(Bn), float(alpha), float(tol),
                               float(Rdir), int(nDeriv), _ptr(N), _ptr(nb), C.byref(nd))
    if rc != 0:
        raise ValueError("no RSEM row for Bn=%d alpha=%g" % (Bn, alpha))
    return tuple(int(x) for x in N), tuple(int(x) for x in nb), nd.value


def rsem_static_pred_error(q, coord, Bn, alpha, Nder, Rdir, N):
    q, coord = _f64(q), _f64(coord)
    Nn = np.ascontiguousarray(N, dtype=np.int64)
    e, rel = C.c_double(), C.c_int()
    rc = lib().npo_rsem_static_pred_error(len(q), _ptr(q), _ptr(coord), int(Bn), float(alpha),
                                          int(Nder), float(Rdir), _ptr(Nn), C.byref(e), C.byref(rel))
    if rc != 0:
        raise ValueError("no RSEM row")
    return e.value, bool(rel.value)


def rsem_helmholtz(qc, coord, Bn, k0, tol, Rdir, nDeriv=-1):
    qf, coord = _c128_as_f64(qc), _f64(coord)
    N = np.zeros(3, np.int64)
    nb = np.zeros(3, np.int64)
    nd = C.c_int()
    k0 = complex(k0)
    rc = lib().npo_rsem_helmholtz(len(coord), _ptr(qf), _ptr(coord), int(Bn), k0.real, k0.imag,
                                  float(tol), float(Rdir), int(nDeriv), _ptr(N), _ptr(nb),
                                  C.byref(nd))
    if rc != 0:
        raise ValueError("Helmholtz RSEM failed (%d)" % rc)
    return tuple(int(x) for x in N), tuple(int(x) for x in nb), nd.value


def cell_index(coord, Rdir, nNeigh=2, nCellClust1D=3):
    coord = _f64(coord)
    n = len(coord)
    nA = np.zeros(3, np.int64)
    tA = np.zeros(n, np.uint64)
    tB = np.zeros(n, np.uint64)
    rc = lib().npo_cell_index(n, _ptr(coord), float(Rdir), int(nNeigh), int(nCellClust1D),
                              _ptr(nA), _ptr(tA), _ptr(tB))
    if rc != 0:
        raise ValueError("point outside box")
    return nA, tA, tB


def fft1d(z, sign=-1):
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    lib().npo_fft1d(len(z), _ptr(z), int(sign))
    return z


def fft3d(z, sign=-1):
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    n1, n2, n3 = z.shape
    lib().npo_fft3d(n1, n2, n3, _ptr(z), int(sign))
    return z


def fgrid_compact(k, N, a, dl, X):
    N = np.ascontiguousarray(N, dtype=np.int64)
    F = np.zeros((N[0] // 2 + 1, N[1] // 2 + 1, N[2] // 2 + 1), np.complex128)
    rc = lib().npo_fgrid_compact(C.byref(k), _ptr(N), _ptr(_f64(a)), _ptr(_f64(dl)),
                                 _ptr(_f64(X)), _ptr(F))
    if rc != 0:
        raise ValueError("F grid failed")
    return F


def lambda2(N, Bn):
    out = np.zeros(N)
    lib().npo_lambda2(int(N), int(Bn), _ptr(out))
    return out


def spread_basic(q, coord, N, L, Bn, R):
    cx = np.iscomplexobj(q)
    qq = _c128_as_f64(q) if cx else _f64(q)
    coord = _f64(coord)
    N = np.ascontiguousarray(N, dtype=np.int64)
    Q = np.zeros((N[0] // 2, N[1] // 2, N[2] // 2), np.complex128)
    rc = lib().npo_spread_basic(len(coord), _ptr(qq), int(cx), _ptr(coord), _ptr(N),
                                _ptr(_f64(L)), int(Bn), _ptr(_f64(R)), _ptr(Q))
    if rc != 0:
        raise ValueError("spline window outside the compact grid")
    return Q


def spline_start(coord, N, L, Bn, R):
    """mStart[n][3]: first grid index of each charge's spline window"""
    coord = _f64(coord)
    N = np.ascontiguousarray(N, dtype=np.int64)
    ms = np.zeros((len(coord), 3), np.int64)
    rc = lib().npo_spline_start(len(coord), _ptr(coord), _ptr(N), _ptr(_f64(L)), int(Bn),
                                _ptr(_f64(R)), _ptr(ms))
    if rc != 0:
        raise ValueError("spline window outside the compact grid")
    return ms


def theta_8parity(N, Bn, F, Q):
    N = np.ascontiguousarray(N, dtype=np.int64)
    F = np.ascontiguousarray(F, dtype=np.complex128)
    Q = np.ascontiguousarray(Q, dtype=np.complex128)
    T = np.zeros_like(Q)
    lib().npo_theta_8parity(_ptr(N), int(Bn), _ptr(F), _ptr(Q), _ptr(T))
    return T


def theta_fullF_reverse(N, Bn, F, Q, rev=(0, 0, 0)):
    """8-parity theta with the full Fourier grid F[N1][N2][N3] and per-axis
    sign reversal (src/RecSumV.cpp:851-1009)"""
    N = np.ascontiguousarray(N, dtype=np.int64)
    F = np.ascontiguousarray(F, dtype=np.complex128)
    Q = np.ascontiguousarray(Q, dtype=np.complex128)
    rv = np.ascontiguousarray([int(bool(r)) for r in rev], dtype=np.int32)
    T = np.zeros_like(Q)
    lib().npo_theta_fullF_reverse(_ptr(N), int(Bn), _ptr(F), _ptr(Q), _ptr(rv), _ptr(T))
    return T


def fgrid_fulltrans(k, N, a, dl, X, R0, part=FULL):
    """real-space translated kernel grid f(r) = part(|r + R0|), [N1][N2][N3]
    complex (src/RecSumGrid.cpp:663-794)"""
    N = np.ascontiguousarray(N, dtype=np.int64)
    f = np.zeros(tuple(int(x) for x in N), np.complex128)
    rc = lib().npo_fgrid_fulltrans(C.byref(k), int(part), _ptr(N), _ptr(_f64(a)), _ptr(_f64(dl)),
                                   _ptr(_f64(X)), _ptr(_f64(R0)), _ptr(f))
    if rc == -2:
        raise ValueError("box-box separation R0 is too small")
    if rc != 0:
        raise ValueError("translated kernel grid failed")
    return f


def min_distance_r0(R0, l):
    return lib().npo_min_distance_r0(_ptr(_f64(R0)), _ptr(_f64(l)))


def boxpair(k, N, Bn, coordT, coordS, qS, dl, a=(4.0, 4.0, 4.0), rev=(0, 0, 0), part=FULL):
    """V at the targets from the sources through the translated full grid:
    both sets in one local frame (each set relative to the centre of its own
    bounding box), the kernel grid of R0 = cT - cS (reversed per rev),
    spread -> theta_fullF_reverse -> interpolate.  V[nT][4] complex.

    The grid index of a charge runs against its position (the window start
    is -floor(w), src/RecSumSupportFunctions.h:123-127), so the correlation
    over grid indices that theta computes, sum_n Q(n) f(n - p), evaluates f
    at the target-minus-source separation: V_t = sum_s q_s func(|t - s +
    R0|) in local coordinates, which is the global separation for
    R0 = cT - cS."""
    coordT, coordS = _f64(coordT), _f64(coordS)
    _, cT = box_dim_center(coordT)
    _, cS = box_dim_center(coordS)
    lt, ls = coordT - cT, coordS - cS
    X0, R0c = box_dim_center(np.vstack([lt, ls]))
    X, R = pme_corr_box(X0, R0c, N, dl, Bn)
    L = 2 * (X + np.asarray(dl, float))
    f = fgrid_fulltrans(k, N, a, dl, X, cT - cS, part)
    F = fft3d(f, -1)
    Q = spread_basic(qS, ls, N, L, Bn, R)
    T = theta_fullF_reverse(N, Bn, F, Q, rev)
    return interp_basic(lt, T, N, L, Bn, R)


def interp_basic(coord, theta, N, L, Bn, R):
    coord = _f64(coord)
    N = np.ascontiguousarray(N, dtype=np.int64)
    th = np.ascontiguousarray(theta, dtype=np.complex128)
    V1 = np.zeros((len(coord), 4), np.complex128)
    rc = lib().npo_interp_basic(len(coord), _ptr(coord), _ptr(th), _ptr(N), _ptr(_f64(L)),
                                int(Bn), _ptr(_f64(R)), _ptr(V1))
    if rc != 0:
        raise ValueError("spline window outside the compact grid")
    return V1


def _vshape(k, n):
    return (n, 4) if k.kind != HELMHOLTZ else (n, 4)


def direct_sr(k, coord, q):
    coord = _f64(coord)
    cx = k.kind in _COMPLEX_KINDS
    qq = _c128_as_f64(q) if cx else _f64(q)
    V = np.zeros((len(coord), 4), np.complex128 if cx else np.float64)
    lib().npo_direct_sr(C.byref(k), len(coord), _ptr(coord), _ptr(qq), _ptr(V))
    return V


def direct_refpairs(k, coord, q, nNeigh=2, nCellClust1D=3):
    """direct sum over the reference's interacting cell pairs, no cutoff
    (LaplaceOrig's erfc kernel and the rFunc user kernels)"""
    coord = _f64(coord)
    cx = k.kind in _COMPLEX_KINDS
    qq = _c128_as_f64(q) if cx else _f64(q)
    V = np.zeros((len(coord), 4), np.complex128 if cx else np.float64)
    rc = lib().npo_direct_refpairs(C.byref(k), len(coord), _ptr(coord), _ptr(qq), int(nNeigh),
                                   int(nCellClust1D), _ptr(V))
    if rc != 0:
        raise ValueError("direct_refpairs failed (%d)" % rc)
    return V


def calcv_ref(k, N, Bn, coord, q, nblk=(20, 20, 20), dl=None, a=(4.0, 4.0, 4.0),
              nNeigh=2, nCellClust1D=3):
    """CalcV with the reference's direct-sum pair set: the reciprocal sum and
    self term of calcv, and the direct sum over the interacting cell pairs
    (NPME_InterfaceReal_Laplace_Original; GenFunc with the rFunc kernels)"""
    V = calcv(k, N, Bn, coord, q, nblk=nblk, dl=dl, a=a)
    return V - direct_sr(k, coord, q) + direct_refpairs(k, coord, q, nNeigh, nCellClust1D)


calcv_laplace_orig = calcv_ref


def exact(k, coord, q, idx=None):
    coord = _f64(coord)
    cx = k.kind in _COMPLEX_KINDS
    qq = _c128_as_f64(q) if cx else _f64(q)
    if idx is None:
        nt, ip = len(coord), None
    else:
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        nt, ip = len(idx), _ptr(idx)
    V = np.zeros((nt, 4), np.complex128 if cx else np.float64)
    lib().npo_exact(C.byref(k), len(coord), _ptr(coord), _ptr(qq), nt, ip, _ptr(V))
    return V


def count_pairs(coord, Rdir):
    coord = _f64(coord)
    return lib().npo_count_pairs(len(coord), _ptr(coord), float(Rdir))


def calcv(k, N, Bn, coord, q, nblk=(20, 20, 20), dl=None, a=(4.0, 4.0, 4.0)):
    p = RecSumParm()
    for d in range(3):
        p.N[d] = int(N[d])
        p.nblk[d] = int(nblk[d])
        p.del_[d] = float(k.Rdir if dl is None else dl[d])
        p.a[d] = float(a[d])
    p.Bn = int(Bn)
    coord = _f64(coord)
    cx = k.kind in _COMPLEX_KINDS
    qq = _c128_as_f64(q) if cx else _f64(q)
    V = np.zeros((len(coord), 4), np.complex128 if cx else np.float64)
    rc = lib().npo_calcv(C.byref(k), C.byref(p), len(coord), _ptr(qq), _ptr(coord), _ptr(V))
    if rc != 0:
        raise ValueError("oracle CalcV failed")
    return V


def last_calcv_times():
    t = np.zeros(3)
    lib().npo_last_calcv_times(_ptr(t))
    return {"direct": t[0], "setup": t[1], "recip": t[2]}


def compare_v(Vref, V):
    cx = np.iscomplexobj(Vref) or np.iscomplexobj(V)
    if cx:
        a = _c128_as_f64(Vref)
        b = _c128_as_f64(V)
    else:
        a, b = _f64(Vref), _f64(V)
    out = np.zeros(5)
    lib().npo_compare_v(len(Vref), int(cx), _ptr(a), _ptr(b), _ptr(out))
    return {"errV": out[0], "errdV": out[1], "magV": out[2], "magdV": out[3],
            "relV": out[0] / out[2] if out[2] else 0.0,
            "reldV": out[1] / out[3] if out[3] else 0.0, "maxrelV": out[4]}


def splitmix_uniform(seed, n):
    out = np.zeros(n)
    lib().npo_splitmix_uniform(int(seed), int(n), _ptr(out))
    return out
