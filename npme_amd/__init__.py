"""MI355X NPME potential evaluation -- Python mirror of the reference's
library interface (src/Interface.h), over the C ABI of libnpme_amd.so.

The names follow the reference: NPME_KeywordInput, NPME_AddMissingKeywords,
NPME_InterfaceReal_Laplace_DM, NPME_InterfaceReal_GenFunc (r^alpha),
NPME_InterfaceComplex_Helmholtz_DM, NPME_Interface_SetUpRealKernel /
ComplexKernel, with SetUp / ResetCoord / CalcV / CalcV_exact / Print and the
Get_N1 ... getters.  Methods return bool like the reference; the reason of a
failure is in npme_last_error() (the reference prints it on std::cout).

Charges and potentials are numpy arrays: coord[n][3], charge[n] (real) or
complex128[n], V[n][4] float64 or complex128.  The library needs the HIP
build (python -m npme_amd.build); importing without it raises, there is no
CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnpme_amd.so")

NPME_OK = 0
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libnpme_amd.so is not built: run `python -m npme_amd.build`")
        L = C.CDLL(LIB_PATH)
        p, s, i, d = C.c_void_p, C.c_size_t, C.c_int, C.c_double
        L.npme_last_error.restype = C.c_char_p
        L.npme_keywords_create.restype = p
        L.npme_keywords_destroy.argtypes = [p]
        L.npme_keywords_set.argtypes = [p, C.c_char_p, C.c_char_p]
        L.npme_keywords_get.argtypes = [p, C.c_char_p, C.c_char_p, s]
        L.npme_keywords_read_file.argtypes = [p, C.c_char_p]
        L.npme_keywords_count.argtypes = [p]
        L.npme_keywords_key.argtypes = [p, i, C.c_char_p, s]
        L.npme_add_missing_keywords.argtypes = [p, s, i, p, p, p]
        L.npme_create.argtypes = [C.POINTER(p), p, s, p]
        L.npme_reset_coord.argtypes = [p, s, p]
        L.npme_destroy.argtypes = [p]
        L.npme_calcv_real.argtypes = [p, p, s, p, p]
        L.npme_calcv_complex.argtypes = [p, p, s, p, p]
        L.npme_calcv_exact_real.argtypes = [p, p, s, p, p]
        L.npme_calcv_exact_complex.argtypes = [p, p, s, p, p]
        L.npme_calcv_exact_subset.argtypes = [p, p, s, p, p]
        L.npme_calcv_device.argtypes = [p, p, p]
        L.npme_get_params.argtypes = [p, p, p, p, p, p, p]
        L.npme_print.argtypes = [p, C.c_char_p, s]
        L.npme_set_profiling.argtypes = [p, i]
        L.npme_set_calcv_graph.argtypes = [p, i]
        L.npme_count_pairs.argtypes = [p, p]
        L.npme_count_pairs_noalloc.argtypes = [p, p]
        L.npme_create_genfunc_real.argtypes = [C.POINTER(p), p, s, p, p, p, p]
        L.npme_create_genfunc_complex.argtypes = [C.POINTER(p), p, s, p, p, p, p]
        L.npme_get_table_error.argtypes = [p, p]
        L.npme_create_boxpair.argtypes = [C.POINTER(p), p, s, p, s, p]
        L.npme_create_boxpair_genfunc_real.argtypes = [C.POINTER(p), p, s, p, s, p, p]
        L.npme_create_boxpair_genfunc_complex.argtypes = [C.POINTER(p), p, s, p, s, p, p]
        L.npme_boxpair_reset_coord.argtypes = [p, s, p, s, p]
        L.npme_boxpair_calcv_real.argtypes = [p, p, s, s, p, p]
        L.npme_boxpair_calcv_complex.argtypes = [p, p, s, s, p, p]
        L.npme_boxpair_separation.argtypes = [p, p, p]
        L.npme_recsumgrid_fulltrans_func.argtypes = [p, C.c_long, d, d, C.c_long, d, d, C.c_long, d,
                                                     d, d, d, d, p, p]
        L.npme_recsumgrid_fulltrans_func_user.argtypes = [p, C.c_long, d, d, C.c_long, d, d,
                                                          C.c_long, d, d, d, d, d, p, p, p]
        L.npme_rstp_calctheta_fullF_reversesign.argtypes = [p, p, p, C.c_long, C.c_long, C.c_long,
                                                            p, p, p, i, i, i]
        L.npme_get_stage_times.argtypes = [p, p, i]
        L.npme_get_indices.argtypes = [p, s, p, p, p, p]
        L.npme_debug_fail_alloc.argtypes = [C.c_long]
        L.npme_debug_fail_alloc.restype = None
        L.npme_server_create.argtypes = [C.POINTER(p), i, d]
        L.npme_server_submit.argtypes = [p, C.POINTER(_Request), C.POINTER(C.c_long)]
        L.npme_server_wait.argtypes = [p, C.c_long, C.POINTER(_RequestInfo)]
        L.npme_server_clock.argtypes = [p]
        L.npme_server_clock.restype = d
        L.npme_server_destroy.argtypes = [p]
        L.npme_server_set_batching.argtypes = [p, i, s]
        L.npme_host_alloc.argtypes = [s, C.POINTER(p)]
        L.npme_host_free.argtypes = [p]
        pp = C.POINTER(p)
        L.npme_read_coord_charge.argtypes = [C.c_char_p, C.POINTER(s), C.POINTER(i), pp, pp]
        L.npme_write_coord_charge.argtypes = [C.c_char_p, s, i, p, p]
        L.npme_write_v.argtypes = [C.c_char_p, s, i, p, i]
        L.npme_read_v.argtypes = [C.c_char_p, C.POINTER(s), C.POINTER(i), pp]
        L.npme_free.argtypes = [p]
        L.npme_compare_v.argtypes = [s, i, p, p, p]
        _lib = L
    return _lib


class _Request(C.Structure):
    _fields_ = [("funcType", C.c_char_p), ("alpha", C.c_double), ("k0_r", C.c_double),
                ("k0_i", C.c_double), ("tol", C.c_double), ("Rdir", C.c_double),
                ("BnOrder", C.c_long), ("fp64", C.c_int), ("nCharge", C.c_size_t),
                ("coord", C.c_void_p), ("charge", C.c_void_p), ("V", C.c_void_p),
                ("tArrival", C.c_double)]


class _RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("tArrival", C.c_double), ("tSubmit", C.c_double),
                ("tStart", C.c_double), ("tEnd", C.c_double), ("N", C.c_long * 3),
                ("deviceGB", C.c_double), ("msKeywords", C.c_double), ("msSetup", C.c_double),
                ("msCopyIn", C.c_double), ("msCalc", C.c_double), ("msCopyOut", C.c_double),
                ("reservedGB", C.c_double), ("batch", C.c_int)]


def npme_last_error():
    return lib().npme_last_error().decode()


class NPMEError(RuntimeError):
    def __init__(self, code, where):
        super().__init__("%s failed (status %d): %s" % (where, code, npme_last_error()))
        self.code = code


def _check(rc, where):
    if rc != NPME_OK:
        raise NPMEError(rc, where)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _coord(coord):
    c = np.ascontiguousarray(coord, dtype=np.float64)
    if c.ndim != 2 or c.shape[1] != 3:
        c = c.reshape(-1, 3)
    return c


# ----------------------------------------------------------------- keywords
class NPME_KeywordInput:
    """Keyword map (src/ReadPrint.h:28-102)."""

    def __init__(self, keywordFile=None, **kw):
        self._L = lib()  # kept for __del__, which may run during interpreter shutdown
        self._h = C.c_void_p(self._L.npme_keywords_create())
        if keywordFile is not None:
            _check(lib().npme_keywords_read_file(self._h, keywordFile.encode()), "read keywords")
        for k, v in kw.items():
            self.AddKeyword(k, v)

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.npme_keywords_destroy(self._h)
            self._h = None

    def AddKeyword(self, key, value):
        if isinstance(value, float):
            value = repr(value)
        _check(lib().npme_keywords_set(self._h, key.encode(), str(value).encode()), "AddKeyword")

    def GetValue(self, key, default=None):
        buf = C.create_string_buffer(512)
        if lib().npme_keywords_get(self._h, key.encode(), buf, 512):
            return buf.value.decode()
        return default

    def keys(self):
        buf = C.create_string_buffer(512)
        out = []
        for i in range(lib().npme_keywords_count(self._h)):
            lib().npme_keywords_key(self._h, i, buf, 512)
            out.append(buf.value.decode())
        return out

    def as_dict(self):
        return {k: self.GetValue(k) for k in self.keys()}

    def PrintKeywords(self):
        return "\n".join("  %-14s %s" % (k, v) for k, v in self.as_dict().items())


def NPME_AddMissingKeywords(keyword, nCharge, coord, chargeReal=None, chargeComplex=None):
    """Defaults, validation and the RSEM grid choice (src/Interface.cpp:53-517).
    Returns True/False like the reference."""
    coord = _coord(coord)
    isReal = chargeComplex is None
    qr = None if chargeReal is None else np.ascontiguousarray(chargeReal, dtype=np.float64)
    qc = None if chargeComplex is None else np.ascontiguousarray(chargeComplex, dtype=np.complex128)
    rc = lib().npme_add_missing_keywords(keyword._h, int(nCharge), int(isReal), _ptr(coord),
                                         None if qr is None else _ptr(qr),
                                         None if qc is None else _ptr(qc))
    return rc == NPME_OK


# ----------------------------------------------------------------- interfaces
class _InterfaceBase:
    _complex = False
    _funcType = None

    def __init__(self):
        self._L = lib()  # kept for __del__, which may run during interpreter shutdown
        self._h = None
        self.nCharge = 0

    def __del__(self):
        self._free()

    def _free(self):
        if getattr(self, "_h", None):
            self._L.npme_destroy(self._h)
            self._h = None

    def SetUp(self, keyword, nCharge, coord, printLog=False):
        ft = keyword.GetValue("funcType")
        if self._funcType is not None and ft not in self._funcType:
            return False
        self._free()
        coord = _coord(coord)
        h = C.c_void_p()
        rc = lib().npme_create(C.byref(h), keyword._h, int(nCharge), _ptr(coord))
        if rc != NPME_OK:
            return False
        self._h = h
        self.nCharge = int(nCharge)
        return True

    def ResetCoord(self, nCharge, coord, printLog=False):
        coord = _coord(coord)
        return lib().npme_reset_coord(self._h, int(nCharge), _ptr(coord)) == NPME_OK

    def _params(self):
        N = np.zeros(3, np.int64)
        nb = np.zeros(3, np.int64)
        Bn = C.c_long()
        nd = C.c_int()
        R = C.c_double()
        gb = C.c_double()
        _check(lib().npme_get_params(self._h, _ptr(N), _ptr(nb), C.byref(Bn), C.byref(nd),
                                     C.byref(R), C.byref(gb)), "get_params")
        return N, nb, Bn.value, nd.value, R.value, gb.value

    def Get_N1(self): return int(self._params()[0][0])
    def Get_N2(self): return int(self._params()[0][1])
    def Get_N3(self): return int(self._params()[0][2])
    def Get_n1(self): return int(self._params()[1][0])
    def Get_n2(self): return int(self._params()[1][1])
    def Get_n3(self): return int(self._params()[1][2])
    def Get_BnOrder(self): return int(self._params()[2])
    def Get_Nder(self): return int(self._params()[3])
    def GetRdir(self): return float(self._params()[4])
    def GetNumCharge(self): return self.nCharge
    def DeviceGB(self): return float(self._params()[5])

    def Print(self):
        buf = C.create_string_buffer(4096)
        lib().npme_print(self._h, buf, 4096)
        return buf.value.decode()

    def _qarr(self, charge):
        if self._complex:
            return np.ascontiguousarray(charge, dtype=np.complex128)
        return np.ascontiguousarray(charge, dtype=np.float64)

    def _varr(self, V, n):
        dt = np.complex128 if self._complex else np.float64
        if V is None:
            return np.zeros((n, 4), dt)
        if V.dtype != dt or not V.flags.c_contiguous or V.size != 4 * n:
            raise ValueError("V must be a contiguous %s array of shape (n, 4)" % np.dtype(dt).name)
        return V

    def CalcV(self, V, nCharge, charge, coord, printLog=False):
        """Fills V[n][4]; returns True/False (src/Interface.h:373-375)."""
        q = self._qarr(charge)
        V = self._varr(V, nCharge)
        f = lib().npme_calcv_complex if self._complex else lib().npme_calcv_real
        return f(self._h, _ptr(V), int(nCharge), _ptr(q), None) == NPME_OK

    def CalcV_exact(self, V, nCharge, charge, coord, printLog=False):
        q = self._qarr(charge)
        V = self._varr(V, nCharge)
        f = lib().npme_calcv_exact_complex if self._complex else lib().npme_calcv_exact_real
        return f(self._h, _ptr(V), int(nCharge), _ptr(q), None) == NPME_OK

    # conveniences (no reference counterpart)
    def calcv(self, charge):
        V = self._varr(None, self.nCharge)
        if not self.CalcV(V, self.nCharge, charge, None):
            raise NPMEError(-1, "CalcV")
        return V

    def calcv_exact_subset(self, charge, idx):
        q = self._qarr(charge)
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        V = self._varr(None, len(idx))
        _check(lib().npme_calcv_exact_subset(self._h, _ptr(V), len(idx), _ptr(idx), _ptr(q)),
               "exact_subset")
        return V

    def calcv_device(self, dV_ptr, dq_ptr):
        """Device pointers (e.g. torch tensor data_ptr()): charges in input
        order, V [n][4] FP64 (complex interleaved)."""
        _check(lib().npme_calcv_device(self._h, C.c_void_p(dV_ptr), C.c_void_p(dq_ptr)), "calcv_device")

    def count_pairs(self):
        """Ordered pairs i != j with r < Rdir (the direct-sum work)."""
        out = C.c_double()
        _check(lib().npme_count_pairs(self._h, C.byref(out)), "count_pairs")
        return int(out.value)

    def count_pairs_noalloc(self):
        """The same count through npme_count_pairs_noalloc (no device
        allocation; partials in the library's static device array)."""
        out = C.c_double()
        _check(lib().npme_count_pairs_noalloc(self._h, C.byref(out)), "count_pairs_noalloc")
        return int(out.value)

    def get_indices(self):
        """Integer indexing of the set-up, per charge in input order:
        (clusterKey[n], cellKey[n], mStart[n][3], order[n]) -- the embedded
        cluster and cell indices, the first grid index of each spline window
        and the working (sorted) order."""
        n = self.nCharge
        ka = np.zeros(n, np.uint64)
        kb = np.zeros(n, np.uint64)
        ms = np.zeros((n, 3), np.int64)
        order = np.zeros(n, np.int64)
        _check(lib().npme_get_indices(self._h, n, _ptr(ka), _ptr(kb), _ptr(ms), _ptr(order)),
               "get_indices")
        return ka, kb, ms, order

    def set_profiling(self, on=True):
        lib().npme_set_profiling(self._h, int(on))

    def set_calcv_graph(self, on):
        """Replay a captured graph for repeated CalcV calls (default on)."""
        _check(lib().npme_set_calcv_graph(self._h, 1 if on else 0), "set_calcv_graph")

    def stage_times(self):
        ms = np.zeros(16)
        k = lib().npme_get_stage_times(self._h, _ptr(ms), 16)
        return ms[:k]


class NPME_InterfaceReal_Laplace_DM(_InterfaceBase):
    _funcType = ("Laplace",)


# ----------------------------------------------------------------- user kernels
_PD = C.POINTER(C.c_double)
_VAL_R = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t, _PD, _PD, _PD)
_GRAD_R = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t, _PD, _PD, _PD, _PD)
_VAL_C = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t, _PD, _PD, _PD, _PD)
_GRAD_C = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t, _PD, _PD, _PD, _PD, _PD, _PD, _PD, _PD)


class _KfuncRealC(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("value", _VAL_R), ("grad", _GRAD_R)]


class _KfuncComplexC(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("value", _VAL_C), ("grad", _GRAD_C)]


def _callback(obj):
    """A ctypes callback running obj.Calc(N, *arrays) on views of the
    library's arrays.  An exception in Calc is kept on obj (raised again by
    SetUp) and the outputs are set to NaN, which the library refuses."""
    def fn(ctx, n, *ptrs):
        N = int(n)
        if N == 0:
            return
        arrs = [np.ctypeslib.as_array(p, (N,)) for p in ptrs]
        try:
            obj.Calc(N, *arrs)
        except BaseException as e:  # noqa: BLE001 -- reported through SetUp
            obj._calc_error = e
            for a in arrs:
                a[:] = np.nan
    return fn


class NPME_KfuncReal:
    """Base of a user-defined real kernel (src/KernelFunction.h:25-60).
    A subclass defines Calc for both of the reference's forms, told apart by
    the number of arrays (numpy float64 views of length N, changed in place):
      Calc(N, x_f0, y, z):          x_f0 holds x, and f(x, y, z) on return
      Calc(N, f0, x_fX, y_fY, z_fZ): f0 receives f; x_fX, y_fY, z_fZ hold
                                     x, y, z and the gradient on return
    and Print(os).  Calc may be called with several batches at once from the
    library's threads (one at a time, under the interpreter lock)."""
    _cstruct = _KfuncRealC

    def Calc(self, N, *arrays):
        raise NotImplementedError("NPME_KfuncReal.Calc")

    def Print(self, os=None):
        pass

    def _c(self):
        if getattr(self, "_cfun", None) is None:
            self._cfun = self._cstruct(None, _VAL_R(_callback(self)), _GRAD_R(_callback(self)))
        return self._cfun


class NPME_KfuncComplex(NPME_KfuncReal):
    """Base of a user-defined complex kernel (src/KernelFunction.h:62-116):
      Calc(N, x_f0_r, f0_i, y, z)
      Calc(N, f0_r, f0_i, x_fX_r, fX_i, y_fY_r, fY_i, z_fZ_r, fZ_i)"""
    _cstruct = _KfuncComplexC

    def _c(self):
        if getattr(self, "_cfun", None) is None:
            self._cfun = self._cstruct(None, _VAL_C(_callback(self)), _GRAD_C(_callback(self)))
        return self._cfun


class _GenFuncBase(_InterfaceBase):
    def SetUp(self, keyword, nCharge, coord, func=None, funcLR=None, funcSR=None,
              printLog=False):
        """SetUp(keyword, nCharge, coord, func, funcLR, funcSR)
        (src/Interface.h:291-296, :624-629).  With user kernels (subclasses
        of NPME_KfuncReal / NPME_KfuncComplex) the grid keywords must be set
        (N1-3, n1-3, BnOrder, Rdir ...).  Without them the built-in kernel
        of the keywords' funcType is used."""
        if isinstance(func, bool):  # SetUp(keyword, nCharge, coord, printLog)
            func, printLog = None, func
        if func is None and funcLR is None and funcSR is None:
            return _InterfaceBase.SetUp(self, keyword, nCharge, coord, printLog)
        fs = (func, funcLR, funcSR)
        if any(f is None for f in fs):
            return False
        self._free()
        coord = _coord(coord)
        for f in fs:
            f._calc_error = None
        cs = [f._c() for f in fs]
        h = C.c_void_p()
        create = lib().npme_create_genfunc_complex if self._complex else lib().npme_create_genfunc_real
        rc = create(C.byref(h), keyword._h, int(nCharge), _ptr(coord),
                    *[C.cast(C.pointer(c), C.c_void_p) for c in cs])
        for f in fs:
            if f._calc_error is not None:
                err, f._calc_error = f._calc_error, None
                if rc == NPME_OK:
                    lib().npme_destroy(h)
                raise err
        if rc != NPME_OK:
            return False
        self._h = h
        self._kfuncs = fs  # the callbacks must outlive the handle
        self.nCharge = int(nCharge)
        return True

    def table_error(self):
        """(f, f'/r) errors of the short-range table of user kernels."""
        e = np.zeros(2)
        _check(lib().npme_get_table_error(self._h, _ptr(e)), "table_error")
        return float(e[0]), float(e[1])


class NPME_InterfaceReal_GenFunc(_GenFuncBase):
    """Generic real kernel: user kernels, or the built-in r^alpha (funcType
    Ralpha) and Laplace."""
    _funcType = ("Ralpha", "Laplace")


class NPME_InterfaceComplex_GenFunc(_GenFuncBase):
    """Generic complex kernel (src/Interface.h:563-650): user kernels, or the
    built-in Helmholtz."""
    _complex = True
    _funcType = ("Helmholtz",)


# ------------------------------------------------------------ two-box grids
def _user_kernel_ptrs(func):
    """(real, complex) C callback pointers of an NPME_KfuncReal /
    NPME_KfuncComplex"""
    c = C.cast(C.pointer(func._c()), C.c_void_p)
    return (None, c) if isinstance(func, NPME_KfuncComplex) else (c, None)


def _raise_calc_error(func, rc, destroy=None):
    if func is not None and getattr(func, "_calc_error", None) is not None:
        err, func._calc_error = func._calc_error, None
        if rc == NPME_OK and destroy is not None:
            destroy()
        raise err


def NPME_RecSumGrid_FullTrans_Func(N, a, dl, X, func, R0):
    """f[N1][N2][N3] (complex): func(|r + R0|) on the full grid of the box
    [-X - del, X + del) per axis with the 8-image top-hat extension
    (src/RecSumGrid.cpp:663-794).  func is an NPME_KeywordInput naming a
    built-in kernel (funcType ..., sampled on the GPU in FP64) or an
    NPME_KfuncReal / NPME_KfuncComplex (sampled on the host)."""
    N = [int(x) for x in N]
    f = np.zeros(N, np.complex128)
    args = []
    for d in range(3):
        args += [N[d], float(a[d]), float(dl[d])]
    args += [float(X[0]), float(X[1]), float(X[2])]
    R = np.ascontiguousarray(R0, dtype=np.float64)
    if isinstance(func, NPME_KeywordInput):
        rc = lib().npme_recsumgrid_fulltrans_func(_ptr(f), *args, func._h, _ptr(R))
    else:
        func._calc_error = None
        rp, cp = _user_kernel_ptrs(func)
        rc = lib().npme_recsumgrid_fulltrans_func_user(_ptr(f), *args, rp, cp, _ptr(R))
        _raise_calc_error(func, rc)
    _check(rc, "NPME_RecSumGrid_FullTrans_Func")
    return f


def NPME_RSTP_CalcTheta_FullF_ReverseSign(F, Qr, lamda, reverseSign=(0, 0, 0)):
    """Compact theta [N1/2][N2/2][N3/2] (src/RecSumV.cpp:943-1009) on the
    GPU: F the full forward DFT [N1][N2][N3] of the kernel grid, Qr the
    compact charge grid (DFT order), lamda = (|lambda_1|^2, |lambda_2|^2,
    |lambda_3|^2); reverseSign uses F as the grid of R0 with those
    components negated."""
    F = np.ascontiguousarray(F, dtype=np.complex128)
    Q = np.ascontiguousarray(Qr, dtype=np.complex128)
    N1, N2, N3 = F.shape
    if Q.shape != (N1 // 2, N2 // 2, N3 // 2):
        raise ValueError("Qr must be [N1/2][N2/2][N3/2]")
    lam = [np.ascontiguousarray(l, dtype=np.float64) for l in lamda]
    for l, n in zip(lam, F.shape):
        if l.shape != (n,):
            raise ValueError("lamda arrays must have N1, N2, N3 values")
    T = np.zeros_like(Q)
    _check(lib().npme_rstp_calctheta_fullF_reversesign(
        _ptr(T), _ptr(F), _ptr(Q), N1, N2, N3, _ptr(lam[0]), _ptr(lam[1]), _ptr(lam[2]),
        int(bool(reverseSign[0])), int(bool(reverseSign[1])), int(bool(reverseSign[2]))),
        "NPME_RSTP_CalcTheta_FullF_ReverseSign")
    return T


class NPME_BoxPair:
    """V at the targets of one box from the sources of another, entirely on
    the translated kernel grid (no reference driver; the reference has the
    two functions above).  keyword: funcType etc. as for SetUp plus N1-N3;
    func: None (the keywords' built-in kernel) or a user NPME_KfuncReal /
    NPME_KfuncComplex, borrowed for the object's life."""

    def __init__(self, keyword, coordTarget, coordSource, func=None):
        self._L = lib()
        self._h = None
        cT, cS = _coord(coordTarget), _coord(coordSource)
        self.nTarget, self.nSource = len(cT), len(cS)
        h = C.c_void_p()
        if func is None:
            self._complex = keyword.GetValue("funcType") == "Helmholtz"
            rc = lib().npme_create_boxpair(C.byref(h), keyword._h, self.nTarget, _ptr(cT),
                                           self.nSource, _ptr(cS))
        else:
            self._complex = isinstance(func, NPME_KfuncComplex)
            func._calc_error = None
            create = (lib().npme_create_boxpair_genfunc_complex if self._complex
                      else lib().npme_create_boxpair_genfunc_real)
            rc = create(C.byref(h), keyword._h, self.nTarget, _ptr(cT), self.nSource, _ptr(cS),
                        C.cast(C.pointer(func._c()), C.c_void_p))
            _raise_calc_error(func, rc, lambda: lib().npme_destroy(h))
            self._kfunc = func
        _check(rc, "npme_create_boxpair")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.npme_destroy(self._h)
            self._h = None

    def calcv(self, chargeSource, reverseSign=(0, 0, 0)):
        dt = np.complex128 if self._complex else np.float64
        q = np.ascontiguousarray(chargeSource, dtype=dt)
        if q.shape != (self.nSource,):
            raise ValueError("chargeSource must have nSource values")
        V = np.zeros((self.nTarget, 4), dt)
        rv = (C.c_int * 3)(*[int(bool(r)) for r in reverseSign])
        f = lib().npme_boxpair_calcv_complex if self._complex else lib().npme_boxpair_calcv_real
        _check(f(self._h, _ptr(V), self.nTarget, self.nSource, _ptr(q), rv), "boxpair_calcv")
        return V

    def reset_coord(self, coordTarget, coordSource):
        cT, cS = _coord(coordTarget), _coord(coordSource)
        _check(lib().npme_boxpair_reset_coord(self._h, len(cT), _ptr(cT), len(cS), _ptr(cS)),
               "boxpair_reset_coord")
        self.nTarget, self.nSource = len(cT), len(cS)

    def separation(self):
        """(R0, rMin): R0 = target centre - source centre, f(r) = func(|r + R0|)"""
        R0 = np.zeros(3)
        r = C.c_double()
        _check(lib().npme_boxpair_separation(self._h, _ptr(R0), C.byref(r)), "boxpair_separation")
        return R0, r.value

    def set_calcv_graph(self, on):
        _check(lib().npme_set_calcv_graph(self._h, int(bool(on))), "set_calcv_graph")


class NPME_InterfaceComplex_Helmholtz_DM(_InterfaceBase):
    _complex = True
    _funcType = ("Helmholtz",)


def NPME_Interface_SetUpRealKernel(keyword, nCharge, coord, useGenericKernel=False,
                                   printLog=False):
    """Allocates and sets up the interface class for a real kernel
    (src/Interface.h:800-815).  Returns the object or None."""
    ft = keyword.GetValue("funcType")
    cls = NPME_InterfaceReal_Laplace_DM if (ft == "Laplace" and not useGenericKernel) \
        else NPME_InterfaceReal_GenFunc
    obj = cls()
    return obj if obj.SetUp(keyword, nCharge, coord, printLog=printLog) else None


def NPME_Interface_SetUpComplexKernel(keyword, nCharge, coord, useGenericKernel=False,
                                      printLog=False):
    obj = NPME_InterfaceComplex_Helmholtz_DM()
    return obj if obj.SetUp(keyword, nCharge, coord, printLog=printLog) else None


# ----------------------------------------------------------------- serving
class NPME_Server:
    """Persistent request loop on the current GPU (npme_server_* in
    include/npme_amd.h): nSlots requests run concurrently, each admitted
    when a slot and its device-memory estimate are free.  No reference
    counterpart."""

    def __init__(self, nSlots=4, memGB=0.0, maxBatch=None, batchMaxN=None):
        self._L = lib()  # kept for __del__, which may run during interpreter shutdown
        self._s = C.c_void_p()
        _check(lib().npme_server_create(C.byref(self._s), int(nSlots), float(memGB)), "server_create")
        self._live = {}
        if maxBatch is not None or batchMaxN is not None:
            self.set_batching(8 if maxBatch is None else maxBatch,
                              32768 if batchMaxN is None else batchMaxN)

    def set_batching(self, maxBatch, batchMaxN=32768):
        """Small queued requests (at most batchMaxN charges) run in batches of
        up to maxBatch with one graph launch for their CalcVs; 1 turns it off."""
        _check(lib().npme_server_set_batching(self._s, int(maxBatch), int(batchMaxN)),
               "server_set_batching")

    def clock(self):
        """Seconds since the server was created (the clock of tArrival)."""
        return lib().npme_server_clock(self._s)

    def submit(self, funcType, coord, charge, tol=1e-6, Rdir=8.0, BnOrder=0, alpha=0.0, k0=0.0,
               fp64=False, tArrival=-1.0, V=None):
        """Queues a request; returns its ticket.  V: optional output array
        [n][4] (float64, or complex128 for Helmholtz, C-contiguous) that
        wait() returns filled; by default one is allocated."""
        coord = _coord(coord)
        cx = funcType == "Helmholtz"
        q = np.ascontiguousarray(charge, dtype=np.complex128 if cx else np.float64)
        n = len(q)
        dt = np.complex128 if cx else np.float64
        if V is None:
            V = np.empty((n, 4), dt)  # fully written by the server
        elif V.shape != (n, 4) or V.dtype != dt or not V.flags.c_contiguous:
            raise ValueError("V must be a C-contiguous [n][4] %s array" % np.dtype(dt).name)
        ft = funcType.encode()
        k0 = complex(k0)
        r = _Request(ft, float(alpha), k0.real, k0.imag, float(tol), float(Rdir), int(BnOrder),
                     int(bool(fp64)), n, coord.ctypes.data, q.ctypes.data, V.ctypes.data,
                     float(tArrival))
        t = C.c_long()
        _check(lib().npme_server_submit(self._s, C.byref(r), C.byref(t)), "server_submit")
        self._live[t.value] = (coord, q, V, ft)
        return t.value

    def wait(self, ticket):
        """Blocks until the request is done: returns (V, info dict)."""
        info = _RequestInfo()
        rc = lib().npme_server_wait(self._s, int(ticket), C.byref(info))
        _, _, V, _ = self._live.pop(ticket)
        if rc != NPME_OK:
            raise NPMEError(rc, "request %d" % ticket)
        return V, {"tArrival": info.tArrival, "tSubmit": info.tSubmit, "tStart": info.tStart,
                   "tEnd": info.tEnd, "grid": tuple(info.N), "deviceGB": info.deviceGB,
                   "reservedGB": info.reservedGB, "batch": info.batch,
                   "ms": {"keywords": info.msKeywords, "setup": info.msSetup,
                          "copy_in": info.msCopyIn, "calcv": info.msCalc,
                          "copy_out": info.msCopyOut}}

    def close(self):
        if getattr(self, "_s", None):
            self._L.npme_server_destroy(self._s)
            self._s = None
            self._live.clear()

    def __del__(self):
        self.close()


class _PinnedBlock:
    """Owner of one page-locked host allocation (npme_host_alloc); freed
    with the last array that views it."""

    def __init__(self, nbytes):
        self._L = lib()
        self.ptr = C.c_void_p()
        _check(self._L.npme_host_alloc(int(nbytes), C.byref(self.ptr)), "host_alloc")
        self.nbytes = int(nbytes)

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self._L.npme_host_free(self.ptr)
            self.ptr = None


def pinned_empty(shape, dtype=np.float64):
    """Uninitialised numpy array in page-locked host memory, for request
    buffers whose device copies should run at DMA speed.  Allocate such
    buffers once and reuse them: releasing page-locked memory may wait for
    the device."""
    dtype = np.dtype(dtype)
    count = int(np.prod(shape))
    blk = _PinnedBlock(count * dtype.itemsize)
    raw = (C.c_char * max(blk.nbytes, 1)).from_address(blk.ptr.value)
    raw._npme_owner = blk  # the buffer keeps its allocation alive
    a = np.frombuffer(raw, dtype=dtype, count=count).reshape(shape)
    return a


# ----------------------------------------------------------------- files
# The on-disk formats of the command-line tools (src/ReadPrint.cpp:234-739)
# and the compareV metric (src/SupportFunctions.cpp:88-305), through the
# library's readers and writers.
def _take(ptr, count, dtype=np.float64):
    a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=(count,)).copy()
    lib().npme_free(ptr)
    return a


def NPME_ReadSingleBoxChargeCoord(filename):
    """(coord[n][3], charge) of an @Coord / @ChargeReal|@ChargeComplex file;
    charge is float64[n] or complex128[n]."""
    L = lib()
    n, real = C.c_size_t(), C.c_int()
    pc, pq = C.c_void_p(), C.c_void_p()
    _check(L.npme_read_coord_charge(os.fsencode(filename), C.byref(n), C.byref(real), C.byref(pc),
                                    C.byref(pq)), "NPME_ReadSingleBoxChargeCoord")
    coord = _take(pc, 3 * n.value).reshape(-1, 3)
    q = _take(pq, (1 if real.value else 2) * n.value)
    return coord, (q if real.value else q.view(np.complex128))


def NPME_PrintSingleBoxChargeCoord(filename, coord, charge):
    c = _coord(coord)
    q = np.asarray(charge)
    real = not np.iscomplexobj(q)
    qd = np.ascontiguousarray(q, dtype=np.float64 if real else np.complex128)
    if qd.shape[0] != c.shape[0]:
        raise ValueError("coord and charge sizes differ")
    _check(lib().npme_write_coord_charge(os.fsencode(filename), c.shape[0], int(real), _ptr(c),
                                         _ptr(qd)), "NPME_PrintSingleBoxChargeCoord")


def NPME_PrintSingleBox_V(filename, V, printHighPrecision=False):
    """@Vreal / @Vcomplex file of V[n][4] (float64 or complex128)."""
    real = not np.iscomplexobj(V)
    v = np.ascontiguousarray(V, dtype=np.float64 if real else np.complex128).reshape(-1, 4)
    _check(lib().npme_write_v(os.fsencode(filename), v.shape[0], int(real), _ptr(v),
                              int(bool(printHighPrecision))), "NPME_PrintSingleBox_V")


def NPME_ReadSingleBox_V(filename):
    L = lib()
    n, real = C.c_size_t(), C.c_int()
    pv = C.c_void_p()
    _check(L.npme_read_v(os.fsencode(filename), C.byref(n), C.byref(real), C.byref(pv)),
           "NPME_ReadSingleBox_V")
    v = _take(pv, (4 if real.value else 8) * n.value)
    return v.reshape(-1, 4) if real.value else v.view(np.complex128).reshape(-1, 4)


def NPME_CompareV1Arrays(V, Vref):
    """compareV metric: dict with the absolute errors errorV, errordVdr and
    the averages V_mag, dVdr_mag of Vref (relative error = ratio)."""
    real = not (np.iscomplexobj(V) or np.iscomplexobj(Vref))
    dt = np.float64 if real else np.complex128
    a = np.ascontiguousarray(V, dtype=dt).reshape(-1, 4)
    b = np.ascontiguousarray(Vref, dtype=dt).reshape(-1, 4)
    if a.shape != b.shape:
        raise ValueError("V and Vref sizes differ")
    out = np.zeros(4)
    _check(lib().npme_compare_v(a.shape[0], int(real), _ptr(a), _ptr(b), _ptr(out)),
           "NPME_CompareV1Arrays")
    return {"errorV": out[0], "errordVdr": out[1], "V_mag": out[2], "dVdr_mag": out[3],
            "relV": out[0] / out[2], "reldV": out[1] / out[3]}
