"""MI355X-native BootCMatchGX single-GPU solve path.

Python mirror of the reference's settings/preconditioner/solver interface
(BCMGX/src/config/Params.h, src/preconditioner/prec_setup.h,
src/preconditioner/prec_apply.h, src/solver/solve.h) over the C ABI of
``lib/libbcmgx.so`` (declared in ``include/bcmgx.h``).  Every computation runs
in the HIP library; this module only marshals arguments.  If the library is
missing the import of the handles fails loudly -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import OK, MEM_HOST, MEM_DEVICE  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbcmgx.so")

_lib = None

_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_VP = C.c_void_p


class BcmgxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"bcmgx status {status}: {msg}")
        self.status = status


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C bootcmatchgx_amd/csrc` "
                           "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    sig = {
        "bcmgx_params_default": (C.c_int, [C.POINTER(abi.Params)]),
        "bcmgx_params_from_properties": (C.c_int, [C.c_char_p, C.POINTER(abi.Params)]),
        "bcmgx_params_set": (C.c_int, [C.POINTER(abi.Params), C.c_char_p, C.c_char_p]),
        "bcmgx_enum_from_string": (C.c_int, [C.c_char_p, C.c_char_p]),
        "bcmgx_enum_to_string": (C.c_char_p, [C.c_char_p, C.c_int]),
        "bcmgx_context_create": (C.c_int, [C.c_int, C.POINTER(_VP)]),
        "bcmgx_context_destroy": (C.c_int, [_VP]),
        "bcmgx_context_synchronize": (C.c_int, [_VP]),
        "bcmgx_error_string": (C.c_char_p, [C.c_int]),
        "bcmgx_last_error": (C.c_char_p, [_VP]),
        "bcmgx_device_alloc": (C.c_int, [_VP, C.c_size_t, C.POINTER(_VP)]),
        "bcmgx_device_free": (C.c_int, [_VP, _VP]),
        "bcmgx_memcpy": (C.c_int, [_VP, _VP, _VP, C.c_size_t, C.c_int]),
        "bcmgx_csr_create": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, _VP, _VP, _VP, C.c_int, C.POINTER(_VP)]),
        "bcmgx_csr_destroy": (C.c_int, [_VP]),
        "bcmgx_csr_info": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "bcmgx_csr_copy_to_host": (C.c_int, [_VP, _I32P, _I32P, _F64P]),
        "bcmgx_csr_update_values": (C.c_int, [_VP, _VP, C.c_int]),
        "bcmgx_csr_analyse": (C.c_int, [_VP, C.POINTER(C.c_double)]),
        "bcmgx_gen_laplacian3d": (C.c_int, [C.c_int] * 7 + [C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                             C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int)),
                                                             C.POINTER(C.POINTER(C.c_double))]),
        "bcmgx_gen_diffusion7p": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64,
                                            C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.POINTER(C.c_int)),
                                            C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_double))]),
        "bcmgx_read_mtx": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int)),
                                     C.POINTER(C.POINTER(C.c_double))]),
        "bcmgx_free": (None, [_VP]),
        "bcmgx_prec_setup": (C.c_int, [_VP, _VP, C.POINTER(abi.Params), C.POINTER(_VP)]),
        "bcmgx_prec_apply": (C.c_int, [_VP, _VP, _VP, C.c_int, C.c_int, C.c_int]),
        "bcmgx_prec_finalize": (C.c_int, [_VP]),
        "bcmgx_prec_get_info": (C.c_int, [_VP, C.POINTER(abi.PrecInfo)]),
        "bcmgx_prec_level_size": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "bcmgx_prec_level_matrix": (C.c_int, [_VP, C.c_int, _I32P, _I32P, _F64P]),
        "bcmgx_prec_level_prolongator": (C.c_int, [_VP, C.c_int, _I32P, _F64P]),
        "bcmgx_prec_hrc_levels": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_int)]),
        "bcmgx_prec_hrc_level_size": (C.c_int, [_VP, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "bcmgx_prec_hrc_level_matrix": (C.c_int, [_VP, C.c_int, C.c_int, _I32P, _I32P, _F64P]),
        "bcmgx_prec_hrc_level_prolongator": (C.c_int, [_VP, C.c_int, C.c_int, _I32P, _F64P]),
        "bcmgx_prec_afsai_factor": (C.c_int, [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int), _VP, _VP, _VP]),
        "bcmgx_solve": (C.c_int, [_VP, _VP, _VP, C.POINTER(abi.Params), _VP, _VP, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(abi.SolverOut)]),
        "bcmgx_solve_columns": (C.c_int, [_VP, C.c_int, _I32P, _F64P, _F64P]),
        "bcmgx_dump_info": (C.c_int, [C.c_char_p, C.POINTER(abi.Params), _VP, C.POINTER(abi.SolverOut)]),
        "bcmgx_spmv": (C.c_int, [_VP, _VP, _VP, _VP, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]),
        "bcmgx_spmm": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.c_int]),
        "bcmgx_spmm_noalloc": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                         C.c_int]),
        "bcmgx_spgemm": (C.c_int, [_VP, _VP, _VP, C.POINTER(_VP)]),
        "bcmgx_transpose": (C.c_int, [_VP, _VP, C.POINTER(_VP)]),
        "bcmgx_matching": (C.c_int, [_VP, _VP, _F64P, _I32P]),
        "bcmgx_make_prolongator": (C.c_int, [_VP, C.c_int, _I32P, _F64P, _I32P, _F64P, C.POINTER(C.c_int)]),
        "bcmgx_stream_copy_gbs": (C.c_int, [_VP, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
        "bcmgx_time_spmv": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "bcmgx_time_smoother": (C.c_int, [_VP, _VP, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "bcmgx_time_spgemm": (C.c_int, [_VP, _VP, _VP, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "bcmgx_sparse_outer_grad": (C.c_int, [_VP, _VP, _VP, _VP, C.c_int, _VP, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return _lib


def declared_symbols():
    """Names of the entry points declared in include/bcmgx.h (for the ABI check)."""
    import re

    hdr = os.path.join(os.path.dirname(_HERE), "include", "bcmgx.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"^(?:int|void|const char\s*\*)\s*\**\s*(bcmgx_\w+)\(", txt, re.M)))


def _check(status, ctx=None):
    if status != OK:
        L = lib()
        msg = L.bcmgx_last_error(ctx.h if ctx is not None else None) or L.bcmgx_error_string(status)
        raise BcmgxError(status, msg.decode() if isinstance(msg, bytes) else str(msg))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- settings ---
class Params:
    """Settings (struct params, Params.h:238-291) with reference defaults."""

    def __init__(self, **kw):
        self.s = abi.Params()
        _check(lib().bcmgx_params_default(C.byref(self.s)))
        for k, v in kw.items():
            self.set(k, v)

    @classmethod
    def init_from_properties_file(cls, path):
        """Params::initFromPropertiesFile (Params.cu:326-404)."""
        p = cls()
        _check(lib().bcmgx_params_from_properties(os.fsencode(path), C.byref(p.s)))
        return p

    def set(self, key, value):
        _check(lib().bcmgx_params_set(C.byref(self.s), key.encode(), str(value).encode()))
        return self

    def __getattr__(self, name):
        if name == "s":
            raise AttributeError(name)
        return getattr(self.s, name)

    def copy(self):
        q = Params()
        C.pointer(q.s)[0] = self.s
        return q


def enum_string(kind, value):
    r = lib().bcmgx_enum_to_string(kind.encode(), int(value))
    return r.decode() if r else None


# ----------------------------------------------------------------- handles ---
class Context:
    """One HIP device + stream (replaces Handles::init and cudaSetDevice)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().bcmgx_context_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().bcmgx_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().bcmgx_context_synchronize(self.h), self)

    def stream_copy_gbs(self, nbytes=1 << 30, reps=20):
        g = C.c_double()
        _check(lib().bcmgx_stream_copy_gbs(self.h, nbytes, reps, C.byref(g)), self)
        return g.value


class DeviceArray:
    """A float64 device buffer of a context (inputs kept resident in HBM)."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape))
        self.size = int(np.prod(self.shape))
        h = C.c_void_p()
        _check(lib().bcmgx_device_alloc(ctx.h, self.size * 8, C.byref(h)), ctx)
        self.ptr = h

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a, np.float64)
        _check(lib().bcmgx_memcpy(self.ctx.h, C.c_void_p(self.ptr.value + offset * 8), _ptr(a), a.size * 8, 0), self.ctx)

    def download(self):
        a = np.empty(self.size)
        _check(lib().bcmgx_memcpy(self.ctx.h, _ptr(a), self.ptr, self.size * 8, 1), self.ctx)
        return a

    def copy_from(self, other):
        _check(lib().bcmgx_memcpy(self.ctx.h, self.ptr, other.ptr, self.size * 8, 2), self.ctx)

    def at(self, offset):
        """device address of element `offset`"""
        return C.c_void_p(self.ptr.value + offset * 8)

    def close(self):
        if getattr(self, "ptr", None):
            lib().bcmgx_device_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_device(ctx, A, b_ptr, x_ptr, nrhs, params, prec=None, ld=None):
    """solve with right-hand sides and solution already in device memory
    (column-major, leading dimension ld); returns the SolverOut dict."""
    out = abi.SolverOut()
    _check(lib().bcmgx_solve(ctx.h, A.h, prec.h if prec is not None else None, C.byref(params.s), b_ptr, x_ptr, nrhs,
                             ld or A.n, MEM_DEVICE, C.byref(out)), ctx)
    res = {f: getattr(out, f) for f, _ in abi.SolverOut._fields_}
    if prec is not None:
        it = np.zeros(nrhs, np.int32)
        er = np.zeros(nrhs)
        rr = np.zeros(nrhs)
        _check(lib().bcmgx_solve_columns(prec.h, nrhs, it, er, rr), ctx)
        res["niter_cols"] = it.tolist()
        res["exit_res_cols"] = er.tolist()
        res["rel_res_cols"] = rr.tolist()
    return res


class CSR:
    """Device CSR matrix (CSRm::init + copyToDevice, CSR.h:297,347)."""

    def __init__(self, ctx, n, m, row, col, val):
        self.ctx = ctx
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        h = C.c_void_p()
        _check(lib().bcmgx_csr_create(ctx.h, n, m, len(col), _ptr(row), _ptr(col), _ptr(val), MEM_HOST, C.byref(h)), ctx)
        self.h = h
        self.n, self.m, self.nnz = n, m, len(col)

    @classmethod
    def _wrap(cls, ctx, h):
        o = cls.__new__(cls)
        o.ctx = ctx
        o.h = h
        n, m, z = C.c_int(), C.c_int(), C.c_int()
        lib().bcmgx_csr_info(h, C.byref(n), C.byref(m), C.byref(z))
        o.n, o.m, o.nnz = n.value, m.value, z.value
        return o

    @classmethod
    def from_device(cls, ctx, n, m, nnz, row_ptr, col_ptr, val_ptr):
        """A matrix copied from CSR arrays already in device memory (int32 row pointers
        and columns, float64 values; addresses as ints or c_void_p)."""
        h = C.c_void_p()
        vp = lambda x: x if isinstance(x, C.c_void_p) else C.c_void_p(x)
        _check(lib().bcmgx_csr_create(ctx.h, n, m, nnz, vp(row_ptr), vp(col_ptr), vp(val_ptr), MEM_DEVICE, C.byref(h)), ctx)
        return cls._wrap(ctx, h)

    @classmethod
    def from_scipy(cls, ctx, S):
        S = S.tocsr()
        S.sort_indices()
        return cls(ctx, S.shape[0], S.shape[1], S.indptr, S.indices, S.data)

    def close(self):
        if getattr(self, "h", None):
            lib().bcmgx_csr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update_values(self, val):
        """Overwrite the nnz values (same sparsity pattern)."""
        val = np.ascontiguousarray(val, np.float64)
        if val.shape != (self.nnz,):
            raise ValueError(f"expected {self.nnz} values, got shape {val.shape}")
        _check(lib().bcmgx_csr_update_values(self.h, _ptr(val), MEM_HOST), self.ctx)

    def analyse(self):
        """Build the row-tile index of the pattern now; returns the seconds it took."""
        t = C.c_double()
        _check(lib().bcmgx_csr_analyse(self.h, C.byref(t)), self.ctx)
        return t.value

    def to_host(self):
        row = np.zeros(self.n + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz, np.float64)
        _check(lib().bcmgx_csr_copy_to_host(self.h, row, col, val), self.ctx)
        return row, col, val

    def to_scipy(self):
        import scipy.sparse as sp

        row, col, val = self.to_host()
        return sp.csr_matrix((val, col, row), shape=(self.n, self.m))

    # kernel-level operations
    def spmv(self, x, y=None, alpha=1.0, beta=0.0, op=1, lanes=0):
        """_CSR_vector_mul_mini_warp<OP> (CSR.cu:452-492) on the device."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n) if y is None else np.ascontiguousarray(y, np.float64).copy()
        _check(lib().bcmgx_spmv(self.ctx.h, self.h, _ptr(x), _ptr(y), alpha, beta, op, lanes, MEM_HOST), self.ctx)
        return y

    def spmm(self, X, Y=None, alpha=1.0, beta=0.0, op=1, kchunk=0):
        """The op of spmv on the k columns of X (n x k) at once."""
        X = np.asarray(X, np.float64)
        k = X.shape[1]
        ld = max(self.n, self.m)  # one leading dimension for both blocks
        Xp = np.zeros((ld, k), order="F")
        Xp[: self.m] = X
        Yp = np.zeros((ld, k), order="F")
        if Y is not None:
            Yp[: self.n] = Y
        _check(lib().bcmgx_spmm(self.ctx.h, self.h, _ptr(Xp), _ptr(Yp), k, ld, alpha, beta, op, kchunk, MEM_HOST),
               self.ctx)
        return Yp[: self.n].copy()

    def spmm_device(self, x_ptr, y_ptr, k, ld, alpha=1.0, beta=0.0, op=1, kchunk=0):
        """spmm on column-major device blocks (leading dimension ld)."""
        _check(lib().bcmgx_spmm(self.ctx.h, self.h, x_ptr, y_ptr, k, ld, alpha, beta, op, kchunk, MEM_DEVICE),
               self.ctx)

    def time_spmv(self, precision=0, k=1, lanes=0, reps=20):
        """Average device milliseconds of one y = A x launch (driverSpMV)."""
        ms = C.c_double()
        _check(lib().bcmgx_time_spmv(self.ctx.h, self.h, precision, k, lanes, reps, C.byref(ms)), self.ctx)
        return ms.value

    def time_smoother(self, precision=0, k=1, reps=20):
        """Average device milliseconds of one l1-Jacobi sweep on k columns."""
        ms = C.c_double()
        _check(lib().bcmgx_time_smoother(self.ctx.h, self.h, precision, k, reps, C.byref(ms)), self.ctx)
        return ms.value

    def time_spgemm(self, B, reps=5):
        """(average device milliseconds of one product self * B, nonzeros of the product)."""
        ms, nnz = C.c_double(), C.c_int()
        _check(lib().bcmgx_time_spgemm(self.ctx.h, self.h, B.h, reps, C.byref(ms), C.byref(nnz)), self.ctx)
        return ms.value, nnz.value

    def spgemm(self, B):
        h = C.c_void_p()
        _check(lib().bcmgx_spgemm(self.ctx.h, self.h, B.h, C.byref(h)), self.ctx)
        return CSR._wrap(self.ctx, h)

    def transpose(self):
        h = C.c_void_p()
        _check(lib().bcmgx_transpose(self.ctx.h, self.h, C.byref(h)), self.ctx)
        return CSR._wrap(self.ctx, h)

    def matching(self, w):
        M = np.zeros(self.n, np.int32)
        _check(lib().bcmgx_matching(self.ctx.h, self.h, np.ascontiguousarray(w, np.float64), M), self.ctx)
        return M

    def sparse_outer_grad(self, lam, x):
        lam = np.asfortranarray(np.asarray(lam, np.float64).reshape(self.n, -1))
        x = np.asfortranarray(np.asarray(x, np.float64).reshape(self.n, -1))
        g = np.zeros(self.nnz)
        _check(lib().bcmgx_sparse_outer_grad(self.ctx.h, self.h, _ptr(lam), _ptr(x), lam.shape[1], _ptr(g), MEM_HOST), self.ctx)
        return g


def make_prolongator(ctx, M, w):
    """makeP_GPU (matchingPairAggregation.cu:109-170), canonical singleton order."""
    n = len(M)
    agg = np.zeros(n, np.int32)
    pv = np.zeros(n)
    nc = C.c_int()
    _check(lib().bcmgx_make_prolongator(ctx.h, n, np.ascontiguousarray(M, np.int32), np.ascontiguousarray(w, np.float64),
                                        agg, pv, C.byref(nc)), ctx)
    return agg, pv, nc.value


def _take_arrays(n, nnz, r, c, v):
    L = lib()
    row = np.ctypeslib.as_array(r, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(c, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(v, shape=(max(nnz, 1),))[:nnz].copy()
    for p in (r, c, v):
        L.bcmgx_free(C.cast(p, C.c_void_p))
    return row, col, val


def laplacian3d(points, nx, ny, nz, P=1, Q=1, R=1):
    """Host arrays of generateLocalLaplacian3D_{7p,27p} over a P x Q x R grid."""
    n, z = C.c_int(), C.c_int()
    r, c, v = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
    _check(lib().bcmgx_gen_laplacian3d(points, nx, ny, nz, P, Q, R, C.byref(n), C.byref(z), C.byref(r), C.byref(c), C.byref(v)))
    return (n.value,) + _take_arrays(n.value, z.value, r, c, v)


def diffusion7p(nx, ny, nz, kind, p1, p2, seed):
    """Variable-coefficient 7-point diffusion (kind 0: exp(p1*N(0,1)); kind 1: U[p1,p2))."""
    n, z = C.c_int(), C.c_int()
    r, c, v = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
    _check(lib().bcmgx_gen_diffusion7p(nx, ny, nz, kind, p1, p2, seed, C.byref(n), C.byref(z), C.byref(r), C.byref(c), C.byref(v)))
    return (n.value,) + _take_arrays(n.value, z.value, r, c, v)


def read_mtx(path):
    """readMTXDouble + check_and_fix_order (CSR.cu:1973-2152, :1895-1918)."""
    n, m, z = C.c_int(), C.c_int(), C.c_int()
    r, c, v = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
    _check(lib().bcmgx_read_mtx(os.fsencode(path), C.byref(n), C.byref(m), C.byref(z), C.byref(r), C.byref(c), C.byref(v)))
    row, col, val = _take_arrays(n.value, z.value, r, c, v)
    return n.value, m.value, row, col, val


# ---------------------------------------------------------- preconditioner ---
class Prec:
    """cgsprec: the preconditioner state built by prec_setup (prec_setup.h:23-55)."""

    def __init__(self, ctx, A, params):
        self.ctx, self.A, self.params = ctx, A, params
        h = C.c_void_p()
        _check(lib().bcmgx_prec_setup(ctx.h, A.h, C.byref(params.s), C.byref(h)), ctx)
        self.h = h

    def finalize(self):
        """prec_finalize (prec_finalize.h:29)."""
        if getattr(self, "h", None):
            lib().bcmgx_prec_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass

    def apply(self, r, u=None):
        """prec_apply (prec_apply.h:32): u <- B r, u also the initial guess."""
        r = np.asfortranarray(np.asarray(r, np.float64).reshape(self.A.n, -1))
        k = r.shape[1]
        u = np.zeros((self.A.n, k), order="F") if u is None else np.asfortranarray(np.asarray(u, np.float64).reshape(self.A.n, -1)).copy(order="F")
        _check(lib().bcmgx_prec_apply(self.h, _ptr(r), _ptr(u), k, self.A.n, MEM_HOST), self.ctx)
        return u[:, 0] if k == 1 else u

    def info(self):
        i = abi.PrecInfo()
        _check(lib().bcmgx_prec_get_info(self.h, C.byref(i)), self.ctx)
        return {f: getattr(i, f) for f, _ in abi.PrecInfo._fields_}

    def levels(self, h=0):
        """Levels of bootstrap hierarchy h (boot::H_array[h])."""
        if h == 0:
            return self.info()["num_levels"]
        n = C.c_int()
        _check(lib().bcmgx_prec_hrc_levels(self.h, h, C.byref(n)), self.ctx)
        return n.value

    def num_hrc(self):
        """(boot::n_hrc, boot::estimated_ratio)"""
        i = self.info()
        return i["num_hrc"], i["estimated_ratio"]

    def level_size(self, l, h=0):
        n, z = C.c_int(), C.c_int()
        _check(lib().bcmgx_prec_hrc_level_size(self.h, h, l, C.byref(n), C.byref(z)), self.ctx)
        return n.value, z.value

    def level_matrix(self, l, h=0):
        import scipy.sparse as sp

        n, z = self.level_size(l, h)
        row = np.zeros(n + 1, np.int32)
        col = np.zeros(z, np.int32)
        val = np.zeros(z)
        _check(lib().bcmgx_prec_hrc_level_matrix(self.h, h, l, row, col, val), self.ctx)
        return sp.csr_matrix((val, col, row), shape=(n, n))

    def level_prolongator(self, l, h=0):
        n, _ = self.level_size(l, h)
        agg = np.zeros(n, np.int32)
        pv = np.zeros(n)
        _check(lib().bcmgx_prec_hrc_level_prolongator(self.h, h, l, agg, pv), self.ctx)
        return agg, pv

    def afsai_factor(self):
        import scipy.sparse as sp

        n, z = C.c_int(), C.c_int()
        _check(lib().bcmgx_prec_afsai_factor(self.h, C.byref(n), C.byref(z), None, None, None), self.ctx)
        row = np.zeros(n.value + 1, np.int32)
        col = np.zeros(z.value, np.int32)
        val = np.zeros(z.value)
        _check(lib().bcmgx_prec_afsai_factor(self.h, C.byref(n), C.byref(z), _ptr(row), _ptr(col), _ptr(val)), self.ctx)
        return sp.csr_matrix((val, col, row), shape=(n.value, n.value))


def prec_setup(ctx, A, params):
    """prec_setup (prec_setup.h:55); the type comes from params.preconditioner."""
    return Prec(ctx, A, params)


def prec_apply(prec, r, u=None):
    return prec.apply(r, u)


def prec_finalize(prec):
    prec.finalize()


# ------------------------------------------------------------------- solve ---
def solve(ctx, A, rhs, x0, params, prec=None):
    """solve (solve.h:33): FCG, CGHS, CGS, CGS_CUBLAS or PIPELINED_CGS per ``params.solver``.

    ``rhs`` and ``x0`` are (n,) or (n, k) arrays; returns (x, SolverOut dict),
    the SolverOut with the reference's counting (niter = iterations + 1).
    For k > 1 the dict also has per-column lists.
    """
    b = np.asfortranarray(np.asarray(rhs, np.float64).reshape(A.n, -1))
    k = b.shape[1]
    x = np.zeros((A.n, k), order="F") if x0 is None else np.asfortranarray(np.asarray(x0, np.float64).reshape(A.n, -1)).copy(order="F")
    out = abi.SolverOut()
    _check(lib().bcmgx_solve(ctx.h, A.h, prec.h if prec is not None else None, C.byref(params.s), _ptr(b), _ptr(x), k, A.n,
                             MEM_HOST, C.byref(out)), ctx)
    res = {f: getattr(out, f) for f, _ in abi.SolverOut._fields_}
    if prec is not None:
        it = np.zeros(k, np.int32)
        er = np.zeros(k)
        rr = np.zeros(k)
        _check(lib().bcmgx_solve_columns(prec.h, k, it, er, rr), ctx)
        res["niter_cols"] = it.tolist()
        res["exit_res_cols"] = er.tolist()
        res["rel_res_cols"] = rr.tolist()
    return (x[:, 0] if k == 1 else x), res


def dump_info(path, params, prec, out):
    """Info file of SolverOut.cu:9-97."""
    so = abi.SolverOut()
    for f, _ in abi.SolverOut._fields_:
        if f in out:
            setattr(so, f, out[f])
    _check(lib().bcmgx_dump_info(os.fsencode(path), C.byref(params.s), prec.h if prec is not None else None, C.byref(so)))
