"""ctypes access to the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, the smoke check and the CPU-baseline leg of
bench.py use it, always as the checker.
"""
import ctypes as C
import os

import numpy as np

from bootcmatchgx_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_lib = None

I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.or_csr_new.restype = vp
        L.or_csr_new.argtypes = [C.c_int, C.c_int, C.c_int, I32P, I32P, F64P]
        L.or_csr_free.argtypes = [vp]
        L.or_csr_dims.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_csr_copy.argtypes = [vp, I32P, I32P, F64P]
        L.or_gen_laplacian3d.restype = vp
        L.or_gen_laplacian3d.argtypes = [C.c_int] * 7
        L.or_gen_diffusion7p.restype = vp
        L.or_gen_diffusion7p.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_ulonglong]
        L.or_read_mtx.restype = vp
        L.or_read_mtx.argtypes = [C.c_char_p]
        L.or_spmv.argtypes = [vp, F64P, F64P, C.c_double, C.c_double, C.c_int]
        L.or_spgemm.restype = vp
        L.or_spgemm.argtypes = [vp, vp]
        L.or_transpose.restype = vp
        L.or_transpose.argtypes = [vp]
        L.or_matching_weights.restype = C.c_int
        L.or_matching_weights.argtypes = [vp, F64P, I32P, I32P, F64P]
        L.or_suitor.argtypes = [vp, F64P, I32P]
        L.or_greedy_matching.argtypes = [vp, F64P, I32P]
        L.or_make_p.restype = C.c_int
        L.or_make_p.argtypes = [C.c_int, I32P, F64P, I32P, F64P]
        L.or_prec_setup.restype = vp
        L.or_prec_setup.argtypes = [vp, C.POINTER(abi.Params)]
        L.or_prec_free.argtypes = [vp]
        L.or_prec_levels.restype = C.c_int
        L.or_prec_levels.argtypes = [vp]
        L.or_prec_level_A.restype = vp
        L.or_prec_level_A.argtypes = [vp, C.c_int]
        L.or_prec_level_P.restype = C.c_int
        L.or_prec_level_P.argtypes = [vp, C.c_int, I32P, F64P]
        L.or_prec_hierarchy_stats.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.or_prec_afsai_G.restype = vp
        L.or_prec_afsai_G.argtypes = [vp]
        L.or_prec_apply.argtypes = [vp, F64P, F64P]
        L.or_prec_num_hrc.restype = C.c_int
        L.or_prec_num_hrc.argtypes = [vp, C.POINTER(C.c_double)]
        L.or_prec_hrc_levels.restype = C.c_int
        L.or_prec_hrc_levels.argtypes = [vp, C.c_int]
        L.or_prec_hrc_level_A.restype = vp
        L.or_prec_hrc_level_A.argtypes = [vp, C.c_int, C.c_int]
        L.or_prec_hrc_level_P.restype = C.c_int
        L.or_prec_hrc_level_P.argtypes = [vp, C.c_int, C.c_int, I32P, F64P]
        L.or_prec_apply_hrc.argtypes = [vp, C.c_int, F64P, F64P]
        L.or_prec_setup_w.restype = vp
        L.or_prec_setup_w.argtypes = [vp, C.POINTER(abi.Params), F64P]
        L.or_boot_random.argtypes = [C.c_int, C.c_ulonglong, F64P]
        L.or_fcg.restype = C.c_int
        L.or_fcg.argtypes = [vp, vp, C.POINTER(abi.Params), F64P, F64P, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_cghs.restype = C.c_int
        L.or_cghs.argtypes = L.or_fcg.argtypes
        L.or_cgs.restype = C.c_int
        L.or_cgs.argtypes = L.or_fcg.argtypes
        L.or_pipecgs.restype = C.c_int
        L.or_pipecgs.argtypes = L.or_fcg.argtypes
        L.or_num_threads.restype = C.c_int
        _lib = L
    return _lib


class Csr:
    """Host CSR owned by the oracle."""

    def __init__(self, handle, owned=True):
        if not handle:
            raise RuntimeError("oracle returned NULL")
        self.h = handle
        self.owned = owned

    @classmethod
    def from_arrays(cls, n, m, row, col, val):
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        return cls(lib().or_csr_new(n, m, len(col), row, col, val))

    @classmethod
    def from_scipy(cls, S):
        S = S.tocsr()
        S.sort_indices()
        return cls.from_arrays(S.shape[0], S.shape[1], S.indptr, S.indices, S.data)

    def __del__(self):
        if getattr(self, "owned", False) and self.h and _lib is not None:
            _lib.or_csr_free(self.h)
            self.h = None

    def dims(self):
        n, m, z = C.c_int(), C.c_int(), C.c_int()
        lib().or_csr_dims(self.h, C.byref(n), C.byref(m), C.byref(z))
        return n.value, m.value, z.value

    def arrays(self):
        n, m, z = self.dims()
        row = np.zeros(n + 1, np.int32)
        col = np.zeros(z, np.int32)
        val = np.zeros(z, np.float64)
        lib().or_csr_copy(self.h, row, col, val)
        return row, col, val

    def to_scipy(self):
        import scipy.sparse as sp

        n, m, _ = self.dims()
        row, col, val = self.arrays()
        return sp.csr_matrix((val, col, row), shape=(n, m))


def laplacian3d(points, nx, ny, nz, P=1, Q=1, R=1):
    return Csr(lib().or_gen_laplacian3d(points, nx, ny, nz, P, Q, R))


def diffusion7p(nx, ny, nz, kind, p1, p2, seed):
    return Csr(lib().or_gen_diffusion7p(nx, ny, nz, kind, p1, p2, seed))


def read_mtx(path):
    return Csr(lib().or_read_mtx(path.encode()))


def spmv(A, x, y=None, alpha=1.0, beta=0.0, op=1):
    n, _, _ = A.dims()
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(n) if y is None else np.ascontiguousarray(y, np.float64).copy()
    lib().or_spmv(A.h, x, y, alpha, beta, op)
    return y


def spgemm(A, B):
    return Csr(lib().or_spgemm(A.h, B.h))


def transpose(A):
    return Csr(lib().or_transpose(A.h))


def matching_weights(A, w):
    n, _, z = A.dims()
    r = np.zeros(n + 1, np.int32)
    c = np.zeros(z, np.int32)
    v = np.zeros(z, np.float64)
    k = lib().or_matching_weights(A.h, np.ascontiguousarray(w, np.float64), r, c, v)
    return r, c[:k], v[:k]


def suitor(A, w):
    n, _, _ = A.dims()
    M = np.zeros(n, np.int32)
    lib().or_suitor(A.h, np.ascontiguousarray(w, np.float64), M)
    return M


def greedy_matching(A, w):
    n, _, _ = A.dims()
    M = np.zeros(n, np.int32)
    lib().or_greedy_matching(A.h, np.ascontiguousarray(w, np.float64), M)
    return M


def make_p(M, w):
    n = len(M)
    agg = np.zeros(n, np.int32)
    pv = np.zeros(n, np.float64)
    nc = lib().or_make_p(n, np.ascontiguousarray(M, np.int32), np.ascontiguousarray(w, np.float64), agg, pv)
    return agg, pv, nc


def boot_random(n, seed):
    x = np.zeros(n)
    lib().or_boot_random(n, seed, x)
    return x


class Prec:
    def __init__(self, A, params, w=None):
        """w: the smooth vector the matching starts from (one hierarchy; default 1)."""
        self.A = A
        self.params = params
        if w is None:
            self.h = lib().or_prec_setup(A.h, C.byref(params))
        else:
            self.h = lib().or_prec_setup_w(A.h, C.byref(params), np.ascontiguousarray(w, np.float64))
        if not self.h:
            raise RuntimeError("oracle prec setup failed")

    def num_hrc(self):
        """(number of hierarchies, last estimated convergence ratio)"""
        r = C.c_double()
        n = lib().or_prec_num_hrc(self.h, C.byref(r))
        return n, r.value

    def hrc_levels(self, h):
        return lib().or_prec_hrc_levels(self.h, h)

    def hrc_level_A(self, h, l):
        c = Csr(lib().or_prec_hrc_level_A(self.h, h, l), owned=False)
        c.parent = self
        return c

    def hrc_level_P(self, h, l):
        n, _, _ = self.hrc_level_A(h, l).dims()
        agg = np.zeros(n, np.int32)
        pv = np.zeros(n, np.float64)
        lib().or_prec_hrc_level_P(self.h, h, l, agg, pv)
        return agg, pv

    def apply_hrc(self, h, r, u=None):
        n, _, _ = self.A.dims()
        u = np.zeros(n) if u is None else np.ascontiguousarray(u, np.float64).copy()
        lib().or_prec_apply_hrc(self.h, h, np.ascontiguousarray(r, np.float64), u)
        return u

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_prec_free(self.h)
            self.h = None

    def levels(self):
        return lib().or_prec_levels(self.h)

    def level_A(self, l):
        c = Csr(lib().or_prec_level_A(self.h, l), owned=False)
        c.parent = self  # the matrix lives inside this preconditioner
        return c

    def level_P(self, l):
        n, _, _ = self.level_A(l).dims()
        agg = np.zeros(n, np.int32)
        pv = np.zeros(n, np.float64)
        lib().or_prec_level_P(self.h, l, agg, pv)
        return agg, pv

    def stats(self):
        a, b, c, d = C.c_double(), C.c_double(), C.c_double(), C.c_int()
        lib().or_prec_hierarchy_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return dict(op_cmplx=a.value, op_wcmplx=b.value, avg_cratio=c.value, max_coarse_size=d.value)

    def afsai_G(self):
        c = Csr(lib().or_prec_afsai_G(self.h), owned=False)
        c.parent = self
        return c

    def apply(self, r, u=None):
        n, _, _ = self.A.dims()
        u = np.zeros(n) if u is None else np.ascontiguousarray(u, np.float64).copy()
        lib().or_prec_apply(self.h, np.ascontiguousarray(r, np.float64), u)
        return u


def pipecgs(A, prec, params, b, x0=None):
    n, _, _ = A.dims()
    x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    it, er, d0 = C.c_int(), C.c_double(), C.c_double()
    retv = lib().or_pipecgs(A.h, prec.h, C.byref(params), np.ascontiguousarray(b, np.float64), x, C.byref(it), C.byref(er),
                            C.byref(d0))
    return x, dict(retv=retv, niter=it.value, exit_res=er.value, del0=d0.value)


def cgs(A, prec, params, b, x0=None):
    n, _, _ = A.dims()
    x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    it, er, d0 = C.c_int(), C.c_double(), C.c_double()
    retv = lib().or_cgs(A.h, prec.h, C.byref(params), np.ascontiguousarray(b, np.float64), x, C.byref(it), C.byref(er), C.byref(d0))
    return x, dict(retv=retv, niter=it.value, exit_res=er.value, del0=d0.value)


def cghs(A, prec, params, b, x0=None):
    n, _, _ = A.dims()
    x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    it, er, d0 = C.c_int(), C.c_double(), C.c_double()
    retv = lib().or_cghs(A.h, prec.h, C.byref(params), np.ascontiguousarray(b, np.float64), x, C.byref(it), C.byref(er), C.byref(d0))
    return x, dict(retv=retv, niter=it.value, exit_res=er.value, del0=d0.value)


def fcg(A, prec, params, b, x0=None):
    n, _, _ = A.dims()
    x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    it, er, d0 = C.c_int(), C.c_double(), C.c_double()
    retv = lib().or_fcg(A.h, prec.h, C.byref(params), np.ascontiguousarray(b, np.float64), x, C.byref(it), C.byref(er), C.byref(d0))
    return x, dict(retv=retv, niter=it.value, exit_res=er.value, del0=d0.value)
