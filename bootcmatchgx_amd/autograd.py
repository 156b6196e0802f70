"""torch.autograd shim over the C ABI: x = A^{-1} b, differentiable in b and
in the stored nonzero values of A.

Forward runs the library's FCG (solve.h:33 -> FCG.cu:27-175) with the
preconditioner of an :class:`SPDSystem`; backward performs the adjoint solve
A^T lambda = dL/dx with the *same* hierarchy (A is symmetric positive
definite, so A^T = A) and returns

    dL/db       = lambda
    dL/dA_ij    = -lambda_i x_j     for every stored (i, j)   (pattern-restricted)

the latter from the library's sparse outer-product kernel (summed over
right-hand-side columns when b has k columns).

Tensors are CPU tensors: the shim moves them through host buffers of the C
ABI.  The library is loaded before torch so that it binds the system HIP
runtime; torch itself is not asked to touch the GPU.
"""
import numpy as np

from . import CSR, Context, Params, lib, prec_setup, solve

lib()  # bind libbcmgx (and its HIP runtime) before torch loads its own copy
import torch  # noqa: E402


class SPDSystem:
    """One sparsity pattern on one device with a preconditioner.

    ``row``/``col`` are the CSR pattern (sorted columns, diagonal present),
    ``values`` the initial nonzero values the hierarchy is built on.  When a
    forward call brings other values, the device matrix is updated in place;
    with ``rebuild=True`` the preconditioner is set up again on them,
    otherwise the existing hierarchy keeps serving (FCG is flexible, so a
    preconditioner of a nearby matrix still converges to the stated
    tolerance).
    """

    def __init__(self, n, row, col, values, params=None, device=0, ctx=None, rebuild=False):
        self.ctx = ctx if ctx is not None else Context(device)
        self.n = int(n)
        self.row = np.ascontiguousarray(row, np.int32)
        self.col = np.ascontiguousarray(col, np.int32)
        vals = np.ascontiguousarray(values, np.float64)
        self.A = CSR(self.ctx, self.n, self.n, self.row, self.col, vals)
        self.params = params if params is not None else Params(solver="FCG", preconditioner="BCMG", rtol=1e-10,
                                                                dispnorm=0)
        self.rebuild = rebuild
        self._vals = vals.copy()
        self.prec = prec_setup(self.ctx, self.A, self.params)
        self.last_info = None

    @property
    def nnz(self):
        return self.A.nnz

    def set_values(self, values):
        v = np.ascontiguousarray(values, np.float64)
        if v.shape != self._vals.shape:
            raise ValueError(f"expected {self._vals.shape[0]} values, got {v.shape}")
        if np.array_equal(v, self._vals):
            return
        self.A.update_values(v)
        self._vals = v.copy()
        if self.rebuild:
            self.prec.finalize()
            self.prec = prec_setup(self.ctx, self.A, self.params)

    def solve(self, b):
        """x = A^{-1} b for b of shape (n,) or (n, k) (numpy, FP64)."""
        x, info = solve(self.ctx, self.A, b, None, self.params, self.prec)
        self.last_info = info
        if info["retv"] != 0:
            raise RuntimeError(f"FCG did not reach rtol {self.params.rtol} within itnlim: {info}")
        return x

    def outer_grad(self, lam, x):
        """g_(ij) = -sum_c lambda[i, c] x[j, c] over the stored pattern."""
        return self.A.sparse_outer_grad(lam, x)

    def close(self):
        if getattr(self, "prec", None) is not None:
            self.prec.finalize()
            self.prec = None
        if getattr(self, "A", None) is not None:
            self.A.close()
            self.A = None


def _np(t):
    return t.detach().to(device="cpu", dtype=torch.float64).contiguous().numpy()


class SparseSolve(torch.autograd.Function):
    """x = A(values)^{-1} b on the pattern of ``system``."""

    @staticmethod
    def forward(fctx, values, b, system):
        if values.device.type != "cpu" or b.device.type != "cpu":
            raise ValueError("SparseSolve takes CPU tensors")
        vals = _np(values)
        system.set_values(vals)
        x = system.solve(_np(b))
        fctx.system = system
        # the adjoint needs this forward's matrix: another forward on the same
        # system may change the device values before backward runs
        fctx.values = vals.copy()
        out = torch.from_numpy(np.ascontiguousarray(x)).to(b.dtype)
        fctx.save_for_backward(out)
        return out

    @staticmethod
    def backward(fctx, gx):
        (x,) = fctx.saved_tensors
        system = fctx.system
        system.set_values(fctx.values)
        # adjoint solve with the forward hierarchy: A^T = A
        lam = system.solve(_np(gx))
        g_b = torch.from_numpy(np.ascontiguousarray(lam)).to(gx.dtype) if fctx.needs_input_grad[1] else None
        g_v = None
        if fctx.needs_input_grad[0]:
            g_v = torch.from_numpy(system.outer_grad(lam, _np(x))).to(gx.dtype)
        return g_v, g_b, None


def sparse_solve(values, b, system):
    """Differentiable x = A^{-1} b (see :class:`SparseSolve`)."""
    return SparseSolve.apply(values, b, system)
