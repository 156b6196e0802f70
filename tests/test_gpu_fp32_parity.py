"""The measured configuration's numerics against the oracle (needs an MI355X):
the FP32 hierarchy's V-cycle applied to batched columns, and the batched FCG
whose preconditioned residual stays in FP32 (the headline solve), column by
column against the oracle's FP32 cycle and FCG.

Tolerances: the FP32 cycle performs the oracle's FP32 operations in the same
order (the coarsest level in FP64 on both sides), so its columns are checked
to 1e-6 relative (FP32 rounding is 6e-8; the measured difference is printed);
the FCG solutions to 1e-9 relative with iteration counts within 1 (the FP64
dot products are summed in other fixed groups than the oracle's, which moves
the iterates in the last bits)."""
import numpy as np
import pytest

import bootcmatchgx_amd as bx
import oracle_py as O
from bootcmatchgx_amd import abi

pytestmark = pytest.mark.gpu


def dev(ctx, ocsr):
    n, m, _ = ocsr.dims()
    r, c, v = ocsr.arrays()
    return bx.CSR(ctx, n, m, r, c, v)


def params(prec, **kw):
    p = bx.Params(solver="FCG", preconditioner=prec, dispnorm=0)
    for k, v in kw.items():
        p.set(k, v)
    return p


def oparams(p):
    q = abi.Params()
    __import__("ctypes").pointer(q)[0] = p.s
    return q


CYCLES = [
    (1, dict(post_galerkin=1)),
    (8, dict(post_galerkin=1)),
    (32, dict(post_galerkin=1)),
    (32, dict(post_galerkin=0)),
    (1, dict(post_galerkin=0)),
    (8, dict(post_galerkin=1, cycle_type="W_CYCLE", postrelax_sweeps=2)),
    (1, dict(post_galerkin=1, cycle_type="VARIABLE_V_CYCLE", postrelax_sweeps=2)),
]


@pytest.mark.parametrize("k,kw", CYCLES)
def test_fp32_cycle_matches_oracle_fp32_cycle(ctx, k, kw):
    """27-point 42^3 (74,088 rows: level 0 takes the autotuned kernels).  With
    post_galerkin the first post-smoothing sweep of each level runs on A P (k >= 8:
    the pipelined SpMM's fused epilogue; k = 1: the product, then the epilogue)."""
    oa = O.laplacian3d(27, 42, 42, 42)
    A = dev(ctx, oa)
    kw = dict(dict(prerelax_sweeps=1, postrelax_sweeps=1), **kw)
    p = params("BCMG", precision="FP32", **kw)
    P = bx.prec_setup(ctx, A, p)
    R = np.random.default_rng(31).normal(size=(A.n, k))
    got = P.apply(R).reshape(A.n, k)
    op = O.Prec(oa, oparams(p))
    p64 = params("BCMG", **kw)
    op64 = O.Prec(oa, oparams(p64))
    worst, vs64 = 0.0, 0.0
    for j in range(k):
        ref = op.apply(R[:, j].copy())
        worst = max(worst, np.linalg.norm(got[:, j] - ref) / np.linalg.norm(ref))
        r64 = op64.apply(R[:, j].copy())
        vs64 = max(vs64, np.linalg.norm(got[:, j] - r64) / np.linalg.norm(r64))
    print(f"k={k} {kw}: FP32 cycle vs oracle FP32 cycle {worst:.2e}, vs oracle FP64 cycle {vs64:.2e}")
    assert worst <= 1e-6
    assert 1e-9 < vs64 <= 1e-4  # it ran in FP32, and stays the FP64 operator to FP32 accuracy
    P.finalize()


@pytest.mark.parametrize("k", [1, 32, 64])
def test_fp32_fcg_matches_oracle(ctx, k):
    """27-point 24^3, rtol 1e-8, 1/1 sweeps, FP32 hierarchy, k columns N(0,1)."""
    oa = O.laplacian3d(27, 24, 24, 24)
    A = dev(ctx, oa)
    p = params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, precision="FP32", fcg_z32=1)
    P = bx.prec_setup(ctx, A, p)
    B = np.random.default_rng(32).normal(size=(A.n, k))
    X, out = bx.solve(ctx, A, B, None, p, P)
    X = X.reshape(A.n, k)
    op = O.Prec(oa, oparams(p))
    worst, dit = 0.0, 0
    its = out["niter_cols"] if k > 1 else [out["niter"]]
    for j in range(k):
        xo, oi = O.fcg(oa, op, oparams(p), B[:, j].copy())
        dit = max(dit, abs(its[j] - oi["niter"]))
        worst = max(worst, np.linalg.norm(X[:, j] - xo) / np.linalg.norm(xo))
        r = B[:, j] - O.spmv(oa, X[:, j])
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(B[:, j]) * 1.0000001
    print(f"k={k}: FCG with FP32 cycle vs oracle: iterations differ by <= {dit}, solutions {worst:.2e}")
    assert out["retv"] == 0 and dit <= 1
    assert worst <= 1e-9
    P.finalize()
