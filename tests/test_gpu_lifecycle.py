"""Handle lifetimes, reuse of one preconditioner by solves with other stop
rules, and in-place value updates, through the C ABI (needs an MI355X).

* A preconditioner serves any number of solves (the reference's cgsprec is
  set up once in driverSolve.cu and passed to solve()); a later solve with a
  tighter rtol or another itnlim follows its own stop rule (FCG.cu:146-162,
  CG_HS.cu:134-171) even when the iterations replay a captured graph.
* Destroying a context, a matrix or a preconditioner in any order is safe;
  a preconditioner whose matrix (its level 0) or context is gone answers
  BCMGX_ERR_STATE and is still freed by prec_finalize.
* After bcmgx_csr_update_values the level-0 divisors follow the new values:
  the cycle equals the dense restatement with level 0 replaced and the coarse
  levels kept.
"""
import numpy as np
import pytest

import bootcmatchgx_amd as bx
import oracle_py as O
from bootcmatchgx_amd import abi
from dense_cycle import DenseHierarchy, np_cycle

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


def true_rel(S, B, X):
    B = B.reshape(S.shape[0], -1)
    X = X.reshape(S.shape[0], -1)
    return np.linalg.norm(B - S @ X, axis=0) / np.linalg.norm(B, axis=0)


@pytest.mark.parametrize("solver", ["FCG", "CGHS"])
@pytest.mark.parametrize("k", [1, 8])
@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_prec_reused_with_other_stop_rules(ctx, solver, k, precision):
    oa = O.laplacian3d(27, 20, 20, 20)
    S = oa.to_scipy()
    A = dev(ctx, oa)
    B = np.random.default_rng(21).normal(size=(A.n, k))
    p = params("BCMG", solver=solver, rtol=1e-4, prerelax_sweeps=1, postrelax_sweeps=1, precision=precision)
    P = bx.prec_setup(ctx, A, p)
    _, o1 = bx.solve(ctx, A, B, None, p, P)
    assert o1["retv"] == 0 and np.all(true_rel(S, B, _) <= 1e-4 * 1.0001)
    p.set("rtol", 1e-10)
    X2, o2 = bx.solve(ctx, A, B, None, p, P)
    print(f"{solver} k={k} {precision}: rtol 1e-4 -> {o1['niter']} iterations, 1e-10 -> {o2['niter']}")
    assert o2["retv"] == 0 and o2["niter"] > o1["niter"]
    assert np.all(true_rel(S, B, X2) <= 1e-10 * 1.001)
    # a lower iteration limit stops the same solve there (FCG counts iter + 1, CG itnlim)
    p.set("itnlim", 3)
    _, o3 = bx.solve(ctx, A, B, None, p, P)
    assert o3["retv"] == -1 and o3["niter"] == (4 if solver == "FCG" else 3)
    p.set("itnlim", 2000)
    X4, o4 = bx.solve(ctx, A, B, None, p, P)
    assert o4["retv"] == 0 and o4["niter"] == o2["niter"]
    assert np.array_equal(X4, X2)  # the same iterates, bit for bit
    P.finalize()
    A.close()


def _raises_state(fn):
    with pytest.raises(bx.BcmgxError) as e:
        fn()
    assert e.value.status == abi.ERR_STATE, e.value


def test_destroy_matrix_before_its_preconditioners():
    c = bx.Context(0)
    oa = O.laplacian3d(7, 12, 12, 12)
    n = oa.dims()[0]
    r = np.random.default_rng(22).normal(size=n)
    A, other = dev(c, oa), dev(c, oa)
    p = params("BCMG")
    P = bx.prec_setup(c, A, p)
    P1 = bx.prec_setup(c, A, params("L1_JACOBI"))
    Q = bx.prec_setup(c, other, p)
    A.close()  # P and P1 lose their level 0
    _raises_state(lambda: P.apply(r))
    _raises_state(lambda: P.info())
    _raises_state(lambda: P1.apply(r))
    _raises_state(lambda: bx.solve(c, other, r, None, p, P))
    P.finalize()
    P1.finalize()
    # the context and its other handles are untouched
    x, out = bx.solve(c, other, r, None, p, Q)
    assert out["retv"] == 0
    assert np.all(np.isfinite(Q.apply(r)))
    c.close()
    # handles of a destroyed context answer BCMGX_ERR_STATE and free cleanly
    _raises_state(lambda: Q.apply(r))
    _raises_state(lambda: other.to_host())
    Q.finalize()
    other.close()


@pytest.mark.parametrize("order", ["context_first", "preconditioner_first", "matrix_first"])
def test_destruction_orders(order):
    c = bx.Context(0)
    oa = O.laplacian3d(27, 10, 10, 10)
    A = dev(c, oa)
    p = params("BCMG", precision="FP32")
    P = bx.prec_setup(c, A, p)
    x, out = bx.solve(c, A, np.ones((A.n, 8)), None, p, P)  # builds graphs and batched buffers
    assert out["retv"] == 0
    if order == "context_first":
        c.close()
        P.finalize()
        A.close()
    elif order == "preconditioner_first":
        P.finalize()
        A.close()
        c.close()
    else:
        A.close()
        P.finalize()
        c.close()
    # a fresh context on the same device works afterwards
    c2 = bx.Context(0)
    A2 = dev(c2, oa)
    x2, o2 = bx.solve(c2, A2, np.ones((A2.n, 8)), None, p, bx.prec_setup(c2, A2, p))
    assert o2["retv"] == 0 and np.array_equal(x2, x)
    A2.close()
    c2.close()


@pytest.mark.parametrize("precision,tol", [("FP64", 1e-12), ("FP32", 2e-5)])
def test_update_values_refreshes_level0_divisors(ctx, precision, tol):
    oa = O.diffusion7p(12, 11, 10, 1, 0.5, 2.0, 4)
    S = oa.to_scipy()
    A = dev(ctx, oa)
    p = params("BCMG", prerelax_sweeps=1, postrelax_sweeps=2, precision=precision)
    P = bx.prec_setup(ctx, A, p)
    pj = params("L1_JACOBI")
    J = bx.prec_setup(ctx, A, pj)
    r = np.random.default_rng(23).normal(size=A.n)
    new = 3.0 * S.data
    A.update_values(new)
    # level 0 is the new matrix, its divisor the new l1 row sums; the coarse
    # operators and prolongators stay those of the setup
    pro, coarse = [], []
    for l in range(P.levels() - 1):
        agg, pv = P.level_prolongator(l)
        nc = P.level_size(l + 1)[0]
        pro.append((agg, pv, nc))
        coarse.append(P.level_matrix(l + 1).toarray())
    H = DenseHierarchy(3.0 * S.toarray(), pro, coarse)
    ref = np_cycle(H, p.s, 0, r)
    got = P.apply(r)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"{precision}: cycle after the value update vs dense restatement {err:.2e}")
    assert err <= tol
    # one-level l1-Jacobi (set up on the old values): the same as a setup on the new values
    B2 = dev(ctx, O.Csr.from_arrays(A.n, A.n, S.indptr, S.indices, new))
    assert np.array_equal(J.apply(r), bx.prec_setup(ctx, B2, pj).apply(r))
    P.finalize()
    J.finalize()
    A.close()
    B2.close()


@pytest.mark.parametrize("prec", ["NONE", "L1_JACOBI", "BCMG", "AFSAI"])
def test_empty_or_rectangular_system_is_refused(prec):
    """A matrix with no rows, or a rectangular one, gets BCMGX_ERR_ARG from setup and
    from solve before any device work; the context stays usable."""
    c = bx.Context(0)
    empty = bx.CSR(c, 0, 0, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    rect = bx.CSR(c, 3, 4, np.arange(4, dtype=np.int32), np.array([0, 1, 2], np.int32), np.ones(3))
    for M in (empty, rect):
        with pytest.raises(bx.BcmgxError) as e:
            bx.prec_setup(c, M, params(prec))
        assert e.value.status == abi.ERR_ARG, e.value
        B, X = bx.DeviceArray(c, (4, 1)), bx.DeviceArray(c, (4, 1))
        with pytest.raises(bx.BcmgxError) as e:
            bx.solve_device(c, M, B.ptr, X.ptr, 1, params(prec), ld=4)
        assert e.value.status == abi.ERR_ARG, e.value
        B.close()
        X.close()
    oa = O.laplacian3d(7, 6, 6, 6)
    A = dev(c, oa)
    x, out = bx.solve(c, A, np.ones(A.n), None, params(prec), bx.prec_setup(c, A, params(prec)))
    assert out["retv"] == 0
    c.close()
