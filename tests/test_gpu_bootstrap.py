"""Bootstrap AMG (several hierarchies) on the HIP path vs the CPU oracle, through the C ABI.

bootstrap (bootstrap.cu:11-46) builds hierarchies from the slowest error of
the composition of those already built; bcmg_apply (bcmg.cu:51-90) composes
them.  The first hierarchy is the plain one and bit-exact; the later ones
start their matching from a smooth vector made by preconditioner
applications and dot products, whose sums the GPU groups differently, so
their aggregates must be equal and their values within 1e-10.  Compositions
are checked by prec_apply against the oracle (1e-10), FCG solves by
iteration count (+-1) and solution (1e-8)."""
import numpy as np
import pytest

import bootcmatchgx_amd as bx
import oracle_py as O
from bootcmatchgx_amd import abi

pytestmark = pytest.mark.gpu

COMPS = ["MULTIPLICATIVE", "SYMMETRIZED_MULTIPLICATIVE", "ADDITIVE"]


def dev(ctx, ocsr):
    n, m, _ = ocsr.dims()
    r, c, v = ocsr.arrays()
    return bx.CSR(ctx, n, m, r, c, v)


def params(comp, **kw):
    p = bx.Params(solver="FCG", preconditioner="BCMG", dispnorm=0, bootstrap_type=comp, max_hrc=3, conv_ratio=0.0,
                  rtol=1e-9)
    for k, v in kw.items():
        p.set(k, v)
    return p


def oparams(p):
    import ctypes as C

    q = abi.Params()
    C.pointer(q)[0] = p.s
    return q


def check_hierarchies(P, op, exact_first=True, tol=1e-10, ratio_tol=1e-8):
    """tol: later hierarchies' values and prolongators (relative to the largest);
    ratio_tol: the estimated convergence ratio (relative)."""
    nh, ratio = P.num_hrc()
    onh, oratio = op.num_hrc()
    assert nh == onh
    assert abs(ratio - oratio) <= ratio_tol * max(oratio, 1e-300)
    for h in range(nh):
        assert P.levels(h) == op.hrc_levels(h), h
        for l in range(P.levels(h)):
            g = P.level_matrix(l, h)
            o = op.hrc_level_A(h, l).to_scipy()
            assert g.shape == o.shape and np.array_equal(g.indptr, o.indptr) and np.array_equal(g.indices, o.indices)
            if h == 0 and exact_first:
                assert np.array_equal(g.data, o.data), (h, l)
            else:
                assert np.all(np.abs(g.data - o.data) <= tol * np.abs(o.data).max()), (h, l)
            if l < P.levels(h) - 1:
                ga, gv = P.level_prolongator(l, h)
                oa_, ov = op.hrc_level_P(h, l)
                assert np.array_equal(ga, oa_), (h, l)
                assert np.all(np.abs(gv - ov) <= tol * np.abs(ov).max()), (h, l)


@pytest.mark.parametrize("comp", COMPS)
@pytest.mark.parametrize("mat", ["diff", "lap27"])
def test_bootstrap_hierarchies_apply_and_solve(ctx, comp, mat):
    oa = O.diffusion7p(16, 15, 14, 1, 0.5, 2.0, 4) if mat == "diff" else O.laplacian3d(27, 14, 14, 14)
    print(f"bootstrap {comp} {mat}", flush=True)
    A = dev(ctx, oa)
    p = params(comp)
    P = bx.prec_setup(ctx, A, p)
    op = O.Prec(oa, oparams(p))
    assert op.num_hrc()[0] == 3
    check_hierarchies(P, op)
    n = A.n
    rng = np.random.default_rng(7)
    r, u = rng.normal(size=n), rng.normal(size=n)
    for guess in (None, u):
        got = P.apply(r, guess)
        ref = op.apply(r, guess)
        assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref)
    b = rng.normal(size=n)
    x, out = bx.solve(ctx, A, b, None, p, P)
    xo, oi = O.fcg(oa, op, oparams(p), b)
    assert out["retv"] == 0
    assert abs(out["niter"] - oi["niter"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    P.finalize()
    A.close()


@pytest.mark.parametrize("comp", ["SYMMETRIZED_MULTIPLICATIVE", "ADDITIVE"])
def test_bootstrap_batched_columns_equal_single_solves(ctx, comp):
    oa = O.diffusion7p(16, 15, 14, 1, 0.5, 2.0, 4)
    A = dev(ctx, oa)
    p = params(comp, max_hrc=2)
    P = bx.prec_setup(ctx, A, p)
    rng = np.random.default_rng(9)
    B = rng.normal(size=(A.n, 4))
    X, out = bx.solve(ctx, A, B, None, p, P)
    assert out["retv"] == 0
    for j in range(4):
        x1, o1 = bx.solve(ctx, A, B[:, j], None, p, P)
        assert abs(o1["niter"] - out["niter_cols"][j]) <= 1
        assert np.linalg.norm(X[:, j] - x1) <= 1e-8 * np.linalg.norm(x1)
    P.finalize()
    A.close()


@pytest.mark.parametrize("comp", COMPS)
def test_bootstrap_fp32_hierarchies(ctx, comp):
    """FP32 hierarchy values and cycles under the FP64 FCG: the oracle's FP32 cycle
    makes the same hierarchies; solves agree in iterations (+-1) and to 1e-7.  The
    smooth vectors of the later hierarchies come from FP32 cycles, whose sums the GPU
    groups differently (FP32 rounding 6e-8 per operation): the estimated ratio is
    checked to 1e-6 and the later hierarchies' values to 1e-5, their aggregates
    exactly; the first hierarchy is bit-exact."""
    oa = O.diffusion7p(16, 15, 14, 1, 0.5, 2.0, 4)
    A = dev(ctx, oa)
    p = params(comp, max_hrc=2, precision="FP32")
    P = bx.prec_setup(ctx, A, p)
    op = O.Prec(oa, oparams(p))
    nh, ratio = P.num_hrc()
    print(f"{comp}: estimated ratio {ratio:.10f} (oracle {op.num_hrc()[1]:.10f})")
    check_hierarchies(P, op, tol=1e-5, ratio_tol=1e-6)
    b = np.random.default_rng(11).normal(size=A.n)
    x, out = bx.solve(ctx, A, b, None, p, P)
    xo, oi = O.fcg(oa, op, oparams(p), b)
    assert out["retv"] == 0
    assert abs(out["niter"] - oi["niter"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)
    P.finalize()
    A.close()


def test_bootstrap_stop_rule_and_info(ctx, tmp_path):
    oa = O.laplacian3d(7, 12, 12, 12)
    A = dev(ctx, oa)
    p = params("SYMMETRIZED_MULTIPLICATIVE", max_hrc=2, conv_ratio=0.999)
    P = bx.prec_setup(ctx, A, p)
    op = O.Prec(oa, oparams(p))
    nh, ratio = P.num_hrc()
    assert nh == 1 and (nh, ) == (op.num_hrc()[0], )
    assert abs(ratio - op.num_hrc()[1]) <= 1e-8 * ratio and 0 < ratio <= 0.999
    x, out = bx.solve(ctx, A, np.ones(A.n), None, p, P)
    f = tmp_path / "info.txt"
    bx.dump_info(str(f), p, P, out)
    txt = f.read_text()
    assert "Number of hierarchies                        : 1" in txt
    assert f"Estimated ratio                              : {ratio:f}" in txt
    P.finalize()
    A.close()


def test_bootstrap_zero_weight_singletons(ctx):
    """Rows holding only a unit diagonal are smoothed exactly (one l1-Jacobi sweep
    leaves their error at 0), so the second hierarchy's smooth vector is 0 there:
    those vertices are singletons of zero weight, which join the aggregate before
    them (makeP), giving restriction rows of more than two entries.  The hierarchies
    still equal the oracle's (first bit-exact, second to 1e-10)."""
    import scipy.sparse as sp

    lap = O.laplacian3d(27, 10, 10, 10).to_scipy()
    S = sp.block_diag([lap, sp.identity(40)]).tocsr()
    S.sort_indices()
    oa = O.Csr.from_scipy(S)
    A = dev(ctx, oa)
    p = params("MULTIPLICATIVE", max_hrc=2)
    P = bx.prec_setup(ctx, A, p)
    op = O.Prec(oa, oparams(p))
    assert op.num_hrc()[0] == 2
    _, pv = op.hrc_level_P(1, 0)
    assert np.count_nonzero(pv[-40:] == 0.0) == 40  # the zero-weight singletons are there
    check_hierarchies(P, op)
    b = np.random.default_rng(13).normal(size=A.n)
    x, out = bx.solve(ctx, A, b, None, p, P)
    xo, oi = O.fcg(oa, op, oparams(p), b)
    assert out["retv"] == 0
    assert abs(out["niter"] - oi["niter"]) <= 1
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    P.finalize()
    A.close()
