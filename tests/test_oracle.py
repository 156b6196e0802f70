"""The CPU oracle pinned against the reference's own fixtures, closed forms,
brute force on tiny inputs and independent re-computations (no GPU)."""
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import oracle_py as O
from bootcmatchgx_amd import abi


def fcg_params(prec, **kw):
    p = abi.reference_defaults()
    p.solver = abi.SOLVER["FCG"]
    p.preconditioner = abi.PREC[prec]
    p.dispnorm = 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def mtx(path):
    M = scipy.io.mmread(path).tocsr()
    M.sort_indices()
    return M


# ---------------------------------------------------------------- generators
@pytest.mark.parametrize("fname,args", [
    ("lap3d_7p_2x2x5_3x4x2.mtx", (7, 2, 2, 5, 3, 4, 2)),
    ("lap3d_27p_5x5x5_1x2x2.mtx", (27, 5, 5, 5, 1, 2, 2)),
    ("lap3d_27p_5x5x5_4x2x1.mtx", (27, 5, 5, 5, 4, 2, 1)),
    ("lap3d_7p_4x14x4_2x2x1.mtx", (7, 4, 14, 4, 2, 2, 1)),
])
def test_generator_matches_matlab_fixture(golden, fname, args):
    """laplacian.cu block generators vs the reference's MATLAB-made matrices."""
    ref = mtx(os.path.join(golden, fname))
    A = O.laplacian3d(*args).to_scipy()
    assert A.shape == ref.shape
    assert (A != ref).nnz == 0


@pytest.mark.parametrize("m", [2, 3, 7, 10])
def test_generator_closed_forms(m):
    A7 = O.laplacian3d(7, m, m, m).to_scipy()
    A27 = O.laplacian3d(27, m, m, m).to_scipy()
    assert A7.nnz == m ** 3 + 6 * (m - 1) * m * m
    assert A27.nnz == (3 * m - 2) ** 3
    assert np.all(A7.diagonal() == 6) and np.all(A27.diagonal() == 26)
    for A in (A7, A27):
        assert (A - A.T).nnz == 0
        rs = np.asarray(A.sum(1)).ravel()
        assert rs.min() >= 0  # interior rows sum to 0, boundary rows > 0
    # interior row sums are exactly zero
    idx = (m // 2) * m * m + (m // 2) * m + m // 2
    if m >= 3:
        assert np.asarray(A7.sum(1)).ravel()[idx] == 0 and np.asarray(A27.sum(1)).ravel()[idx] == 0


def test_diffusion_generator_properties():
    A = O.diffusion7p(6, 5, 4, 1, 0.5, 2.0, 4).to_scipy()
    assert A.shape == (120, 120) and (A - A.T).nnz == 0
    off = A - sp.diags(A.diagonal())
    assert off.data.max() < 0 and -off.data.min() <= 2.0 and -off.data.max() >= 0.5
    # diagonal = sum of the six face conductivities: equal to |off-diagonal row
    # sum| inside the box, larger on rows that touch a Dirichlet face
    offsum = -np.asarray(off.sum(1)).ravel()
    assert np.all(A.diagonal() >= offsum * (1 - 1e-15))
    assert np.isclose(A.diagonal()[1 * 30 + 1 * 6 + 1], offsum[1 * 30 + 1 * 6 + 1], rtol=1e-14)
    B = O.diffusion7p(6, 5, 4, 1, 0.5, 2.0, 4).to_scipy()
    assert (A != B).nnz == 0  # seeded, reproducible


def test_mtx_reader_symmetric_expansion(golden):
    A = O.read_mtx(os.path.join(golden, "bcsstk09.mtx")).to_scipy()
    ref = mtx(os.path.join(golden, "bcsstk09.mtx"))
    assert abs(A - ref).max() == 0


# ------------------------------------------------------------ sparse kernels
def rand_spd(n, density, seed, ties=False):
    S = sp.random(n, n, density=density, random_state=seed)
    S = S + S.T
    S.data[:] = -(np.round(S.data * 3) / 3 if ties else S.data)
    S = S + sp.diags(np.asarray(abs(S).sum(1)).ravel() + 1.0)
    S = S.tocsr()
    S.sort_indices()
    return S


def test_spmv_ops_vs_dense():
    S = rand_spd(60, 0.1, 1)
    A = O.Csr.from_scipy(S)
    rng = np.random.default_rng(0)
    x, y = rng.normal(size=60), rng.normal(size=60)
    D = S.toarray()
    assert np.allclose(O.spmv(A, x, op=1), D @ x, rtol=1e-14, atol=1e-13)
    assert np.allclose(O.spmv(A, x, y, op=2), y - D @ x, rtol=1e-14, atol=1e-13)
    assert np.allclose(O.spmv(A, x, y, alpha=0.5, beta=-2.0, op=0), 0.5 * D @ x - 2 * y, rtol=1e-14, atol=1e-13)


def test_spgemm_and_transpose_vs_scipy():
    S1, S2 = rand_spd(80, 0.05, 2), sp.random(80, 50, density=0.1, random_state=3, format="csr")
    S2.sort_indices()
    A, B = O.Csr.from_scipy(S1), O.Csr.from_scipy(S2)
    C = O.spgemm(A, B).to_scipy()
    ref = (S1 @ S2).tocsr()
    ref.sort_indices()
    assert abs(C - ref).max() < 1e-13
    T = O.transpose(B).to_scipy()
    assert abs(T - S2.T).max() == 0
    assert abs(O.transpose(O.transpose(B)).to_scipy() - S2).max() == 0


# ----------------------------------------------------------------- matching
MATCH_CASES = [
    ("lap7_10", lambda: O.laplacian3d(7, 10, 10, 10)),
    ("lap7_9x7x5", lambda: O.laplacian3d(7, 9, 7, 5)),
    ("lap27_8", lambda: O.laplacian3d(27, 8, 8, 8)),
    ("diff_exp", lambda: O.diffusion7p(9, 8, 7, 0, 2.0, 0.0, 2)),
    ("diff_unif", lambda: O.diffusion7p(9, 8, 7, 1, 0.5, 2.0, 4)),
    ("rand_ties", lambda: O.Csr.from_scipy(rand_spd(300, 0.03, 5, ties=True))),
    ("rand", lambda: O.Csr.from_scipy(rand_spd(300, 0.03, 6))),
]


@pytest.mark.parametrize("name,make", MATCH_CASES)
def test_suitor_equals_greedy_and_invariants(name, make):
    """Sequential Suitor (the GPU semantics) == greedy matching on (W, max id, min id)."""
    A = make()
    n, _, _ = A.dims()
    w = np.random.default_rng(1).uniform(0.5, 2.0, n) if name.startswith("rand") else np.ones(n)
    M = O.suitor(A, w)
    G = O.greedy_matching(A, w)
    assert np.array_equal(M, G)
    # involution
    m = M >= 0
    assert np.all(M[M[m]] == np.nonzero(m)[0])
    # maximality over positive edges between unmatched vertices
    r, c, v = O.matching_weights(A, w)
    for i in np.nonzero(~m)[0]:
        for j in range(r[i], r[i + 1]):
            if v[j] > 0:
                assert M[c[j]] >= 0, "positive edge between two unmatched vertices"


def test_matching_weights_quadratic_form_identity():
    """|AH_ij| is the energy of the pair vector v = (w_i, -w_j) under the 2x2
    principal submatrix A2 of (i, j), relative to the diagonal energy of
    u = (w_i, w_j): |AH_ij| = |v' A2 v| / (u' diag(A2) u).  Computed here with
    numpy matrix products, independent of the oracle's formula code; W must be
    its log wherever it is above the clamp, and W is symmetric."""
    S = rand_spd(40, 0.15, 7)
    A = O.Csr.from_scipy(S)
    w = np.random.default_rng(3).uniform(0.5, 2, 40)
    r, c, v = O.matching_weights(A, w)
    W = sp.csr_matrix((v, c, r), shape=(40, 40))
    assert abs(W - W.T).max() == 0
    Sd = S.toarray()
    checked = 0
    for i in range(40):
        for j in np.nonzero(Sd[i])[0]:
            if j == i:
                continue
            A2 = Sd[np.ix_([i, j], [i, j])]
            vv = np.array([w[i], -w[j]])
            uu = np.array([w[i], w[j]])
            ah = abs(vv @ A2 @ vv) / (uu @ np.diag(np.diag(A2)) @ uu)
            assert np.isclose(np.exp(W[i, j]), ah, rtol=1e-12, atol=0), (i, j)
            checked += 1
    assert checked > 100


def test_tied_laplacian_pairs_along_x():
    """All weights tie on the constant-coefficient 7-point Laplacian: the greedy
    order (W, max id, min id) pairs every even x index with the next one."""
    M = O.suitor(O.laplacian3d(7, 10, 6, 4), np.ones(240))
    assert np.array_equal(M, np.arange(240) ^ 1)


def test_make_p_canonical_order():
    M = np.array([1, 0, -1, 5, -1, 3, -1], np.int32)
    w = np.array([1.0, 2.0, 3.0, 1.0, 0.0, 1.0, -2.0])
    agg, pv, nc = O.make_p(M, w)
    # pairs by master (0,1) -> 0, (3,5) -> 1; good singletons ascending: 2 -> 2, 6 -> 3
    assert nc == 4
    assert agg.tolist() == [0, 0, 2, 1, 2, 1, 3]  # bad singleton 4 -> (running count 3) - 1
    assert np.allclose(pv, [1 / np.sqrt(5), 2 / np.sqrt(5), 1.0, 1 / np.sqrt(2), 0.0, 1 / np.sqrt(2), -1.0])


# ---------------------------------------------------------------- hierarchy
@pytest.mark.parametrize("N,expect", [(125000, 1999), (2097152, 5119), (4096000, 6399), (64000, 1599), (1000, 399)])
def test_reference_coarse_size_truncation(N, expect):
    assert int(40 * np.power(float(N), 1.0 / 3.0)) == expect


def test_galerkin_matches_dense_triple_product():
    A = O.laplacian3d(7, 6, 5, 4)
    p = fcg_params("BCMG", aggrsweeps=0, max_coarse_size=10)
    P = O.Prec(A, p)
    S = A.to_scipy()
    for l in range(P.levels() - 1):
        agg, pv = P.level_P(l)
        Al = P.level_A(l).to_scipy().toarray()
        Pm = sp.csr_matrix((pv, agg, np.arange(len(agg) + 1)), shape=(len(agg), P.level_A(l + 1).dims()[0])).toarray()
        Ac = P.level_A(l + 1).to_scipy().toarray()
        assert np.allclose(Pm.T @ Al @ Pm, Ac, rtol=1e-13, atol=1e-12)
        assert np.allclose(Ac, Ac.T, rtol=0, atol=1e-13)
    assert S.shape[0] == 120


def test_hierarchy_shape_27p():
    A = O.laplacian3d(27, 16, 16, 16)
    P = O.Prec(A, fcg_params("BCMG"))
    sizes = [P.level_A(l).dims()[0] for l in range(P.levels())]
    assert sizes == [4096, 512]  # 3 pairwise sweeps -> 2x2x2 aggregates; 512 <= 639
    st = P.stats()
    assert st["max_coarse_size"] == 639 and abs(st["avg_cratio"] - 8.0) < 1e-12


# --------------------------------------------------------------- whole solve
@pytest.mark.parametrize("pts", [7, 27])
@pytest.mark.parametrize("prec", ["BCMG", "AFSAI", "L1_JACOBI", "NONE"])
def test_solve_matches_reference_fixture(golden, pts, prec):
    """launchAndCompare.mk: ||x_ref - x|| / ||x_ref|| <= 1e-5, A x = 1, x0 = 0."""
    A = O.laplacian3d(pts, 10, 10, 10)
    p = fcg_params(prec)
    P = O.Prec(A, p)
    x, info = O.fcg(A, P, p, np.ones(1000))
    ref = np.loadtxt(os.path.join(golden, f"lap3d_{pts}p_10x10x10_1x1x1_sol.txt"))
    assert info["retv"] == 0
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-5


@pytest.mark.parametrize("pts", [7, 27])
@pytest.mark.parametrize("grid", [(2, 1, 1), (2, 2, 1), (2, 2, 2)])
def test_solve_matches_block_fixture(golden, pts, grid):
    """The P x Q x R block-ordered problems (each 10^3 subdomain one block,
    laplacian.cu:14-25) against the reference's solution files."""
    P, Q, R = grid
    A = O.laplacian3d(pts, 10, 10, 10, P, Q, R)
    n = 1000 * P * Q * R
    p = fcg_params("BCMG")
    x, info = O.fcg(A, O.Prec(A, p), p, np.ones(n))
    ref = np.loadtxt(os.path.join(golden, f"lap3d_{pts}p_10x10x10_{P}x{Q}x{R}_sol.txt"))
    assert info["retv"] == 0
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-5


@pytest.mark.parametrize("prec", ["BCMG", "AFSAI", "L1_JACOBI"])
def test_solve_bcsstk09_fixture(golden, prec):
    A = O.read_mtx(os.path.join(golden, "bcsstk09.mtx"))
    n = A.dims()[0]
    p = fcg_params(prec, itnlim=5000)
    x, info = O.fcg(A, O.Prec(A, p), p, np.ones(n))
    ref = np.loadtxt(os.path.join(golden, "bcsstk09_sol.txt"))
    assert info["retv"] == 0
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-5


def test_fcg_reaches_1e8_and_counts_like_reference():
    A = O.laplacian3d(7, 20, 20, 20)
    n = 8000
    xs = np.random.default_rng(0).uniform(-1, 1, n)
    b = O.spmv(A, xs)
    p = fcg_params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1)
    x, info = O.fcg(A, O.Prec(A, p), p, b)
    r = b - O.spmv(A, x)
    assert np.linalg.norm(r) <= 1e-8 * info["del0"] * (1 + 1e-9)
    assert abs(info["exit_res"] - np.linalg.norm(r)) <= 1e-6 * np.linalg.norm(r) + 1e-12
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-6
    # itnlim stops with retv -1 and niter = itnlim + 1
    p2 = fcg_params("BCMG", rtol=1e-30, itnlim=3)
    _, info2 = O.fcg(A, O.Prec(A, p2), p2, b)
    assert info2["retv"] == -1 and info2["niter"] == 4


def test_fp32_hierarchy_oracle_converges():
    A = O.laplacian3d(7, 20, 20, 20)
    b = O.spmv(A, np.random.default_rng(1).uniform(-1, 1, 8000))
    p64 = fcg_params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1)
    p32 = fcg_params("BCMG", rtol=1e-8, prerelax_sweeps=1, postrelax_sweeps=1, precision=1)
    _, i64 = O.fcg(A, O.Prec(A, p64), p64, b)
    _, i32 = O.fcg(A, O.Prec(A, p32), p32, b)
    assert i32["retv"] == 0 and abs(i32["niter"] - i64["niter"]) <= 1


# ------------------------------------------------------------ CG (CG_HS.cu)
def cghs_params(prec, **kw):
    p = fcg_params(prec, **kw)
    p.solver = abi.SOLVER["CGHS"]
    return p


@pytest.mark.parametrize("pts", [7, 27])
@pytest.mark.parametrize("prec", ["BCMG", "AFSAI", "L1_JACOBI", "NONE"])
def test_cghs_matches_reference_fixture(golden, pts, prec):
    """The fixtures' A x = 1 solutions hold for any converged solver (launchAndCompare.mk, 1e-5)."""
    A = O.laplacian3d(pts, 10, 10, 10)
    p = cghs_params(prec)
    x, info = O.cghs(A, O.Prec(A, p), p, np.ones(1000))
    ref = np.loadtxt(os.path.join(golden, f"lap3d_{pts}p_10x10x10_1x1x1_sol.txt"))
    assert info["retv"] == 0
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-5


def textbook_pcg(S, b, apply_m, rtol, itnlim):
    """Preconditioned CG as printed in Saad, Iterative Methods, Alg. 9.1 (independent of the oracle's code)."""
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_m(r)
    p = z.copy()
    rz = r @ z
    tol = rtol * np.linalg.norm(r)
    for it in range(itnlim):
        Ap = S @ p
        a = rz / (p @ Ap)
        x += a * p
        r -= a * Ap
        if np.linalg.norm(r) < tol:
            return x, it + 1
        z = apply_m(r)
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    return x, itnlim


@pytest.mark.parametrize("prec", ["NONE", "L1_JACOBI"])
def test_cghs_matches_textbook_pcg(prec):
    """Iteration count and solution against Saad's PCG, with the l1-Jacobi preconditioner
    (4 sweeps from zero, divisor a_ii + sum_{j!=i} |a_ij|, l1jacobi.cu:484-523) restated in numpy."""
    oa = O.laplacian3d(27, 12, 11, 10)
    S = oa.to_scipy()
    n = S.shape[0]
    b = np.random.default_rng(7).normal(size=n)
    d = S.diagonal()
    m = d + (abs(S).sum(axis=1).A1 - abs(d))

    def l1j(r):
        u = np.zeros_like(r)
        for _ in range(4):
            u = u + (r - S @ u) / m
        return u

    apply_m = (lambda r: r.copy()) if prec == "NONE" else l1j
    p = cghs_params(prec, rtol=1e-10, itnlim=2000, l1jacsweeps=4)
    x, info = O.cghs(oa, O.Prec(oa, p), p, b)
    xt, it = textbook_pcg(S, b, apply_m, 1e-10, 2000)
    assert info["retv"] == 0 and abs(info["niter"] - it) <= 1
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-8
    if prec == "NONE":  # scipy's CG stops on ||r|| <= rtol ||b|| from x0 = 0: the same test up to the tie
        import scipy.sparse.linalg as spl
        count = [0]
        spl.cg(S, b, rtol=1e-10, maxiter=2000, callback=lambda _: count.__setitem__(0, count[0] + 1))
        assert abs(count[0] - info["niter"]) <= 1


def test_cghs_finite_termination_on_distinct_eigenvalues():
    """CG on an SPD matrix with m distinct eigenvalues converges in m steps (exact arithmetic)."""
    ev = np.array([1.0, 2.0, 5.0, 7.0, 11.0])
    n = 400
    diag = ev[np.arange(n) % 5]
    oa = O.Csr.from_arrays(n, n, np.arange(n + 1), np.arange(n), diag)
    b = np.random.default_rng(2).uniform(0.5, 1.5, n)
    p = cghs_params("NONE", rtol=1e-12, itnlim=50)
    x, info = O.cghs(oa, O.Prec(oa, p), p, b)
    assert info["retv"] == 0 and info["niter"] in (5, 6)
    assert np.allclose(x, b / diag, rtol=1e-10)


def test_cghs_stop_rules():
    A = O.laplacian3d(7, 12, 12, 12)
    b = np.ones(A.dims()[0])
    # itnlim: retv -1 and niter = itnlim (iter + 1 + excnlim, CG_HS.cu:165-171)
    p = cghs_params("L1_JACOBI", rtol=1e-30, itnlim=3)
    _, info = O.cghs(A, O.Prec(A, p), p, b)
    assert info["retv"] == -1 and info["niter"] == 3
    # stop_criterion 0: absolute residual, delta0 = 1 (CG_HS.cu:74-79)
    p = cghs_params("L1_JACOBI", rtol=1e-6, stop_criterion=0)
    x, info = O.cghs(A, O.Prec(A, p), p, b)
    assert info["retv"] == 0 and info["del0"] == 1.0
    true_res = np.linalg.norm(b - O.spmv(A, x))
    assert info["exit_res"] < 1e-6 and abs(true_res - info["exit_res"]) <= 1e-9


# ------------------------------------------------- s-step CG (CGs.cu, mpk.cu)
def cgs_params(prec, s, **kw):
    p = fcg_params(prec, **kw)
    p.solver = abi.SOLVER["CGS"]
    p.sstep = s
    return p


@pytest.mark.parametrize("prec", ["NONE", "L1_JACOBI", "BCMG", "AFSAI"])
def test_sstep_one_is_pcg(prec):
    """With s = 1 the s-step recurrences are PCG's: same counts and iterates as CGHS."""
    A = O.laplacian3d(7, 12, 12, 11)
    b = np.random.default_rng(3).normal(size=A.dims()[0])
    pc = cghs_params(prec, rtol=1e-10, itnlim=3000)
    ps = cgs_params(prec, 1, rtol=1e-10, itnlim=3000)
    xc, ic = O.cghs(A, O.Prec(A, pc), pc, b)
    xs, is_ = O.cgs(A, O.Prec(A, ps), ps, b)
    assert is_["retv"] == 0 and abs(is_["niter"] - ic["niter"]) <= 1
    assert np.linalg.norm(xs - xc) / np.linalg.norm(xc) < 1e-8


@pytest.mark.parametrize("prec,s", [("NONE", 2), ("NONE", 3), ("L1_JACOBI", 2), ("L1_JACOBI", 3), ("BCMG", 2)])
def test_sstep_tracks_cg_every_s_steps(prec, s):
    """s-step CG reproduces CG's iterate every s steps (exact arithmetic): about niter_CG / s
    outer iterations, and the same solution, against Saad's PCG restated in numpy."""
    oa = O.laplacian3d(7, 10, 9, 8)
    S = oa.to_scipy()
    n = S.shape[0]
    b = np.random.default_rng(4).normal(size=n)
    p = cgs_params(prec, s, rtol=1e-8, itnlim=3000)
    x, info = O.cgs(oa, O.Prec(oa, p), p, b)
    assert info["retv"] == 0
    P = O.Prec(oa, cghs_params(prec))
    _, it = textbook_pcg(S, b, lambda r: r.copy() if prec == "NONE" else P.apply(r), 1e-8, 3000)
    assert abs(info["niter"] - int(np.ceil(it / s))) <= 1
    assert np.linalg.norm(b - S @ x) <= 1e-8 * np.linalg.norm(b) * (1 + 1e-6)


@pytest.mark.parametrize("pts", [7, 27])
def test_sstep_matches_reference_fixture(golden, pts):
    A = O.laplacian3d(pts, 10, 10, 10)
    p = cgs_params("BCMG", 2)
    x, info = O.cgs(A, O.Prec(A, p), p, np.ones(1000))
    ref = np.loadtxt(os.path.join(golden, f"lap3d_{pts}p_10x10x10_1x1x1_sol.txt"))
    assert info["retv"] == 0
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-5


def test_sstep_residual_options_and_breakdown():
    A = O.laplacian3d(7, 9, 9, 9)
    n = A.dims()[0]
    b = np.random.default_rng(5).normal(size=n)
    base = cgs_params("L1_JACOBI", 2, rtol=1e-8)
    xb, ib = O.cgs(A, O.Prec(A, base), base, b)
    # ru_res = 0: the residual is recomputed as b - A x after each update (CGs.cu:148-152)
    p0 = cgs_params("L1_JACOBI", 2, rtol=1e-8, ru_res=0)
    x0, i0 = O.cgs(A, O.Prec(A, p0), p0, b)
    # rec_res_int: periodic recomputation (CGs.cu:121-131)
    p1 = cgs_params("L1_JACOBI", 2, rtol=1e-8, rec_res_int=3)
    x1, i1 = O.cgs(A, O.Prec(A, p1), p1, b)
    for x, i in ((x0, i0), (x1, i1)):
        assert i["retv"] == 0 and abs(i["niter"] - ib["niter"]) <= 1
        assert np.linalg.norm(x - xb) / np.linalg.norm(xb) < 1e-6
    # a negative definite operator: the s x s Cholesky fails, retv > 0 (scalarWorkMO)
    N = O.Csr.from_arrays(50, 50, np.arange(51), np.arange(50), -np.linspace(1, 2, 50))
    pn = cgs_params("NONE", 1)
    _, inn = O.cgs(N, O.Prec(N, pn), pn, np.ones(50))
    assert inn["retv"] > 0


# ------------------------------------------- pipelined s-step CG (pipelinedCGs.cu)
def pipe_params(prec, s, **kw):
    p = cgs_params(prec, s, **kw)
    p.solver = abi.SOLVER["PIPELINED_CGS"]
    return p


@pytest.mark.parametrize("prec", ["NONE", "L1_JACOBI", "BCMG", "AFSAI"])
@pytest.mark.parametrize("s", [1, 2, 3])
def test_pipelined_sstep_equals_sstep(prec, s):
    """The pipelined form re-associates the s-step recurrences (the basis of the next block is
    recombined from the current one): same outer iterations and solution as CGs, and PCG's for s = 1."""
    A = O.laplacian3d(7, 11, 10, 12)
    b = np.random.default_rng(6).normal(size=A.dims()[0])
    ps = cgs_params(prec, s, rtol=1e-9, itnlim=3000)
    pp = pipe_params(prec, s, rtol=1e-9, itnlim=3000)
    xs, is_ = O.cgs(A, O.Prec(A, ps), ps, b)
    xp, ip = O.pipecgs(A, O.Prec(A, pp), pp, b)
    assert ip["retv"] == 0 and abs(ip["niter"] - is_["niter"]) <= 1
    assert np.linalg.norm(xp - xs) / np.linalg.norm(xs) < 1e-7
    if s == 1:
        pc = cghs_params(prec, rtol=1e-9, itnlim=3000)
        _, ic = O.cghs(A, O.Prec(A, pc), pc, b)
        assert abs(ip["niter"] - ic["niter"]) <= 1


@pytest.mark.parametrize("pts", [7, 27])
def test_pipelined_sstep_matches_reference_fixture(golden, pts):
    A = O.laplacian3d(pts, 10, 10, 10)
    p = pipe_params("BCMG", 2)
    x, info = O.pipecgs(A, O.Prec(A, p), p, np.ones(1000))
    ref = np.loadtxt(os.path.join(golden, f"lap3d_{pts}p_10x10x10_1x1x1_sol.txt"))
    assert info["retv"] == 0
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) <= 1e-5


# -------------------------------------------------------------------- AFSAI
def test_afsai_factor_invariants():
    A = O.laplacian3d(7, 8, 7, 6)
    P = O.Prec(A, fcg_params("AFSAI"))
    G = P.afsai_G().to_scipy()
    S = A.to_scipy()
    assert sp.triu(G, 1).nnz == 0
    assert np.diff(G.indptr).max() <= 31
    d = 1 / np.sqrt(S.diagonal())
    As = sp.diags(d) @ S @ sp.diags(d)
    dd = (G @ As @ G.T).diagonal()
    assert np.allclose(dd, 1.0, rtol=1e-10)
    # first step on the tied 7-point stencil: the entry with the smallest column wins
    assert G.indices[G.indptr[S.shape[0] - 1]] < S.shape[0] - 1
