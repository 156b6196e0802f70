"""Pins of the oracle's solve path by what the mathematics fixes (no GPU).

* The V-cycle (GAMG_cycle.cu:58-236 with l1-Jacobi smoothing,
  l1jacobi.cu:344-479, and the divisor M_i = sum_j |a_ij| of relaxPrepare,
  matchingAggregation.cu:40-60) is restated as dense linear algebra
  (tests/dense_cycle.py): smoothing x <- x + M^{-1}(f - A x), the Galerkin
  operator P^T A P formed with numpy from the oracle's prolongator, the coarse
  correction x <- x + P x_c.
  Its output must equal the oracle's prec_apply; with an exact coarse solve it
  must equal the closed-form two-grid operator, and the V(nu, nu) operator must
  be symmetric.
* FCG(1) (FCG.cu:27-175) with a fixed symmetric positive definite
  preconditioner produces the iterates of preconditioned CG (Saad, Iterative
  Methods, Alg. 9.1): same iteration count (+-1) and iterates to 1e-8, for no
  preconditioner, l1-Jacobi and the BCMG V-cycle (the numpy restatement above
  as the PCG's preconditioner, so nothing is shared with the oracle's code).
* AFSAI (kapGradCUDA_WiP.cu:189-368, afsai.cu:287-1090) replayed row by row
  with dense algebra: each step adds the column of largest |Kaporin gradient|
  (A_s g)_j over j < i outside the pattern, solves A_s[P,P] x = -A_s[P,i], and
  stops on the Kaporin-number test; G must equal the oracle's pattern and values.

A plausible mistake in the oracle (the diagonal instead of the l1 row sum as
divisor, a sweep count off by one, a dropped coarse correction, a wrong theta /
theta2 recurrence, a wrong AFSAI gradient or argmax) fails one of these tests.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_py as O
from bootcmatchgx_amd import abi
from dense_cycle import DenseHierarchy, np_cycle


def params(prec="BCMG", solver="FCG", **kw):
    p = abi.reference_defaults()
    p.solver = abi.SOLVER[solver]
    p.preconditioner = abi.PREC[prec]
    p.dispnorm = 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------- V-cycle
def diffusion(nx, ny, nz, seed):
    return O.diffusion7p(nx, ny, nz, 1, 0.5, 2.0, seed)


CYCLES = [
    ("V", dict(prerelax_sweeps=1, postrelax_sweeps=1)),
    ("V", dict(prerelax_sweeps=2, postrelax_sweeps=3)),
    ("V", dict(prerelax_sweeps=4, postrelax_sweeps=4, relaxnumber_coarse=7)),
    ("W", dict(prerelax_sweeps=1, postrelax_sweeps=1)),
    ("H", dict(prerelax_sweeps=2, postrelax_sweeps=1)),
    ("VARIABLE_V", dict(prerelax_sweeps=1, postrelax_sweeps=2)),
]


@pytest.mark.parametrize("cycle,kw", CYCLES)
@pytest.mark.parametrize("mat", ["lap27", "diff7"])
def test_cycle_equals_dense_restatement(cycle, kw, mat):
    """prec_apply of the FP64 hierarchy (u = 0 on entry, as FCG zero-fills it) is the
    dense cycle built from P alone, to 1e-12, on hierarchies of three or more levels."""
    A = O.laplacian3d(27, 9, 8, 7) if mat == "lap27" else diffusion(10, 9, 8, 11)
    p = params(cycle_type=abi.CYCLE[cycle + "_CYCLE"], aggrsweeps=0, max_coarse_size=12, **kw)
    P = O.Prec(A, p)
    H = DenseHierarchy.from_oracle(A, P)
    assert H.levels() >= 3
    for l in range(1, H.levels()):  # the oracle's Galerkin levels are the dense triple products
        Ao = P.level_A(l).to_scipy().toarray()
        assert np.allclose(Ao, H.A[l], rtol=1e-12, atol=1e-13 * np.abs(H.A[l]).max())
    rng = np.random.default_rng(5)
    for _ in range(2):
        r = rng.normal(size=H.A[0].shape[0])
        u = P.apply(r)
        ref = np_cycle(H, p, 0, r)
        assert np.linalg.norm(u - ref) <= 1e-12 * np.linalg.norm(ref)


def test_fp32_cycle_equals_dense_restatement_to_fp32_rounding():
    """The FP32 hierarchy's cycle (values, divisors, transfers and level vectors in
    FP32; the coarsest sweeps in FP64) is the same operator up to FP32 rounding."""
    A = diffusion(10, 9, 8, 12)
    p = params(aggrsweeps=0, max_coarse_size=12, prerelax_sweeps=1, postrelax_sweeps=1, precision=1)
    P = O.Prec(A, p)
    H = DenseHierarchy.from_oracle(A, P)
    r = np.random.default_rng(6).normal(size=H.A[0].shape[0])
    ref = np_cycle(H, p, 0, r)
    err = np.linalg.norm(P.apply(r) - ref) / np.linalg.norm(ref)
    assert err <= 2e-5, err
    assert err > 0  # it really ran in FP32


@pytest.mark.parametrize("cycle,kw", CYCLES)
def test_fp32_post_galerkin_cycle_equals_dense_restatement(cycle, kw):
    """The FP32 cycle whose first post-smoothing sweep takes its residual from the
    pre-smoothing residual and the product A P (f - A (x + P e) = (f - A x) - (A P) e)
    is the same operator as the plain cycle: both agree with the dense restatement to
    FP32 rounding, and the two FP32 paths really differ (they round differently)."""
    A = diffusion(10, 9, 8, 14)
    base = dict(cycle_type=abi.CYCLE[cycle + "_CYCLE"], aggrsweeps=0, max_coarse_size=12, precision=1, **kw)
    pg = O.Prec(A, params(post_galerkin=1, **base))
    plain = O.Prec(A, params(post_galerkin=0, **base))
    H = DenseHierarchy.from_oracle(A, pg)
    assert H.levels() >= 3
    r = np.random.default_rng(7).normal(size=H.A[0].shape[0])
    ref = np_cycle(H, params(**base), 0, r)
    u_pg, u_plain = pg.apply(r), plain.apply(r)
    nr = np.linalg.norm(ref)
    assert np.linalg.norm(u_pg - ref) <= 2e-5 * nr
    assert np.linalg.norm(u_plain - ref) <= 2e-5 * nr
    assert np.linalg.norm(u_pg - u_plain) > 0


def test_two_grid_with_exact_coarse_solve_is_closed_form():
    """Two levels and enough coarsest sweeps to converge: the cycle is the two-grid
    method whose error propagation is E = S^nu2 (I - P A_c^{-1} P^T A) S^nu1, with
    the smoothing iteration S = I - M^{-1} A, so B = (I - E) A^{-1}."""
    A = diffusion(6, 5, 6, 13)
    nu1, nu2 = 2, 1
    p = params(aggrsweeps=2, max_levels=2, max_coarse_size=10, prerelax_sweeps=nu1, postrelax_sweeps=nu2,
               relaxnumber_coarse=4000)
    P = O.Prec(A, p)
    assert P.levels() == 2
    agg, pv = P.level_P(0)
    S_ = A.to_scipy().toarray()
    n = S_.shape[0]
    nc = P.level_A(1).dims()[0]
    Pm = np.zeros((n, nc))
    Pm[np.arange(n), agg] = pv
    Minv = np.diag(1.0 / np.abs(S_).sum(axis=1))
    I = np.eye(n)
    S = I - Minv @ S_
    # error propagation E = S^nu2 (I - P A_c^{-1} P^T A) S^nu1, and B = (I - E) A^{-1}
    Ac = Pm.T @ S_ @ Pm
    C = I - Pm @ np.linalg.solve(Ac, Pm.T @ S_)
    E = np.linalg.matrix_power(S, nu2) @ C @ np.linalg.matrix_power(S, nu1)
    B = (I - E) @ np.linalg.inv(S_)
    rng = np.random.default_rng(7)
    for _ in range(3):
        r = rng.normal(size=n)
        u = P.apply(r)
        assert np.linalg.norm(u - B @ r) <= 1e-11 * np.linalg.norm(B @ r)


@pytest.mark.parametrize("nu", [1, 2])
def test_symmetric_cycle_operator(nu):
    """V(nu, nu) with a symmetric coarsest solve is a symmetric operator: u'Bv = v'Bu."""
    A = diffusion(9, 8, 8, 14)
    p = params(prerelax_sweeps=nu, postrelax_sweeps=nu)
    P = O.Prec(A, p)
    rng = np.random.default_rng(8)
    n = A.dims()[0]
    for _ in range(3):
        u, v = rng.normal(size=n), rng.normal(size=n)
        a, b = u @ P.apply(v), v @ P.apply(u)
        assert abs(a - b) <= 1e-12 * (abs(a) + abs(b))


# ------------------------------------------------------------------- FCG
def textbook_pcg(S, b, apply_m, rtol, itnlim):
    """Preconditioned CG as printed in Saad, Iterative Methods, Alg. 9.1."""
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
        if np.linalg.norm(r) <= tol:
            return x, it + 1
        z = apply_m(r)
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    return x, itnlim


@pytest.mark.parametrize("prec", ["NONE", "L1_JACOBI", "BCMG", "BCMG22"])
def test_fcg_equals_pcg_for_fixed_spd_preconditioner(prec):
    """FCG(1) with a fixed SPD preconditioner is PCG: counts +-1, iterates 1e-8."""
    A = diffusion(11, 10, 9, 15)
    S = A.to_scipy()
    n = S.shape[0]
    b = np.random.default_rng(9).normal(size=n)
    kw = dict(rtol=1e-10, itnlim=3000, l1jacsweeps=3)
    if prec.startswith("BCMG"):
        nu = 2 if prec == "BCMG22" else 1
        kw.update(prerelax_sweeps=nu, postrelax_sweeps=nu)
    p = params("BCMG" if prec.startswith("BCMG") else prec, **kw)
    P = O.Prec(A, p)
    x, info = O.fcg(A, P, p, b)
    if prec == "NONE":
        apply_m = lambda r: r.copy()
    elif prec == "L1_JACOBI":  # l1jacobi.cu:484-523: divisor a_ii + sum_{j != i} |a_ij|, sweeps from 0
        d = S.diagonal()
        m = d + (abs(S).sum(axis=1).A1 - abs(d))

        def apply_m(r):
            u = np.zeros_like(r)
            for _ in range(3):
                u = u + (r - S @ u) / m
            return u
    else:
        H = DenseHierarchy.from_oracle(A, P)
        apply_m = lambda r: np_cycle(H, p, 0, r)
    xt, it = textbook_pcg(S, b, apply_m, 1e-10, 3000)
    assert info["retv"] == 0
    assert abs(info["niter"] - it) <= 1, (info["niter"], it)
    assert np.linalg.norm(x - xt) <= 1e-8 * np.linalg.norm(xt)


def test_fcg_flexibility_with_variable_preconditioner():
    """FCG keeps the new direction A-orthogonal to the previous one even when the
    preconditioner changes between iterations (the reason for gamma/theta); the
    FP32 V-cycle is such a preconditioner, and FCG still reaches the FP64 bar."""
    A = diffusion(12, 11, 10, 16)
    b = np.random.default_rng(10).normal(size=A.dims()[0])
    p = params(rtol=1e-10, prerelax_sweeps=1, postrelax_sweeps=1, precision=1)
    x, info = O.fcg(A, O.Prec(A, p), p, b)
    S = A.to_scipy()
    assert info["retv"] == 0
    assert np.linalg.norm(b - S @ x) <= 1e-10 * np.linalg.norm(b) * (1 + 1e-6)


# ------------------------------------------------------------------- AFSAI
def afsai_replay(S, maxiter=30, eps=1e-3):
    """Row-by-row dense replay of adaptive FSAI on A_s = D^{-1/2} A D^{-1/2}."""
    A = S.toarray()
    d = 1.0 / np.sqrt(np.diag(A))
    As = A * d[:, None] * d[None, :]
    n = A.shape[0]
    rows = []
    for i in range(n):
        pat, x, dko = [], np.zeros(0), 0.0
        dk = 0.0
        for _ in range(maxiter):
            g = As[i, :i].copy()  # (A_s [x; 1])_j for j < i, x on the pattern
            if pat:
                g += x @ As[np.ix_(pat, np.arange(i))]
                g[pat] = 0.0
            if i == 0 or len(pat) >= maxiter or not np.any(g):
                break
            j = int(np.argmax(np.abs(g)))  # first maximum: ties to the smallest column
            pat = sorted(pat + [j])
            rhs = -As[pat, i]
            x = np.linalg.solve(As[np.ix_(pat, pat)], rhs)
            dk = rhs @ x
            stop = abs(dk - dko) < eps * dko or dk == 0.0
            dko = dk
            if stop:
                break
        sc = As[i, i] - dk
        sc = 1.0 / np.sqrt(sc) if sc > 0 else 1.0
        rows.append((pat + [i], np.concatenate([sc * x, [sc]]) if pat else np.array([sc])))
    return rows


@pytest.mark.parametrize("name", ["diff7", "rand"])
def test_afsai_pattern_is_kaporin_argmax(name):
    if name == "diff7":
        S = diffusion(6, 5, 4, 17).to_scipy()
    else:
        rng = np.random.default_rng(18)
        B = sp.random(90, 90, density=0.06, random_state=rng)
        B = B + B.T
        S = (B + sp.diags(abs(B).sum(axis=1).A1 + 1.0)).tocsr()
        S.sort_indices()
    A = O.Csr.from_scipy(S)
    G = O.Prec(A, params("AFSAI")).afsai_G().to_scipy()
    rows = afsai_replay(S)
    grown = 0
    for i, (cols, vals) in enumerate(rows):
        gc = G.indices[G.indptr[i]:G.indptr[i + 1]]
        gv = G.data[G.indptr[i]:G.indptr[i + 1]]
        assert gc.tolist() == cols, (i, gc, cols)
        assert np.allclose(gv, vals, rtol=1e-10, atol=1e-13), i
        grown += len(cols) > 2
    assert grown > 10  # the pattern grows beyond one step on many rows
