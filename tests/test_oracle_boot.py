"""Pins of the oracle's bootstrap (several hierarchies) by what the mathematics fixes (no GPU).

The reference grows hierarchies in `bootstrap` (bootstrap.cu:11-46) while the
estimated convergence ratio exceeds conv_ratio and fewer than max_hrc exist,
and applies them in `bcmg_apply` (bcmg.cu:51-90).  It never updates the
estimate (Boot::init sets 1.0, AMG.cu:393-399) and stops at one hierarchy
(assert, bootstrap.cu:37-38); the estimate and the next smooth vector follow
BootCMatch: solver_it = 15 applications (AMG.cu:342) of the composition to
A x = 0 from a random x.

Every check below is linear algebra on dense operators built from the oracle's
prolongators with numpy (tests/dense_cycle.py), not a retyping of the oracle:

* each hierarchy's cycle is the dense cycle of its own prolongators;
* MULTIPLICATIVE: I - B A = E_{n-1} ... E_0 with E_h = I - B_h A;
  SYMMETRIZED_MULTIPLICATIVE: I - B A = E_0 ... E_{n-1} ... E_0 and B symmetric;
  ADDITIVE: B = sum_h B_h, symmetric;
  from a nonzero guess u the composition returns E u + B r;
* the estimated ratio is sqrt(x_15' A x_15 / x_14' A x_14) for x_k = E^k x_0,
  which for an A-self-adjoint E never exceeds its spectral radius;
* hierarchy h+1 is the one-hierarchy setup from w = x_15 / max |x_15|.
"""
import numpy as np
import pytest

import oracle_py as O
from bootcmatchgx_amd import abi
from dense_cycle import DenseHierarchy, np_cycle

COMPS = ["MULTIPLICATIVE", "SYMMETRIZED_MULTIPLICATIVE", "ADDITIVE"]


def bparams(comp, max_hrc=3, conv_ratio=0.0, **kw):
    p = abi.reference_defaults()
    p.solver = abi.SOLVER["FCG"]
    p.preconditioner = abi.PREC["BCMG"]
    p.dispnorm = 0
    p.bootstrap_type = abi.BOOTSTRAP[comp]
    p.max_hrc = max_hrc
    p.conv_ratio = conv_ratio
    p.aggrsweeps = 0
    p.max_coarse_size = 12
    p.prerelax_sweeps = 1
    p.postrelax_sweeps = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def matrix():
    return O.diffusion7p(8, 7, 6, 1, 0.5, 2.0, 21)


def dense_B(A, P, p, h):
    """B_h as a dense matrix from the dense cycle of hierarchy h's prolongators."""
    H = DenseHierarchy.from_oracle(A, P, h)
    n = H.A[0].shape[0]
    return np_cycle(H, p, 0, np.eye(n))


def composed_error(Es, comp):
    """The error propagator of the composition, E_h applied in order h = 0, 1, ..."""
    order = list(range(len(Es)))
    if comp == "SYMMETRIZED_MULTIPLICATIVE":
        order += list(range(len(Es) - 2, -1, -1))
    E = np.eye(Es[0].shape[0])
    for h in order:
        E = Es[h] @ E
    return E


def test_one_hierarchy_is_the_plain_cycle():
    A = matrix()
    n = A.dims()[0]
    r = np.random.default_rng(1).normal(size=n)
    base = O.Prec(A, bparams("MULTIPLICATIVE", max_hrc=1)).apply(r)
    for comp in COMPS:
        P = O.Prec(A, bparams(comp, max_hrc=1))
        assert P.num_hrc() == (1, 1.0)  # never estimated: Boot::init's 1.0
        assert np.array_equal(P.apply(r), base)


@pytest.mark.parametrize("comp", COMPS)
def test_compositions_equal_their_closed_forms(comp):
    A = matrix()
    Ad = A.to_scipy().toarray()
    n = Ad.shape[0]
    p = bparams(comp, max_hrc=3)
    P = O.Prec(A, p)
    nh, ratio = P.num_hrc()
    assert nh == 3 and 0 < ratio < 1
    Bs = []
    for h in range(nh):
        Bh = dense_B(A, P, p, h)
        Bo = np.stack([P.apply_hrc(h, e) for e in np.eye(n)], axis=1)
        assert np.abs(Bo - Bh).max() <= 1e-12 * np.abs(Bh).max()
        Bs.append(Bh)
    # the hierarchies differ: each starts its matching from another smooth vector
    assert not np.array_equal(P.hrc_level_P(0, 0)[0], P.hrc_level_P(1, 0)[0])
    B = np.stack([P.apply(e) for e in np.eye(n)], axis=1)
    Es = [np.eye(n) - Bh @ Ad for Bh in Bs]
    if comp == "ADDITIVE":
        ref = sum(Bs)
        assert np.abs(B - ref).max() <= 1e-12 * np.abs(ref).max()
    else:
        lhs = np.eye(n) - B @ Ad
        ref = composed_error(Es, comp)
        assert np.abs(lhs - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())
    if comp != "MULTIPLICATIVE":
        assert np.abs(B - B.T).max() <= 1e-12 * np.abs(B).max()
    # from a nonzero guess u: E u + B r (E = I - B A)
    rng = np.random.default_rng(2)
    r, u = rng.normal(size=n), rng.normal(size=n)
    got = P.apply(r, u)
    want = (np.eye(n) - B @ Ad) @ u + B @ r
    assert np.linalg.norm(got - want) <= 1e-11 * np.linalg.norm(want)


@pytest.mark.parametrize("comp", COMPS)
def test_ratio_and_next_smooth_vector(comp):
    A = matrix()
    Ad = A.to_scipy().toarray()
    n = Ad.shape[0]
    p = bparams(comp, max_hrc=3)
    P = O.Prec(A, p)
    Bs = [dense_B(A, P, p, h) for h in range(2)]
    Es = [np.eye(n) - Bh @ Ad for Bh in Bs]
    for built in (1, 2):  # the estimate over the first `built` hierarchies made hierarchy `built`
        E = composed_error(Es[:built], comp) if comp != "ADDITIVE" else np.eye(n) - sum(Bs[:built]) @ Ad
        x = O.boot_random(n, built)
        xs = [x]
        for _ in range(15):
            xs.append(E @ xs[-1])
        a = [v @ Ad @ v for v in xs]
        ratio = np.sqrt(a[15] / a[14])
        # the estimate itself is only exported for the last one; the last
        # hierarchy's matching comes from the smooth vector of this estimate
        w = xs[15] / np.abs(xs[15]).max()
        ref = O.Prec(A, bparams(comp, max_hrc=1), w=w)
        assert P.hrc_levels(built) == ref.levels()
        for l in range(ref.levels() - 1):
            ga, gv = P.hrc_level_P(built, l)
            ra, rv = ref.level_P(l)
            assert np.array_equal(ga, ra)
            assert np.allclose(gv, rv, rtol=1e-9, atol=0)
        if built == 2:
            assert abs(P.num_hrc()[1] - ratio) <= 1e-9 * ratio
        if comp != "MULTIPLICATIVE" or built == 1:
            # E is A-self-adjoint here: ||E y||_A <= rho(E) ||y||_A
            rho = np.abs(np.linalg.eigvals(E)).max()
            assert ratio <= rho * (1 + 1e-9)
            assert ratio >= 0.5 * rho


def test_stops_once_the_ratio_is_good_enough():
    A = matrix()
    p = bparams("SYMMETRIZED_MULTIPLICATIVE", max_hrc=4)
    nh, ratio = O.Prec(A, p).num_hrc()
    assert nh == 4
    one = O.Prec(A, bparams("SYMMETRIZED_MULTIPLICATIVE", max_hrc=2, conv_ratio=0.999)).num_hrc()
    # the first estimate is below 0.999: no second hierarchy, and the estimate is reported
    assert one[0] == 1 and 0 < one[1] <= 0.999


@pytest.mark.parametrize("comp", COMPS)
def test_bootstrap_fcg_solves(comp):
    """FCG with each composition reaches the tolerance on a 27-point Laplacian and the
    reference's 10^3 solution fixture (compareSolutions at 1e-5)."""
    import os

    A = O.laplacian3d(7, 10, 10, 10)
    p = bparams(comp, max_hrc=2, max_coarse_size=0, aggrsweeps=2, prerelax_sweeps=4, postrelax_sweeps=4)
    p.rtol = 1e-8
    P = O.Prec(A, p)
    assert P.num_hrc()[0] == 2
    n = A.dims()[0]
    x, info = O.fcg(A, P, p, np.ones(n))
    ref = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "lap3d_7p_10x10x10_1x1x1_sol.txt"))
    assert info["retv"] == 0
    assert np.linalg.norm(x - ref) <= 1e-5 * np.linalg.norm(ref)
    one = O.fcg(A, O.Prec(A, bparams(comp, max_hrc=1, max_coarse_size=0, aggrsweeps=2, prerelax_sweeps=4,
                                         postrelax_sweeps=4, rtol=1e-8)), p, np.ones(n))[1]
    # a multiplicative composition is at least one full cycle more per application;
    # the additive one sums corrections from one residual and may take an iteration more
    assert info["niter"] <= one["niter"] + (1 if comp == "ADDITIVE" else 0)
