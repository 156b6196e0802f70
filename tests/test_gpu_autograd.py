"""The torch.autograd shim (x = A^{-1} b with the adjoint solve in backward)
on the 7-point diffusion operator with per-face coefficients U[0.5, 2]
(seed 4) at 40^3, loss L = 0.5 ||x - x_target||^2 with x_target ~ N(0,1)
(seed 5) and b ~ N(0,1) (seed 6).

The gradients are pinned two ways that do not share code with the shim:
  * a direct sparse LU factorisation (SuperLU through scipy) of A gives
    x, lambda = A^{-T}(x - x_target), dL/db = lambda and
    dL/dA_ij = -lambda_i x_j, to which the shim must agree within 1e-6
    relative (norm over all entries); the same reference with the outer
    product transposed (-lambda_j x_i) must NOT agree, so a swapped operand
    fails the test;
  * central finite differences (h = 1e-6) of L in 200 random stored values
    and 20 random entries of b, within 1e-6 relative (norm over the sample).
Solves use rtol 1e-12 (1e-14 for the finite differences) so that the
solver's truncation is far below the bar."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import bootcmatchgx_amd as bx

AG = pytest.importorskip("bootcmatchgx_amd.autograd")
torch = AG.torch

pytestmark = pytest.mark.gpu

H = 1e-6


def c4_problem(N):
    n, row, col, val = bx.diffusion7p(N, N, N, 1, 0.5, 2.0, 4)
    xt = np.random.default_rng(5).normal(size=n)
    b = np.random.default_rng(6).normal(size=n)
    return n, row, col, val, xt, b


def tight_params():
    return bx.Params(solver="FCG", preconditioner="BCMG", rtol=1e-12, itnlim=5000, dispnorm=0)


def loss_difference(xp, xm, xt):
    """L(x+) - L(x-) for L = 0.5 ||x - xt||^2, as 0.5 sum (x+ - x-)(x+ + x- - 2 xt):
    the same number without subtracting two losses of size ~1e5."""
    xp, xm, xt = (np.asarray(a, np.float64).ravel() for a in (xp, xm, xt))
    return 0.5 * math.fsum((xp - xm) * (xp + xm - 2.0 * xt))


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module")
def c4(ctx):
    n, row, col, val, xt, b = c4_problem(40)
    S = sp.csr_matrix((val, col, row), shape=(n, n))
    lu = spla.splu(S.tocsc())
    x_ref = lu.solve(b)
    lam_ref = lu.solve(x_ref - xt, trans="T")
    rows = np.repeat(np.arange(n), np.diff(row))
    system = AG.SPDSystem(n, row, col, val, tight_params(), ctx=ctx)
    yield dict(n=n, row=row, col=col, val=val, xt=xt, b=b, rows=rows, x_ref=x_ref, lam_ref=lam_ref, system=system)
    system.close()


def test_adjoint_gradient_matches_direct_factorisation(c4):
    s = c4
    v = torch.tensor(s["val"], requires_grad=True)
    bt = torch.tensor(s["b"], requires_grad=True)
    x = AG.sparse_solve(v, bt, s["system"])
    L = 0.5 * ((x - torch.tensor(s["xt"])) ** 2).sum()
    L.backward()
    assert rel(x.detach().numpy(), s["x_ref"]) < 1e-9
    g_ref = -s["lam_ref"][s["rows"]] * s["x_ref"][s["col"]]
    g_swapped = -s["lam_ref"][s["col"]] * s["x_ref"][s["rows"]]
    print(f"direct reference: x {rel(x.detach().numpy(), s['x_ref']):.3e}, dL/db {rel(bt.grad.numpy(), s['lam_ref']):.3e}, "
          f"dL/dA {rel(v.grad.numpy(), g_ref):.3e}")
    assert rel(bt.grad.numpy(), s["lam_ref"]) < 1e-6
    assert rel(v.grad.numpy(), g_ref) < 1e-6
    assert rel(v.grad.numpy(), g_swapped) > 1e-3  # the check sees a transposed outer product


def test_adjoint_gradient_matches_finite_differences(c4, ctx):
    s = c4
    val, b, xt = s["val"], s["b"], s["xt"]
    # a perturbed single entry makes A slightly nonsymmetric; the +-h solves run
    # to 1e-14 so that the solver's error (smooth in h, but not symmetric-exact)
    # stays far below the 1e-6 bar of the difference quotient
    p = bx.Params(solver="FCG", preconditioner="BCMG", rtol=1e-14, itnlim=5000, dispnorm=0)
    system = AG.SPDSystem(s["n"], s["row"], s["col"], val, p, ctx=ctx)
    v = torch.tensor(val, requires_grad=True)
    bt = torch.tensor(b, requires_grad=True)
    x = AG.sparse_solve(v, bt, system)
    (0.5 * ((x - torch.tensor(xt)) ** 2).sum()).backward()
    gv, gb = v.grad.numpy().copy(), bt.grad.numpy().copy()

    rng = np.random.default_rng(7)
    idx = rng.choice(len(val), size=200, replace=False)
    fd = np.zeros(len(idx))
    with torch.no_grad():
        for t, k in enumerate(idx):
            xs = []
            for sgn in (1.0, -1.0):
                vp = val.copy()
                vp[k] += sgn * H
                xs.append(AG.sparse_solve(torch.tensor(vp), torch.tensor(b), system).numpy())
            fd[t] = loss_difference(xs[0], xs[1], xt) / (2 * H)
        bidx = rng.choice(len(b), size=20, replace=False)
        fdb = np.zeros(len(bidx))
        for t, k in enumerate(bidx):
            xs = []
            for sgn in (1.0, -1.0):
                bp = b.copy()
                bp[k] += sgn * H
                xs.append(AG.sparse_solve(torch.tensor(val), torch.tensor(bp), system).numpy())
            fdb[t] = loss_difference(xs[0], xs[1], xt) / (2 * H)
    system.close()
    print(f"finite differences: values {rel(gv[idx], fd):.3e}, b {rel(gb[bidx], fdb):.3e}")
    assert rel(gv[idx], fd) < 1e-6, np.max(np.abs(gv[idx] - fd))
    assert rel(gb[bidx], fdb) < 1e-6


def test_adjoint_multiple_right_hand_sides(c4):
    """b with 3 columns: dL/dA_ij = -sum_c lambda_ic x_jc."""
    s = c4
    n = s["n"]
    B = np.random.default_rng(8).normal(size=(n, 3))
    Xt = np.random.default_rng(9).normal(size=(n, 3))
    S = sp.csr_matrix((s["val"], s["col"], s["row"]), shape=(n, n))
    lu = spla.splu(S.tocsc())
    Xr = lu.solve(B)
    Lr = lu.solve(Xr - Xt, trans="T")
    v = torch.tensor(s["val"], requires_grad=True)
    bt = torch.tensor(B, requires_grad=True)
    X = AG.sparse_solve(v, bt, s["system"])
    (0.5 * ((X - torch.tensor(Xt)) ** 2).sum()).backward()
    g_ref = -(Lr[s["rows"]] * Xr[s["col"]]).sum(axis=1)
    assert rel(X.detach().numpy(), Xr) < 1e-9
    assert rel(bt.grad.numpy(), Lr) < 1e-6
    assert rel(v.grad.numpy(), g_ref) < 1e-6


def test_new_values_with_rebuilt_hierarchy(ctx):
    n, row, col, val, xt, b = c4_problem(16)
    system = AG.SPDSystem(n, row, col, val, tight_params(), ctx=ctx, rebuild=True)
    try:
        # a different s.p.d. operator on the same pattern: scale the off-diagonals
        rows = np.repeat(np.arange(n), np.diff(row))
        val2 = np.where(rows == col, val, 0.7 * val)
        x = AG.sparse_solve(torch.tensor(val2), torch.tensor(b), system).numpy()
        S2 = sp.csr_matrix((val2, col, row), shape=(n, n))
        assert rel(x, spla.spsolve(S2.tocsc(), b)) < 1e-9
    finally:
        system.close()


def test_backward_uses_its_own_forward_values(c4):
    """Two forwards on one system with different values, then the first one's
    backward: its adjoint runs on the first forward's matrix, not the values
    the second forward left on the device."""
    s = c4
    v1 = torch.tensor(s["val"], requires_grad=True)
    rows = np.repeat(np.arange(s["n"]), np.diff(s["row"]))
    val2 = np.where(rows == s["col"], s["val"], 0.8 * s["val"])
    bt = torch.tensor(s["b"])
    x1 = AG.sparse_solve(v1, bt, s["system"])
    AG.sparse_solve(torch.tensor(val2), bt, s["system"])  # another matrix on the same system
    (0.5 * ((x1 - torch.tensor(s["xt"])) ** 2).sum()).backward()
    g_ref = -s["lam_ref"][s["rows"]] * s["x_ref"][s["col"]]
    print(f"backward after a second forward: dL/dA {rel(v1.grad.numpy(), g_ref):.3e}")
    assert rel(v1.grad.numpy(), g_ref) < 1e-6
