"""Dense restatement of the BCMG cycle (test helper, no GPU).

GAMG_cycle (GAMG_cycle.cu:58-236) with l1-Jacobi smoothing (l1jacobi.cu:344-479)
written as dense linear algebra from its definition: smoothing
x <- x + M^{-1} (f - A x) with M_i = sum_j |a_ij| (relaxPrepare,
matchingAggregation.cu:40-60), restriction P^T r, coarse correction
x <- x + P x_c, the coarsest level smoothed relaxnumber_coarse times, and the
number of coarse visits of the W-, H- and variable-V cycles (AMG.cu:306-331).
"""
import numpy as np

from bootcmatchgx_amd import abi


class DenseHierarchy:
    """Level operators A_l (dense), prolongators P_l (dense, one entry per row)
    and l1 divisors M_l = row sums of |A_l|.  Coarse operators are the dense
    Galerkin products P^T A P unless given."""

    def __init__(self, A0, prolongators, coarse=None):
        self.A = [np.asarray(A0, np.float64)]
        self.P = []
        for l, (agg, pv, nc) in enumerate(prolongators):
            Pm = np.zeros((len(agg), nc))
            Pm[np.arange(len(agg)), agg] = pv
            self.P.append(Pm)
            self.A.append(np.asarray(coarse[l], np.float64) if coarse is not None else Pm.T @ self.A[-1] @ Pm)
        self.M = [np.abs(a).sum(axis=1) for a in self.A]

    @classmethod
    def from_oracle(cls, A, P, h=0):
        """The oracle's prolongators (of bootstrap hierarchy h); everything else recomputed here."""
        pro = []
        for l in range(P.hrc_levels(h) - 1):
            agg, pv = P.hrc_level_P(h, l)
            pro.append((agg, pv, P.hrc_level_A(h, l + 1).dims()[0]))
        return cls(A.to_scipy().toarray(), pro)

    def levels(self):
        return len(self.A)


def np_cycle(H, p, l, f, x0=None):
    """One cycle on level l of H with settings p, right-hand side f, guess x0 (zero when None)."""
    nl = H.levels()
    A, M = H.A[l], H.M[l]
    x = np.zeros_like(f) if x0 is None else x0.copy()

    Md = M if f.ndim == 1 else M[:, None]  # f may hold several right-hand sides as columns

    def smooth(x, sweeps):
        for _ in range(sweeps):
            x = x + (f - A @ x) / Md
        return x

    if l == nl - 1:
        return smooth(x, p.relaxnumber_coarse)
    coeff = 2 ** l if p.cycle_type == abi.CYCLE["VARIABLE_V_CYCLE"] else 1
    x = smooth(x, p.prerelax_sweeps * coeff)
    fc = H.P[l].T @ (f - A @ x)
    if nl > 2 or p.relaxnumber_coarse > 0:
        visits = 1
        if p.cycle_type == abi.CYCLE["W_CYCLE"] and l < nl - 2:
            visits = 2
        if p.cycle_type == abi.CYCLE["H_CYCLE"] and l % 2 == 0:
            visits = 2
        if l + 1 == nl - 1:
            visits = 1
        xc = None
        for _ in range(visits):
            xc = np_cycle(H, p, l + 1, fc, xc)
        x = x + H.P[l] @ xc
    return smooth(x, p.postrelax_sweeps * coeff)
