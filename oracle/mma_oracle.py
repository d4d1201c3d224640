"""TEST INFRASTRUCTURE ONLY — CPU oracle of the MMA87 design update.

numpy restatement of Svanberg's method of moving asymptotes (1987) for one
inequality constraint with gradient dg >= 0 (the volume constraint):
asymptotes, the separable convex approximation, the box [0.9 L + 0.1 x,
0.9 U + 0.1 x] within [0, 1], and the dual solved for the multiplier with a
scalar root finder (scipy brentq on F(lam) = T), independently of the
device's bracketing passes (both stop at lam = 1e15 when the constraint
cannot be met inside the move limits).  The reference's optimizer comes from the
external NonconvexMMA package (not vendored, no Manifest), so its exact
iterates are parity unpinned; this restatement is pinned by the KKT
conditions of every subproblem and by convergence to the analytic optimum of
a separable convex problem (tests/test_mma_oracle.py).
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import brentq


class MMA87:
    def __init__(self, s_init=0.5, s_incr=1.2, s_decr=0.7):
        self.s_init, self.s_incr, self.s_decr = s_init, s_incr, s_decr
        self.k = 0
        self.L = self.U = self.x1 = self.x2 = None

    def subproblem(self, x, dc, dg, g0):
        """Asymptotes and the subproblem data at iterate x (updates L, U)."""
        if self.k < 2:
            L, U = x - self.s_init, x + self.s_init
        else:
            osc = (x - self.x1) * (self.x1 - self.x2)
            s = np.where(osc < 0, self.s_decr, np.where(osc > 0, self.s_incr, 1.0))
            L = x - s * (self.x1 - self.L)
            U = x + s * (self.U - self.x1)
        self.L, self.U = L, U
        a = np.maximum(0.0, 0.9 * L + 0.1 * x)
        b = np.minimum(1.0, 0.9 * U + 0.1 * x)
        p0 = (U - x) ** 2 * np.maximum(dc, 0.0)
        q0 = (x - L) ** 2 * np.maximum(-dc, 0.0)
        p1 = (U - x) ** 2 * dg
        T = np.sum(p1 / (U - x)) - g0
        return dict(L=L, U=U, a=a, b=b, p0=p0, q0=q0, p1=p1, T=T)

    @staticmethod
    def y_of(sp, lam, x):
        P = sp["p0"] + lam * sp["p1"]
        rp, rq = np.sqrt(P), np.sqrt(sp["q0"])
        den = rp + rq
        y = np.where(den > 0, (sp["L"] * rp + sp["U"] * rq) / np.where(den > 0, den, 1.0), x)
        return np.clip(y, sp["a"], sp["b"])

    def step(self, x, dc, dg, g0):
        """One MMA87 iteration: returns (new design, multiplier, subproblem)."""
        x = np.asarray(x, float)
        sp = self.subproblem(x, dc, dg, g0)
        F = lambda lam: np.sum(sp["p1"] / (sp["U"] - self.y_of(sp, lam, x)))
        if F(0.0) <= sp["T"]:
            lam = 0.0
        else:
            hi = 1.0
            while F(hi) > sp["T"] and hi < 1e15:
                hi *= 10.0
            if F(hi) > sp["T"]:  # infeasible within the move limits: the bracket's end
                lam = hi
            else:
                lam = brentq(lambda l: F(l) - sp["T"], 0.0, hi, xtol=1e-300, rtol=1e-15,
                             maxiter=500)
        y = self.y_of(sp, lam, x)
        self.x2 = self.x1 if self.x1 is not None else x.copy()
        self.x1 = x.copy()
        self.k += 1
        return y, lam, sp
