"""CPU checks of the MMA87 oracle (oracle/mma_oracle.py): the KKT conditions
of each convex subproblem, and convergence to the analytic optimum of a
separable convex problem.  The reference's MMA comes from the external
NonconvexMMA package, so its iterates themselves are parity unpinned."""
import numpy as np
import pytest

from oracle.mma_oracle import MMA87


def _kkt_check(opt, sp, x, y, lam):
    """y minimises the subproblem Lagrangian over [a, b] for lam, and the
    approximate constraint is active (lam > 0) or satisfied (lam = 0)."""
    P = sp["p0"] + lam * sp["p1"]
    Q = sp["q0"]
    d = P / (sp["U"] - y) ** 2 - Q / (y - sp["L"]) ** 2  # dLagrangian / dy
    scale = np.abs(P / (sp["U"] - y) ** 2) + np.abs(Q / (y - sp["L"]) ** 2) + 1e-300
    inner = (y > sp["a"] + 1e-12) & (y < sp["b"] - 1e-12)
    assert np.all(np.abs(d[inner]) <= 1e-9 * scale[inner])
    assert np.all(d[np.isclose(y, sp["a"], rtol=0, atol=1e-12)] >= -1e-9 * scale.max())
    assert np.all(d[np.isclose(y, sp["b"], rtol=0, atol=1e-12)] <= 1e-9 * scale.max())
    gt = np.sum(sp["p1"] / (sp["U"] - y)) - sp["T"]  # g~(y)
    ref = np.sum(sp["p1"] / (sp["U"] - x))
    if lam > 0:
        assert abs(gt) <= 1e-10 * ref
    else:
        assert gt <= 1e-12 * ref


def test_mma_converges_to_separable_optimum():
    """min sum c_j / x_j  s.t.  mean(x) <= V, 0 <= x <= 1: the optimum is
    x_j = min(1, sqrt(c_j / mu)) with mu fixed by the constraint."""
    g = np.random.default_rng(0)
    n, V = 200, 0.4
    c = g.uniform(0.1, 2.0, n)
    opt = MMA87()
    x = np.full(n, V)
    for _ in range(60):
        dc = -c / x ** 2
        dg = np.full(n, 1.0 / n)
        xn, lam, sp = opt.step(x, dc, dg, x.mean() - V)
        _kkt_check(opt, sp, x, xn, lam)
        x = xn
    lo, hi = 1e-12, 1e12
    for _ in range(200):
        mu = np.sqrt(lo * hi)
        if np.minimum(1.0, np.sqrt(c / mu)).mean() > V:
            lo = mu
        else:
            hi = mu
    xs = np.minimum(1.0, np.sqrt(c / mu))
    np.testing.assert_allclose(x, xs, rtol=1e-6, atol=1e-8)


def test_mma_inactive_constraint_and_asymptotes():
    """A constraint that stays slack gives lam = 0; the asymptotes follow the
    1987 rule (s_init for two iterations, then widen / narrow)."""
    n = 10
    opt = MMA87(s_init=0.5, s_incr=1.2, s_decr=0.7)
    x = np.full(n, 0.5)
    dc = np.full(n, 1.0)  # increasing objective: x moves down, volume stays below V
    y, lam, sp = opt.step(x, dc, np.full(n, 1.0 / n), x.mean() - 0.9)
    assert lam == 0.0
    np.testing.assert_allclose(sp["L"], x - 0.5)
    np.testing.assert_allclose(y, sp["a"])  # pushed to the lower move limit
    x1 = y
    y2, _, sp2 = opt.step(x1, dc, np.full(n, 1.0 / n), x1.mean() - 0.9)
    np.testing.assert_allclose(sp2["L"], x1 - 0.5)
    y3, _, sp3 = opt.step(y2, dc, np.full(n, 1.0 / n), y2.mean() - 0.9)
    # two steps in the same direction: the asymptote distance grows by s_incr
    np.testing.assert_allclose(y2 - sp3["L"], 1.2 * (x1 - sp2["L"]))


def test_oracle_simp_mma_keeps_volume(O):
    prob = O.Cantilever((30, 10))
    r = O.simp(prob, 0.5, 2.0, 30, solver="direct", optimizer="mma")
    flt = O.Filter((30, 10), 2.0)
    assert (flt.colsum @ r["x"]) / prob.nel <= 0.5 + 1e-6
    assert r["obj"][-1] < 0.6 * r["obj"][0]
    oc = O.simp(prob, 0.5, 2.0, 30, solver="direct")
    assert abs(r["obj"][-1] - oc["obj"][-1]) < 0.1 * oc["obj"][-1]
