"""CPU checks of the heat-conduction oracle (oracle/heat_oracle.py) and of the
host-side problem mirror (topopt_jl_amd.heat): closed forms, invariants, the
reference's own ThermalComplianceFun test facts
(test/Functions/test_thermal_compliance.jl) and central differences."""
import numpy as np
import pytest

from oracle import heat_oracle as H


def test_ke_closed_forms():
    # unit square, Ferrite corner order: k/6 [4 -1 -2 -1; ...]
    K2 = H.heat_ke(2, (1.0, 1.0), 2.0)
    ref = 2.0 / 6 * np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]])
    np.testing.assert_allclose(K2, ref, atol=1e-15)
    # unit cube: diagonal 1/3, edge neighbours 0, face and body diagonals -1/12
    K3 = H.heat_ke(3, (1.0, 1.0, 1.0))
    c = np.array(H.CORNERS[3])
    dist = np.abs(c[:, None, :] - c[None, :, :]).sum(-1)
    expect = np.select([dist == 0, dist == 1], [1 / 3, 0.0], -1 / 12)
    np.testing.assert_allclose(K3, expect, atol=1e-15)
    for dim, h in ((2, (2.0, 0.5)), (3, (1.0, 0.5, 2.0))):
        K = H.heat_ke(dim, h, 1.5)
        np.testing.assert_allclose(K, K.T, atol=1e-14)
        np.testing.assert_allclose(K.sum(1), 0.0, atol=1e-13)  # constants are in the null space
        np.testing.assert_allclose(K, H.heat_ke(dim, h, 1.5, order=4), atol=1e-13)
        assert np.all(np.linalg.eigvalsh(K)[1:] > 1e-12)  # one zero mode only


def test_numbering_matches_package(T):
    from topopt_jl_amd.problems import ferrite_numbering
    for nels in ((5, 3), (3, 2, 2)):
        P = H.HeatProblem(nels)
        nd, cd = ferrite_numbering(nels, ncomp=1)
        np.testing.assert_array_equal(P.node_dof, nd[0] - 1)
        np.testing.assert_array_equal(P.cell_dofs.T, cd - 1)


def test_load_and_dirichlet_sets_match_package(T):
    cases = [((8, 6), (1.0, 0.5), dict(Tleft=0.0, Tright=2.0, heatflux={"top": 100.0})),
             ((6, 4), None, dict(Tleft=None, Tright=None, Tbottom=0.0, Ttop=1.0,
                                 heatflux={"left": 2.0, "bottom": -1.0}, cload={12: 3.0},
                                 Tfix={5: 7.0})),
             ((4, 2, 3), (1.0, 2.0, 0.5), dict(heatflux={"top": 1.5, "front": 2.0},
                                               cload={7: 1.0}))]
    for nels, sizes, kw in cases:
        P = H.HeatProblem(nels, sizes, **kw)
        Q = T.HeatConductionProblem(nels, sizes, **kw)
        np.testing.assert_allclose(Q.fixedload, P.Q, atol=1e-12)
        np.testing.assert_array_equal(Q.fixed.astype(bool), P.fixed)
        np.testing.assert_allclose(Q.fixed_values, P.values)
        # total flux = q x face area (+ point sources)
        h = sizes or (1.0,) * len(nels)
        L = [n * hh for n, hh in zip(nels, h)]
        axis = {"left": 0, "right": 0, "front": 1, "back": 1, "bottom": 1 if len(nels) == 2 else 2,
                "top": 1 if len(nels) == 2 else 2}
        tot = sum(q * np.prod([L[d] for d in range(len(nels)) if d != axis[n]])
                  for n, q in kw.get("heatflux", {}).items()) + sum(kw.get("cload", {}).values())
        assert P.Q.sum() == pytest.approx(tot)


def test_linear_conduction_exact():
    """Uniform bar, T = 1 at x = 0 and 0 at x = L, no source: T is linear in x
    (bilinear / trilinear cells reproduce it exactly)."""
    for nels in ((10, 3), (6, 2, 3)):
        P = H.HeatProblem(nels, Tleft=1.0, Tright=0.0)
        T, _ = P.solve(np.ones(P.ncells))
        x = P.coords[:, 0][np.argsort(P.node_dof)]
        np.testing.assert_allclose(T, 1.0 - x / nels[0], atol=1e-12)


def test_reference_objective_facts():
    # J == Q'T > 0 (homogeneous), and T'KT > J + 1 with Tleft = 100
    P = H.HeatProblem((8, 4), Tleft=100.0, Tright=0.0, heatflux={"top": 1.0})
    J, g, cell, T, K = P.thermal_compliance(np.ones(32), p=3.0, xmin=0.01)
    assert J == pytest.approx(P.Q @ T, rel=1e-12)
    assert T @ (K @ T) > J + 1.0
    # no heat source: J = 0 and a zero gradient
    P0 = H.HeatProblem((4, 4), Tleft=100.0, Tright=0.0)
    J0, g0, *_ = P0.thermal_compliance(np.ones(16), p=3.0, xmin=0.01)
    assert abs(J0) < 1e-10 and np.linalg.norm(g0) < 1e-8


@pytest.mark.parametrize("Tleft,p", [(0.0, 1.0), (100.0, 3.0)])
def test_thermal_gradient_central_differences(Tleft, p):
    P = H.HeatProblem((8, 6), Tleft=Tleft, Tright=0.0, heatflux={"top": 100.0 if Tleft == 0 else 1.0})
    x = np.clip(np.random.default_rng(42).uniform(0, 1, P.ncells), 0.2, 1.0)
    J, g, *_ = P.thermal_compliance(x, p=p, xmin=0.01)
    eps = 1e-6
    fd = np.empty(P.ncells)
    for e in range(P.ncells):
        xp, xm = x.copy(), x.copy()
        xp[e] += eps
        xm[e] -= eps
        fd[e] = (P.thermal_compliance(xp, p=p, xmin=0.01)[0] -
                 P.thermal_compliance(xm, p=p, xmin=0.01)[0]) / (2 * eps)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8 * np.abs(fd).max())


def test_temperature_pullback_central_differences():
    P = H.HeatProblem((6, 4), Tleft=5.0, Tright=None, Tbottom=0.0, heatflux={"top": 2.0})
    g = np.random.default_rng(3)
    x = g.uniform(0.2, 1.0, P.ncells)
    D = g.standard_normal(P.nnodes)
    pb = P.temperature_pullback(x, D, p=3.0, xmin=0.01)
    eps = 1e-6
    fd = np.empty(P.ncells)
    for e in range(P.ncells):
        xp, xm = x.copy(), x.copy()
        xp[e] += eps
        xm[e] -= eps
        fd[e] = (D @ P.solve(xp, 3.0, 0.01)[0] - D @ P.solve(xm, 3.0, 0.01)[0]) / (2 * eps)
    np.testing.assert_allclose(pb, fd, rtol=1e-5, atol=1e-9 * np.abs(fd).max())
