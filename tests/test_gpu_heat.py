"""GPU parity of the heat-conduction path (tohip_heat_* through the C ABI)
against the oracle (oracle/heat_oracle.py): element matrix, operator,
forward solve with prescribed temperatures, ThermalComplianceFun and its
adjoint gradient, TemperatureFun's pullback, and the reference test facts
of test/Functions/test_thermal_compliance.jl."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

CASES = [
    ((8, 6), (1.0, 1.0), 1.0, dict(Tleft=0.0, Tright=0.0, heatflux={"top": 100.0})),
    ((12, 8), (1.0, 0.5), 2.0, dict(Tleft=100.0, Tright=0.0, heatflux={"top": 1.0},
                                    cload={20: 3.0})),
    ((6, 4, 4), (1.0, 1.0, 1.0), 1.0, dict(Tleft=None, Tright=None, Tbottom=0.0,
                                           heatflux={"top": 1.0}, Tfix={3: 2.0})),
    ((5, 4, 6), (0.5, 1.0, 2.0), 1.0, dict(Tleft=1.0, Tright=0.0, heatflux={"back": 2.0})),
]


def make(T, nels, sizes, k, kw, p=3.0, xmin=0.01):
    from oracle import heat_oracle as H
    prob = T.HeatConductionProblem(nels, sizes, k, **kw)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=xmin, penalty=T.PowerPenaltyFun(p),
                    abstol=0.0, reltol=1e-13, cg_max_iter=200000)
    return prob, s, H.HeatProblem(nels, sizes, k, **kw)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_heat_operator_matches_oracle(T, gpu, case):
    nels, sizes, k, kw = CASES[case]
    prob, s, P = make(T, nels, sizes, k, kw)
    np.testing.assert_allclose(s.element_matrix(), P.Ke, rtol=0, atol=1e-14 * np.abs(P.Ke).max())
    g = np.random.default_rng(case)
    x = g.uniform(0.05, 1.0, prob.ncells)
    s.set_densities(x)
    v = g.standard_normal(prob.ndofs)
    y = s.matvec(v)
    rho = 0.01 + 0.99 * x ** 3
    K = P.assemble(rho)
    free = ~P.fixed
    ye = np.zeros(prob.ndofs)
    ye[free] = K[free][:, free] @ v[free]
    ye[P.fixed] = P.ncells * np.trace(P.Ke) * v[P.fixed]  # fixed rows: meandiag * x
    assert relerr(y, ye) < 1e-14
    # the free rows on their own (the fixed rows' meandiag * x dominates the norm)
    assert relerr(y[free], ye[free]) < 1e-14


@pytest.mark.parametrize("case", range(len(CASES)))
def test_heat_solve_and_thermal_compliance(T, gpu, case):
    nels, sizes, k, kw = CASES[case]
    prob, s, P = make(T, nels, sizes, k, kw)
    x = np.random.default_rng(10 + case).uniform(0.2, 1.0, prob.ncells)
    comp = T.ThermalComplianceFun(s)
    J = comp(T.PseudoDensities(x))
    Je, ge, cell_e, Te, _ = P.thermal_compliance(x, p=3.0, xmin=0.01)
    assert s.converged
    assert relerr(s.u, Te) < 1e-9
    np.testing.assert_allclose(s.u[P.fixed], P.values[P.fixed], rtol=1e-12, atol=1e-12)
    assert J == pytest.approx(Je, rel=1e-9, abs=1e-12)
    assert relerr(comp.grad, ge) < 1e-8
    assert relerr(comp.cell_comp, cell_e) < 1e-8
    # TemperatureFun and its pullback
    tf = T.TemperatureFun(s)
    Tv = tf(T.PseudoDensities(x))
    assert relerr(Tv, Te) < 1e-9
    D = np.random.default_rng(20 + case).standard_normal(prob.ndofs)
    assert relerr(tf.pullback(D), P.temperature_pullback(x, D, p=3.0, xmin=0.01)) < 1e-8


def test_heat_reference_objective_facts(T, gpu):
    """test/Functions/test_thermal_compliance.jl: J == Q'T, T'KT > J + 1 with
    Tleft = 100, and J == 0 (zero gradient) without a heat source."""
    prob, s, P = make(T, (8, 4), (1.0, 1.0), 1.0, dict(Tleft=100.0, Tright=0.0,
                                                       heatflux={"top": 1.0}))
    comp = T.ThermalComplianceFun(s)
    J = comp(T.PseudoDensities(np.ones(32)))
    assert J == pytest.approx(prob.fixedload @ s.u, rel=1e-10)
    _, K = P.solve(np.ones(32), 3.0, 0.01)
    assert s.u @ (K @ s.u) > J + 1.0
    prob0, s0, _ = make(T, (4, 4), (1.0, 1.0), 1.0, dict(Tleft=100.0, Tright=0.0))
    comp0 = T.ThermalComplianceFun(s0)
    J0 = comp0(T.PseudoDensities(np.ones(16)))
    assert abs(J0) < 1e-10 and np.linalg.norm(comp0.pullback()) < 1e-8


def test_heat_tree_and_errors(T, gpu):
    prob = T.HeatTree((16, 8), q=1.0)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=0.001, penalty=T.PowerPenaltyFun(3.0),
                    abstol=0.0, reltol=1e-10, cg_max_iter=100000)
    comp = T.ThermalComplianceFun(s)
    J = comp(np.full(prob.ncells, 0.5))
    assert J > 0 and np.all(comp.grad < 0)  # more material conducts better everywhere
    assert np.all(s.u >= -1e-12)  # heat flows in at the top, out at the cold bottom
    with pytest.raises(ValueError):
        T.FEASolver(T.HIPMatrixFreeSolver, prob, precision="mixed")
    with pytest.raises(ValueError):
        T.ThermalComplianceFun(T.FEASolver(T.HIPMatrixFreeSolver, T.PointLoadCantilever((4, 2))))
