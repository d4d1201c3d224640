"""GPU: BESO (src/Algorithms/beso.jl) run on the device by tohip_beso_run,
against the oracle's restatement of the reference loop (oracle.beso).

The device and the oracle take the same decisions: the topology is compared
bit for bit, the objective to 1e-9 relative (solves to 1e-12), the iteration
count and the convergence flag exactly.  The reference's own BESO tests
(test/Algorithms/test_beso.jl, test/integration/test_end_to_end.jl:161-183)
supply the invariant checks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make(T, nels, xmin=1e-3):
    prob = T.PointLoadCantilever(nels)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=xmin)
    return prob, s, T.ComplianceFun(s), T.VolumeFun(s)


@pytest.mark.parametrize("nels,V,rmin,p,maxiter,tol,solver",
                         [((20, 10), 0.5, 2.0, 3.0, 40, 1e-3, "direct"),
                          ((12, 6), 0.5, 2.0, 1.0, 15, 0.05, "direct"),
                          ((12, 6, 6), 0.3, 1.5, 3.0, 25, 1e-3, "pcg")])
def test_beso_matches_oracle(T, O, gpu, nels, V, rmin, p, maxiter, tol, solver):
    prob, s, comp, vol = make(T, nels)
    beso = T.BESO(comp, vol, V, T.DensityFilterFun(s, rmin), maxiter=maxiter, tol=tol, p=p,
                  rel_tol=1e-12)
    x0 = np.full(prob.ncells, 0.6)
    res = beso(x0)
    ref = O.beso(O.Cantilever(tuple(nels)), V, rmin, maxiter=maxiter, tol=tol, p=p, x0=x0,
                 solver=solver)
    print(f"{nels}: fevals {res.fevals} converged {res.converged} change {res.change:.3e} "
          f"volume {res.volume:.4f} obj {res.objval:.6g}")
    assert res.fevals == ref["fevals"] and res.converged == bool(ref["converged"])
    assert np.array_equal(res.topology, ref["topology"])
    np.testing.assert_allclose(res.obj_trace, ref["obj"], rtol=1e-9)
    assert abs(res.objval - ref["objval"]) <= 1e-9 * abs(ref["objval"])
    assert abs(res.change - ref["change"]) <= 1e-9
    assert abs(res.volume - ref["volume"]) < 1e-15
    assert np.all((res.topology == 0) | (res.topology == 1))


def test_beso_reference_cases(T, O, gpu):
    """test_beso.jl: construction fields; passive (black / white) cells keep their
    values; length and overlap errors."""
    prob, s, comp, vol = make(T, (12, 6))
    filt = T.DensityFilterFun(s, 2.0)
    beso = T.BESO(comp, vol, 0.4, filt, maxiter=8, tol=0.05, p=1.0, er=0.05)
    assert beso.vol_limit == 0.4 and beso.maxiter == 8 and beso.tol == 0.05
    assert beso.penalty.p == 1.0 and beso.er == 0.05 and len(beso.topology) == prob.ncells
    black = np.zeros(prob.ncells, bool)
    black[:6] = True
    white = np.zeros(prob.ncells, bool)
    white[-6:] = True
    res = beso(np.full(prob.ncells, 0.8), black=black, white=white)
    ref = O.beso(O.Cantilever((12, 6)), 0.4, 2.0, maxiter=8, tol=0.05, p=1.0, er=0.05,
                 x0=np.full(prob.ncells, 0.8), black=black, white=white)
    assert np.all(res.topology[black] == 1) and np.all(res.topology[white] == 0)
    assert np.array_equal(res.topology, ref["topology"]) and res.fevals == ref["fevals"]
    with pytest.raises(ValueError):
        beso(np.full(prob.ncells, 0.8), black=np.ones(3, bool))
    with pytest.raises(ValueError):
        beso(np.full(prob.ncells, 0.8), black=black, white=black)


def test_beso_multigrid_solves(T, gpu):
    """The same loop on the multigrid-preconditioned mixed solver: the design
    follows the FP64 run (solves to the same FP64 residual)."""
    out = {}
    for prec in ("fp64", "mixed_mg"):
        prob, s, comp, vol = make(T, (40, 16, 16), xmin=1e-3)
        beso = T.BESO(comp, vol, 0.3, T.DensityFilterFun(s, 1.5), maxiter=12, tol=1e-3, p=3.0,
                      rel_tol=1e-10, precision=prec)
        out[prec] = beso(np.ones(prob.ncells))
    a, b = out["fp64"], out["mixed_mg"]
    print(f"fevals {a.fevals} / {b.fevals}, differing cells "
          f"{int(np.sum(a.topology != b.topology))} of {a.topology.size}")
    assert a.fevals == b.fevals
    np.testing.assert_allclose(b.obj_trace, a.obj_trace, rtol=1e-6)
    assert np.mean(a.topology != b.topology) < 1e-3
