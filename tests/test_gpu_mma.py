"""GPU parity of the MMA87 design update (tohip_simp_run_mma) against the
oracle's SIMP loop with its MMA87 restatement (oracle/mma_oracle.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nels,rmin,vf,iters,solver", [((30, 10), 2.0, 0.5, 12, "direct"),
                                                      ((12, 6, 6), 1.5, 0.3, 8, "pcg")])
def test_simp_mma_matches_oracle(T, O, gpu, nels, rmin, vf, iters, solver):
    prob = T.PointLoadCantilever(nels)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
    res = T.simp_run(s, volfrac=vf, rmin=rmin, iterations=iters, rel_tol=1e-12,
                     optimizer=T.MMA87())
    ref = O.simp(O.Cantilever(nels), vf, rmin, iters, p=3.0, xmin=1e-9, solver=solver,
                 reltol=1e-13, optimizer="mma")
    assert res.converged
    np.testing.assert_allclose(res.obj, ref["obj"], rtol=1e-7)
    assert np.max(np.abs(res.x - ref["x"])) < 1e-5


def test_simp_mma_design_quality(T, O, gpu):
    """40 MMA iterations: the filtered volume meets volfrac and the compliance
    lands within 10% of the OC run's (the two optimizers' different paths)."""
    nels, rmin, vf = (40, 20), 2.0, 0.4
    prob = T.PointLoadCantilever(nels)
    out = {}
    for name, opt in (("mma", T.MMA87()), ("oc", "oc")):
        s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
        out[name] = T.simp_run(s, volfrac=vf, rmin=rmin, iterations=40, rel_tol=1e-10,
                               optimizer=opt)
    flt = O.Filter(nels, rmin)
    assert flt.colsum @ out["mma"].x / prob.ncells <= vf + 1e-6
    assert out["mma"].obj[-1] < 0.5 * out["mma"].obj[0]
    assert abs(out["mma"].obj[-1] - out["oc"].obj[-1]) < 0.1 * out["oc"].obj[-1]
    with pytest.raises(ValueError):
        T.simp_run(s, volfrac=vf, rmin=rmin, iterations=1, optimizer="gcmma")


def test_simp_mma_resident_continuation_matches_one_run(T, gpu):
    """A run split into resident calls (x0="device" continues on the device)
    keeps the MMA asymptote schedule and design history: 5 + 5 iterations
    give the iterates of one 10-iteration call, bit for bit (same kernels,
    same order)."""
    nels, rmin, vf = (24, 12), 2.0, 0.4
    prob = T.PointLoadCantilever(nels)
    s1 = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
    one = T.simp_run(s1, volfrac=vf, rmin=rmin, iterations=10, rel_tol=1e-10,
                     optimizer=T.MMA87())
    s2 = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
    a = T.simp_run(s2, volfrac=vf, rmin=rmin, iterations=5, rel_tol=1e-10, optimizer=T.MMA87())
    b = T.simp_run(s2, volfrac=vf, rmin=rmin, iterations=5, rel_tol=1e-10, optimizer=T.MMA87(),
                   x0="device")
    assert np.array_equal(np.concatenate([a.obj, b.obj]), one.obj)
    # a new run on the same context (x0 given) restarts the schedule
    c = T.simp_run(s2, volfrac=vf, rmin=rmin, iterations=10, rel_tol=1e-10, optimizer=T.MMA87())
    assert np.array_equal(c.obj, one.obj) and np.array_equal(c.x, one.x)
