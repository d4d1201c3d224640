"""GPU: the multigrid-preconditioned mixed-precision solve (TOHIP_MIXED_MG).

The V-cycle itself is compared with its FP64 oracle (oracle/mg_oracle.py,
pinned in tests/test_mg_oracle.py) at FP32 tolerance: relative error <= 1e-5
(FP32 vectors and operator; the coarsest level runs 30 sweeps).  The solves
are held to the same FP64 residual rule as every other path, and their
compliance to 1e-6 of the FP64 path (the north-star bound for mixed precision).
"""
import numpy as np
import pytest

from conftest import relerr, rng

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def MG(O):
    from oracle import mg_oracle
    return mg_oracle


def node_major(prob):
    dim = prob.dim
    idx = np.empty(prob.ndofs, np.int64)
    for d in range(dim):
        idx[d::dim] = prob.node_dofs[d] - 1
    return idx


def node_mask(prob):
    dim = prob.dim
    fx = prob.fixed.astype(bool)
    return np.array([sum(int(fx[prob.node_dofs[d, n] - 1]) << d for d in range(dim))
                     for n in range(prob.node_dofs.shape[1])], np.uint8)


# (16,8,8): one coarse level in LDS; (40,16,16): three levels; (42,14,14): the
# coarsest (21x7x7) is too large for the one-workgroup kernel (sweep launches);
# (32,16): 2D.
@pytest.mark.parametrize("nels", [(16, 8, 8), (40, 16, 16), (42, 14, 14), (32, 16)])
def test_vcycle_matches_oracle(T, O, MG, gpu, nels):
    prob = T.PointLoadCantilever(nels)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-3, penalty=T.PowerPenaltyFun(3.0))
    g = rng(21)
    x = g.uniform(0.0, 1.0, prob.ncells)
    s.vars = x
    r = g.standard_normal(prob.ndofs)
    r[prob.fixed.astype(bool)] = 0.0
    z = s.mg_apply(r)
    E = 1e-3 + (1 - 1e-3) * x ** 3
    md = O.Oracle(O.Cantilever(tuple(nels)), xmin=1e-3, p=3.0).meandiag
    M = MG.MGOracle(tuple(nels), E, node_mask(prob), md)
    # the device's smoother weights (from its lambda_max estimate) per level
    w, _ = s.mg_weights()
    assert len(w) == len(M.omega)
    M.omega = [float(v) for v in w]
    idx = node_major(prob)
    zo = M.apply(r[idx])
    err = relerr(z[idx], zo)
    print(f"{nels}: levels {M.grids} rel err {err:.2e}")
    assert err < 1e-5
    assert np.all(z[prob.fixed.astype(bool)] == 0.0)


# Box cells: every level takes the general element operator of its 2^l h box.
# (64, 32): three levels, so an intermediate level is smoothed as well.
@pytest.mark.parametrize("nels,sizes", [((64, 32), (1.0, 1.0)), ((64, 32), (1.0, 0.5)),
                                        ((32, 16), (1.0, 0.5)), ((40, 16, 16), (1.0, 2.0, 0.5))])
def test_vcycle_box_cells_matches_oracle(T, O, MG, gpu, nels, sizes):
    prob = T.PointLoadCantilever(nels, sizes=sizes)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-3, penalty=T.PowerPenaltyFun(3.0))
    g = rng(28)
    x = g.uniform(0.0, 1.0, prob.ncells)
    s.vars = x
    r = g.standard_normal(prob.ndofs)
    r[prob.fixed.astype(bool)] = 0.0
    z = s.mg_apply(r)
    E = 1e-3 + (1 - 1e-3) * x ** 3
    md = O.Oracle(O.Cantilever(tuple(nels)), xmin=1e-3, p=3.0, h=sizes).meandiag
    M = MG.MGOracle(tuple(nels), E, node_mask(prob), md, h=list(sizes))
    w, _ = s.mg_weights()
    M.omega = [float(v) for v in w]
    idx = node_major(prob)
    zo = M.apply(r[idx])
    err = relerr(z[idx], zo)
    print(f"{nels} cells {sizes}: levels {M.grids} rel err {err:.2e}")
    assert err < 1e-5


def test_vcycle_symmetric_on_device(T, gpu):
    prob = T.PointLoadCantilever((40, 16, 16))
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-3, penalty=T.PowerPenaltyFun(3.0))
    g = rng(22)
    s.vars = g.uniform(0.0, 1.0, prob.ncells)
    fx = prob.fixed.astype(bool)
    r1, r2 = g.standard_normal(prob.ndofs), g.standard_normal(prob.ndofs)
    r1[fx] = 0
    r2[fx] = 0
    z1 = s.mg_apply(r1)
    a, b = z1 @ r2, r1 @ s.mg_apply(r2, refresh=False)
    assert abs(a - b) <= 1e-5 * abs(a)


@pytest.mark.parametrize("nels", [(40, 16, 16), (64, 32)])
def test_mg_solve_matches_fp64(T, gpu, nels):
    prob = T.PointLoadCantilever(nels, load="centre") if len(nels) == 3 else \
        T.PointLoadCantilever(nels)
    g = rng(23)
    x = g.uniform(0.05, 1.0, prob.ncells)
    res = {}
    for prec in ("fp64", "mixed_mg"):
        s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0),
                        precision=prec, abstol=0.0, reltol=1e-8, cg_max_iter=100000)
        s.vars = x
        s()
        assert s.converged
        comp = T.ComplianceFun(s)
        res[prec] = (s.u.copy(), s.last_iters, comp(x, initially_zero=True))
        # true residual of the accepted solution, in FP64 through the device
        # matvec: the mixed paths stop on it; the FP64 path (the reference's
        # rule) stops on its recursive residual, which drifts from it
        rr = np.linalg.norm(prob.f - s.matvec(s.u)) / np.linalg.norm(prob.f)
        print(f"{prec}: true relative residual {rr:.2e}")
        assert rr <= (1.01e-8 if prec == "mixed_mg" else 1e-7)
    (u64, it64, c64), (umg, itmg, cmg) = res["fp64"], res["mixed_mg"]
    print(f"{nels}: Jacobi PCG {it64} iterations, MG PCG {itmg}")
    assert abs(cmg - c64) <= 1e-6 * abs(c64)
    print(f"relative difference of the displacements {relerr(umg, u64):.1e}")
    assert relerr(umg, u64) < 1e-3
    assert itmg * 5 < it64


def test_mg_rejects_odd_grid(T, gpu):
    prob = T.PointLoadCantilever((15, 6, 6))
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, precision="mixed_mg")
    s.vars = np.full(prob.ncells, 0.5)
    with pytest.raises(T.TohipError, match="even"):
        s()


def test_simp_loop_mixed_mg_matches_fp64(T, gpu):
    nels, rmin, vf, iters = (40, 16, 16), 1.5, 0.3, 8
    prob = T.PointLoadCantilever(nels, load="centre")
    out = {}
    for prec in ("fp64", "mixed_mg"):
        s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
        out[prec] = T.simp_run(s, volfrac=vf, rmin=rmin, iterations=iters, rel_tol=1e-8,
                               precision=prec)
    a, b = out["fp64"], out["mixed_mg"]
    print("PCG iterations fp64", list(a.pcg_iters), "mixed_mg", list(b.pcg_iters))
    assert b.converged
    np.testing.assert_allclose(b.obj, a.obj, rtol=1e-6)
    assert np.max(np.abs(b.x - a.x)) < 1e-5
    assert sum(b.pcg_iters) * 5 < sum(a.pcg_iters)


@pytest.mark.parametrize("nels", [(16, 8, 8), (40, 16, 16)])
def test_dense_coarse_solve_matches_iteration(T, gpu, nels, monkeypatch):
    """The coarsest level's dense operator (built once per operator refresh,
    symmetrised, applied one wave per row) against the one-workgroup Chebyshev
    iteration it stands for (TOHIP_MG_DENSE=0): the same V-cycle up to FP32
    rounding, and prescribed DOFs stay zero."""
    prob = T.PointLoadCantilever(nels)
    g = rng(23)
    x = g.uniform(0.0, 1.0, prob.ncells)
    fx = prob.fixed.astype(bool)
    r = g.standard_normal(prob.ndofs)
    r[fx] = 0.0

    def vcycle():
        s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-3, penalty=T.PowerPenaltyFun(3.0))
        s.vars = x
        return s.mg_apply(r)

    z_dense = vcycle()
    monkeypatch.setenv("TOHIP_MG_DENSE", "0")
    z_iter = vcycle()
    err = relerr(z_dense, z_iter)
    print(f"{nels}: dense vs iteration rel err {err:.2e}")
    assert err < 1e-5
    assert np.all(z_dense[fx] == 0.0)


@pytest.mark.parametrize("nels,design", [((40, 16, 16), "random"), ((48, 16, 16), "layered"),
                                         ((64, 32), "random")])
def test_weight_estimate(T, O, MG, gpu, nels, design):
    """Smoother weights from the device's lambda_max(D^-1 K) estimate, per
    smoothed level, against the exact largest eigenvalue (scipy eigsh on the
    oracle's assembled level operators): the estimate is a Rayleigh quotient,
    so at most lambda_max (FP32 rounding: 1e-4); the weight is
    clamp(1.9 / (1.2 theta), bound weight, 0.8) or the bound's weight, and in
    every case w * lambda_max < 2 (a convergent smoother).  A second refresh
    (2 more warm steps) keeps all of this."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    prob = T.PointLoadCantilever(nels)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-3, penalty=T.PowerPenaltyFun(3.0))
    g = rng(25)
    if design == "random":
        x = g.uniform(0.0, 1.0, prob.ncells)
    else:  # solid and void slabs along x
        x = np.where((np.arange(prob.ncells) // 4) % 2 == 0, 1.0, 0.05)
    E = 1e-3 + (1 - 1e-3) * x ** 3
    md = O.Oracle(O.Cantilever(tuple(nels)), xmin=1e-3, p=3.0).meandiag
    M = MG.MGOracle(tuple(nels), E, node_mask(prob), md)
    lam = []
    for l in range(len(M.K) - 1):
        Dh = sp.diags(np.sqrt(M.dinv[l]))
        lam.append(spla.eigsh(Dh @ M.K[l] @ Dh, k=1, which="LA", return_eigenvectors=False,
                              tol=1e-8)[0])
    r = g.standard_normal(prob.ndofs)
    r[prob.fixed.astype(bool)] = 0.0
    for refresh in range(2):
        s.vars = x
        s.mg_apply(r)
        w, th = s.mg_weights()
        for l, lm in enumerate(lam):
            wb = M.omega[l]  # the design-independent bound's weight (oracle default)
            print(f"{nels} {design} refresh {refresh} level {l}: lambda_max {lm:.4f} "
                  f"estimate {th[l]:.4f} ({th[l] / lm:.3f}) w {w[l]:.4f} w*lambda {w[l] * lm:.3f}")
            assert 0.5 * lm < th[l] <= lm * (1 + 1e-4)
            adapted = min(max(1.9 / (1.2 * th[l]), wb), 0.8)
            assert abs(w[l] - wb) <= 1e-6 * wb or abs(w[l] - adapted) <= 1e-5 * adapted
            assert w[l] >= wb * (1 - 1e-6)
            assert w[l] * lm < 2.0


def test_weight_estimate_after_design_change(T, O, MG, gpu):
    """The estimate is warm-started from the previous design's vector.  After
    a switch from a layered design (lambda_max(D^-1 K) ~ 3.1, weights adapted
    above the bound's) to a random one (~ 4.5), every refresh must keep
    w * lambda_max < 2 on every smoothed level (the stale vector's Rayleigh
    quotient has not levelled off, so the bound's weight is used until it
    has), and the weights adapt to the new design within a few refreshes."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    nels = (40, 16, 16)
    prob = T.PointLoadCantilever(nels)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-3, penalty=T.PowerPenaltyFun(3.0))
    md = O.Oracle(O.Cantilever(nels), xmin=1e-3, p=3.0).meandiag
    g = rng(26)
    r = g.standard_normal(prob.ndofs)
    r[prob.fixed.astype(bool)] = 0.0

    def lambdas(x):
        E = 1e-3 + (1 - 1e-3) * x ** 3
        M = MG.MGOracle(nels, E, node_mask(prob), md)
        out = []
        for l in range(len(M.K) - 1):
            Dh = sp.diags(np.sqrt(M.dinv[l]))
            out.append(spla.eigsh(Dh @ M.K[l] @ Dh, k=1, which="LA", return_eigenvectors=False,
                                  tol=1e-8)[0])
        return out, M.omega

    xa = np.where((np.arange(prob.ncells) // 4) % 2 == 0, 1.0, 0.05)
    xb = g.uniform(0.0, 1.0, prob.ncells)
    la, wb = lambdas(xa)
    lb, _ = lambdas(xb)
    for _ in range(3):
        s.vars = xa
        s.mg_apply(r)
    w, th = s.mg_weights()
    print(f"layered: lambda_max {np.round(la, 3)} weights {np.round(w, 4)}")
    assert w[0] > 1.2 * wb[0]  # adapted above the bound's weight
    for refresh in range(6):
        s.vars = xb
        s.mg_apply(r)
        w, th = s.mg_weights()
        wl = [w[l] * lb[l] for l in range(len(lb))]
        print(f"random, refresh {refresh}: estimate {np.round(th[:len(lb)], 3)} of "
              f"{np.round(lb, 3)}, weights {np.round(w[:len(lb)], 4)}, w*lambda {np.round(wl, 3)}")
        assert max(wl) < 2.0
    assert w[0] > wb[0] * (1 + 1e-6)  # adapted to the random design by now


@pytest.mark.parametrize("nels,sizes", [((40, 16, 16), (1.0, 2.0, 0.5)), ((64, 32), (1.0, 0.5))])
def test_mg_box_cells(T, gpu, nels, sizes):
    """Non-cubic cells take the general element operator in every multigrid
    kernel (smoothing, the weight estimate's power steps, the coarsest
    operator's build).  The oracle's multigrid restatement is for cubes, so
    this checks what holds for any cells: the V-cycle is symmetric, the
    weights stay at or above the bound's, and the multigrid solve reaches the
    FP64 Jacobi solution (compliance 1e-6) in far fewer iterations."""
    prob = T.PointLoadCantilever(nels, sizes=sizes)
    g = rng(27)
    x = g.uniform(0.05, 1.0, prob.ncells)
    fx = prob.fixed.astype(bool)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
    s.vars = x
    r1, r2 = g.standard_normal(prob.ndofs), g.standard_normal(prob.ndofs)
    r1[fx] = 0.0
    r2[fx] = 0.0
    z1 = s.mg_apply(r1)
    z2 = s.mg_apply(r2, refresh=False)
    a, b = z1 @ r2, r1 @ z2
    w, th = s.mg_weights()
    print(f"{nels} cells {sizes}: symmetry {a:.6e} vs {b:.6e}; weights {np.round(w, 4)}, "
          f"estimates {np.round(th, 3)}")
    assert abs(a - b) <= 1e-5 * abs(a)
    assert r1 @ z1 > 0 and np.all(z1[fx] == 0.0)
    res = {}
    for prec in ("fp64", "mixed_mg"):
        sp_ = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0),
                          precision=prec, abstol=0.0, reltol=1e-8, cg_max_iter=100000)
        sp_.vars = x
        sp_()
        assert sp_.converged
        res[prec] = (sp_.last_iters, T.ComplianceFun(sp_)(x, initially_zero=True))
    (it64, c64), (itmg, cmg) = res["fp64"], res["mixed_mg"]
    print(f"Jacobi PCG {it64} iterations, MG PCG {itmg}; compliance {c64:.10g} / {cmg:.10g}")
    assert abs(cmg - c64) <= 1e-6 * abs(c64)
    assert itmg * 5 < it64
