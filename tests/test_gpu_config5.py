"""GPU parity at the large cantilever: 256 x 128 x 128 hex8 (4,194,304 cells,
12,830,211 DOFs), PointLoadCantilever (src/TopOptProblems/problem_types.jl
:162-231) with the load at the x = max face centre, p = 3, rmin = 2.

* FP64 matvec (free DOFs; then the prescribed-DOF rule), compliance and
  sensitivities against the oracle at the same size;
* the density filter (rmin 2) and its pullback against the oracle's closed
  form of the filter (the sparse J would hold ~4e8 entries);
* three SIMP iterations of the mixed-precision multigrid path, without and
  with the learned warm start (the network fine-tuned on 200 pairs at this
  size), against the FP64 Jacobi path (compliance within 1e-6, designs
  within 1e-5), and the device memory the solver holds (its high-water mark:
  buffers are only released at destroy).
Each case prints its sizes and timings (pytest -s)."""
import ctypes as C
import time

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

C5 = (256, 128, 128)


@pytest.fixture(scope="module")
def c5(T, O, gpu):
    t0 = time.perf_counter()
    prob = T.PointLoadCantilever(C5, load="centre")
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
    orc = O.Oracle(O.Cantilever(C5, load="centre"), xmin=1e-9, p=3.0)
    g = np.random.default_rng(0)
    rho = g.uniform(0.01, 1.0, prob.ncells)
    u = g.standard_normal(prob.ndofs)
    u[prob.fixed.astype(bool)] = 0.0
    s.set_densities(rho)
    orc.set_vars(rho)
    print(f"\n[c5] {prob.ncells} cells, {prob.ndofs} DOFs, set-up {time.perf_counter() - t0:.1f} s",
          flush=True)
    yield prob, s, orc, rho, u
    s.close()


def test_c5_matvec(c5):
    prob, s, orc, rho, u = c5
    y = s.matvec(u)
    yo = orc.matvec(u)
    assert relerr(y, yo) < 1e-13
    fx = prob.fixed.astype(bool)
    assert np.all(y[fx] == 0.0)
    w = np.random.default_rng(2).standard_normal(prob.ndofs)
    y2 = s.matvec(w)
    yo2 = orc.matvec(w)
    assert relerr(y2[~fx], yo2[~fx]) < 1e-13
    np.testing.assert_allclose(y2[fx], orc.meandiag * w[fx], rtol=1e-15)
    # symmetry on the free space
    Kw = y2.copy()
    Kw[fx] = 0.0
    uu = u.copy()
    assert abs(Kw @ uu - w[~fx] @ y[~fx]) <= 1e-12 * np.linalg.norm(Kw) * np.linalg.norm(uu)


def test_c5_compliance_and_sensitivity(c5):
    from topopt_jl_amd._lib import dptr
    prob, s, orc, rho, u = c5
    obj = C.c_double()
    ce = np.empty(prob.ncells)
    gr = np.empty(prob.ncells)
    assert s.lib.tohip_compliance(s.ctx.ptr, dptr(u), C.byref(obj), dptr(ce), dptr(gr)) == 0
    obj_o, ce_o, g_o = orc.element_energy(u)
    assert abs(obj.value - obj_o) <= 1e-12 * abs(obj_o)
    assert relerr(ce, ce_o) < 1e-13
    assert relerr(gr, g_o) < 1e-13


def test_c5_filter_and_pullback(T, O, c5):
    prob, s, orc, rho, u = c5
    flt = T.DensityFilterFun(s, 2.0)
    y = flt(T.PseudoDensities(rho)).x
    assert relerr(y, O.filter_closed_form(C5, 2.0, rho)) < 1e-14
    d = np.random.default_rng(3).standard_normal(prob.ncells)
    z = flt.pullback(d)
    assert relerr(z, O.filter_closed_form(C5, 2.0, d, transpose=True)) < 1e-14
    # adjointness <J x, d> = <x, J^T d> and row sums 1, at the full size
    assert abs(y @ d - rho @ z) <= 1e-12 * np.linalg.norm(y) * np.linalg.norm(d)
    np.testing.assert_allclose(flt(np.full(prob.ncells, 0.2)).x, 0.2, rtol=1e-14)


def test_c5_mixed_mg_simp_matches_fp64(T, gpu):
    import os
    from topopt_jl_amd.nn import WarmStartNet
    weights = os.path.join(os.path.dirname(T.__file__), "weights", "warmstart_256x128x128.npz")
    res = {}
    for name in ("fp64", "mixed_mg", "learned_mg"):
        prec = "fp64" if name == "fp64" else "mixed_mg"
        prob = T.PointLoadCantilever(C5, load="centre")
        s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0),
                        precision=prec)
        warm = True
        if name == "learned_mg":  # the network fine-tuned on 200 pairs at this size
            net = WarmStartNet.for_problem(prob, weights)
            s.attach_warm_start(net)
            warm = "learned"
        t0 = time.perf_counter()
        r = T.simp_run(s, volfrac=0.2, rmin=2.0, iterations=3, rel_tol=1e-8, precision=prec,
                       warm_start=warm)
        dt = time.perf_counter() - t0
        assert r.converged
        xp = np.empty(prob.ncells)
        from topopt_jl_amd._lib import dptr
        assert s.lib.tohip_get_xphys(s.ctx.ptr, dptr(xp)) == 0
        mem = s.memory()
        print(f"\n[c5] {name}: 3 SIMP iterations {dt:.2f} s, PCG iterations {list(r.pcg_iters)}, "
              f"obj {list(r.obj)}, network predictions {r.nn_calls}, solver device bytes "
              f"{mem['solver_bytes']}, device used {mem['device_used_bytes']}", flush=True)
        # the solver's device memory (all its buffers, kept until destroy),
        # with the attached network's weights and activations on the learned path
        # (measured on MI355X: 1.64 GB FP64, 1.83 GB mixed_mg, 5.05 GB learned_mg)
        assert mem["solver_bytes"] < (8 if name == "learned_mg" else 4) * 2 ** 30
        res[name] = (r.obj, r.x, xp)
        s.close()
        if name == "learned_mg":
            assert 1 <= r.nn_calls <= 3
            net.close()
    for name in ("mixed_mg", "learned_mg"):
        np.testing.assert_allclose(res[name][0], res["fp64"][0], rtol=1e-6)
        assert np.max(np.abs(res[name][1] - res["fp64"][1])) < 1e-5
        assert np.max(np.abs(res[name][2] - res["fp64"][2])) < 1e-5


def test_c5_vcycle_symmetric(c5):
    """The multigrid V-cycle at this size smooths its finest level by z-marching
    columns (pre-smoothing, prolongation pass, post-smoothing): it must stay a
    symmetric operator, (M r1).r2 = r1.(M r2), or PCG loses its footing."""
    prob, s, orc, rho, u = c5
    g = np.random.default_rng(4)
    fx = prob.fixed.astype(bool)
    r1, r2 = g.standard_normal(prob.ndofs), g.standard_normal(prob.ndofs)
    r1[fx] = 0.0
    r2[fx] = 0.0
    z1, z2 = s.mg_apply(r1), s.mg_apply(r2, refresh=False)
    a, b = z1 @ r2, r1 @ z2
    print(f"\n[c5] V-cycle symmetry: {a:.6e} vs {b:.6e}", flush=True)
    assert abs(a - b) <= 1e-5 * abs(a)
    assert np.all(z1[fx] == 0.0)
    assert r1 @ z1 > 0 and r2 @ z2 > 0


def test_c5_fused_prolongation_same_bits(c5, monkeypatch):
    """The finest post-smoothing forms x' = x + P xc while staging its
    columns (coarse planes through an LDS ring; z-chunks of 13 planes start
    on both parities here): the V-cycle must give the same bits as with the
    separate prolongation pass (TOHIP_ZM_PROL=0)."""
    prob, s, orc, rho, u = c5
    g = np.random.default_rng(5)
    fx = prob.fixed.astype(bool)
    r = g.standard_normal(prob.ndofs)
    r[fx] = 0.0
    s.mg_apply(r)  # densities as the fixture set them, hierarchy built
    z_fused = s.mg_apply(r, refresh=False)
    monkeypatch.setenv("TOHIP_ZM_PROL", "0")
    z_sep = s.mg_apply(r, refresh=False)
    print(f"\n[c5] fused vs separate prolongation: max |diff| "
          f"{np.abs(z_fused - z_sep).max():.3e}", flush=True)
    assert np.array_equal(z_fused, z_sep)
