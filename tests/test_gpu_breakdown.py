"""PCG breakdown on the device: a hard-void design (xmin = 0, every cell
void: test/FEA/test_simulate.jl:17-35 expects an all-NaN displacement) makes
K(vars) zero on the free DOFs, so the Jacobi diagonal is 0 and r'z is not
finite at the first iteration.  The device flags it in that iteration's
direction kernel; the host returns TOHIP_BREAKDOWN with u = NaN (the
reference's convention for a matrix that is not positive definite,
src/FEA/solvers_api.jl:152-158) instead of replaying graphs to maxiter.
A numeric case, not a fault: every kernel exits normally."""
import time
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAXITER = 200000


def solver(T, nels, xmin):
    prob = T.PointLoadCantilever(nels)
    return prob, T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=xmin,
                             penalty=T.PowerPenaltyFun(3.0))


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_mg"])
@pytest.mark.parametrize("nels", [(6, 4), (8, 4, 4)])
def test_hard_void_breaks_down_promptly(T, gpu, nels, precision):
    from topopt_jl_amd import _lib
    prob, s = solver(T, nels, 0.0)
    s.vars = np.zeros(prob.ncells)
    t0 = time.perf_counter()
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        s(abstol=0.0, reltol=1e-8, maxiter=MAXITER, precision=precision)
    dt = time.perf_counter() - t0
    assert any("breakdown" in str(x.message) for x in w)
    assert not s.converged
    assert np.all(np.isnan(s.u))
    assert s.last_iters < 64  # within the first replayed graph, far below maxiter
    assert dt < 5.0
    # the context stays usable: a solid design solves normally afterwards
    s.vars = np.ones(prob.ncells)
    s(abstol=0.0, reltol=1e-8, maxiter=MAXITER, precision=precision)
    assert s.converged and np.all(np.isfinite(s.u))


def test_hard_void_return_code_through_abi(T, gpu):
    import ctypes as C
    from topopt_jl_amd import _lib
    from topopt_jl_amd._lib import dptr
    prob, s = solver(T, (6, 4), 0.0)
    s.set_densities(np.zeros(prob.ncells))
    u = np.zeros(prob.ndofs)
    it = C.c_int64()
    rc = s.lib.tohip_solve(s.ctx.ptr, dptr(np.ascontiguousarray(prob.f)), dptr(u), 0.0, 1e-8,
                           MAXITER, 1, _lib.TOHIP_FP64, C.byref(it), None)
    assert rc == _lib.TOHIP_BREAKDOWN
    assert b"breakdown" in s.lib.tohip_last_error(s.ctx.ptr)
    assert np.all(np.isnan(u)) and it.value == 0


def test_soft_void_still_solves(T, gpu):
    """xmin = 1e-9 (the soft void of the same reference test) stays positive
    definite: no breakdown, a finite displacement."""
    prob, s = solver(T, (6, 4), 1e-9)
    s.vars = np.zeros(prob.ncells)
    s(abstol=0.0, reltol=1e-8, maxiter=MAXITER)
    assert s.converged and np.all(np.isfinite(s.u))


def test_simp_run_stops_at_breakdown(T, gpu):
    from topopt_jl_amd import _lib
    prob, s = solver(T, (12, 6), 0.0)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        res = T.simp_run(s, volfrac=0.3, rmin=1.5, iterations=5, rel_tol=1e-8,
                         x0=np.zeros(prob.ncells), maxiter=MAXITER)
    assert any("breakdown" in str(x.message) for x in w)
    assert not res.converged
    assert np.isnan(res.obj[0]) and np.all(res.pcg_iters < 64)
