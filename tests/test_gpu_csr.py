"""GPU parity of the assembled-matrix comparator (HIPCGAssemblySolver,
TopOpt.jl's CGAssemblySolver) against the oracle's assemble! + apply!
matrix and its CG: the CSR values, the product, the solve with the identity
(the reference's default) and the Jacobi preconditioner, and `safe`."""
import numpy as np
import pytest

from conftest import relerr, rng

pytestmark = pytest.mark.gpu


def make(T, O, nels, xmin=1e-3, p=3.0, **kw):
    prob = T.PointLoadCantilever(nels)
    s = T.FEASolver(T.HIPCGAssemblySolver, prob, xmin=xmin, penalty=T.PowerPenaltyFun(p),
                    cg_max_iter=200000, **kw)
    orc = O.Oracle(O.Cantilever(tuple(nels)), xmin=xmin, p=p)
    return prob, s, orc


@pytest.mark.parametrize("nels", [(8, 4), (6, 4, 4), (9, 6, 4)])
def test_assembled_matrix_matches_oracle(T, O, gpu, nels):
    prob, s, orc = make(T, O, nels)
    x = rng(11).uniform(0.05, 1.0, prob.ncells)
    s.vars = x
    orc.set_vars(x)
    K = s.assembled_matrix()
    Ke, m = orc.assembled()
    scale = abs(Ke).max()
    assert abs(K - Ke).max() <= 1e-14 * scale
    # every row keeps the full neighbourhood pattern (9 or 27 nodes x dim)
    dim = len(nels)
    N = [n + 1 for n in nels]
    full = np.prod([3 * n - 2 for n in N]) * dim * dim  # sum over nodes of (neighbours x dim^2)
    assert K.nnz <= full
    v = rng(12).standard_normal(prob.ndofs)
    y = np.empty_like(v)
    from topopt_jl_amd._lib import dptr
    assert s.lib.tohip_csr_matvec(s.ctx.ptr, dptr(v), dptr(y)) == 0
    assert relerr(y, Ke @ v) < 1e-14
    free = ~prob.fixed.astype(bool)
    assert relerr(y[free], (Ke @ v)[free]) < 1e-14  # free rows alone


@pytest.mark.parametrize("jacobi", [False, True])
def test_assembled_cg_matches_oracle(T, O, gpu, jacobi):
    nels = (12, 6, 6)
    prob, s, orc = make(T, O, nels, preconditioner="jacobi" if jacobi else None)
    x = rng(13).uniform(0.1, 1.0, prob.ncells)
    s.vars = x
    orc.set_vars(x)
    s(abstol=0.0, reltol=1e-10)
    assert s.converged
    Ke, _ = orc.assembled()
    b = np.where(prob.fixed.astype(bool), 0.0, prob.f)
    uo, ito, _ = orc.cg_assembled(Ke, b, reltol=1e-10, jacobi=jacobi)
    # same recurrence, same stopping rule: the counts agree up to rounding
    assert abs(s.last_iters - ito) <= max(2, 0.02 * ito)
    assert relerr(s.u, uo) < 1e-7
    ud, _ = orc.direct(prob.f)
    assert relerr(s.u, ud) < 1e-7
    # the matrix-free path solves the same system (homogeneous conditions)
    mf = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-3, penalty=T.PowerPenaltyFun(3.0),
                     cg_max_iter=200000)
    mf.vars = x
    mf(abstol=0.0, reltol=1e-10)
    assert relerr(s.u, mf.u) < 1e-7


def test_assembled_safe_and_errors(T, O, gpu):
    """xmin = 0 with a void block leaves nodes whose every cell is void: a zero
    diagonal, which `safe` replaces by mean|diag| (solvers_api.jl:232-239)."""
    nels = (8, 4)
    prob, s, orc = make(T, O, nels, xmin=0.0, safe=True)
    x = np.ones(prob.ncells)
    x.reshape(4, 8)[1:3, 3:6] = 0.0  # interior nodes of this block touch only void cells
    s.vars = x
    orc.set_vars(x)
    K = s.assembled_matrix()
    Ke, m = orc.assembled(safe=True)
    assert abs(K - Ke).max() <= 1e-14 * abs(Ke).max()
    assert np.all(K.diagonal() > 0)
    with pytest.raises(ValueError):
        T.FEASolver(T.HIPCGAssemblySolver, prob, precision="mixed")
