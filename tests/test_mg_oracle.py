"""CPU tests: pin the multigrid oracle (oracle/mg_oracle.py) before it checks
the device V-cycle.  TopOpt.jl has no multigrid, so the pins are mathematical:
the fine operator is the pinned oracle matvec, the rediscretised coarse
operator is the Galerkin product P^T K P for uniform moduli (nested Q1
spaces), P reproduces linear fields, the cycle is symmetric positive
definite, and it cuts PCG iteration counts."""
import numpy as np
import pytest

from conftest import relerr, rng


@pytest.fixture(scope="module")
def MG(O):
    from oracle import mg_oracle
    return mg_oracle


def node_major(prob):
    """Caller (Ferrite) DOF index of node-major DOF dim * n + d."""
    dim = prob.dim
    idx = np.empty(prob.ndof, np.int64)
    for d in range(dim):
        idx[d::dim] = prob.node_dofs[d] - 1
    return idx


def node_mask(prob):
    dim = prob.dim
    return np.array([sum(int(prob.fixed[prob.node_dofs[d, n] - 1]) << d for d in range(dim))
                     for n in range(prob.node_dofs.shape[1])], np.uint8)


def test_level_grids_known_sizes(MG):
    assert MG.level_grids((120, 40, 40))[-1] == (15, 5, 5)
    assert len(MG.level_grids((120, 40, 40))) == 4
    assert MG.level_grids((256, 128, 128))[-1] == (8, 4, 4)
    assert MG.level_grids((160, 40)) == [(160, 40), (80, 20), (40, 10)]
    assert MG.level_grids((42, 14, 14))[-1] == (21, 7, 7)  # coarsest too large: sweep fallback
    with pytest.raises(ValueError):
        MG.level_grids((15, 4, 4))


# box cells: per-axis sizes (the oracle matvec's Gauss-quadrature Ke of the box)
@pytest.mark.parametrize("nels,sizes", [((8, 4, 4), None), ((12, 6), None),
                                        ((8, 4, 4), (1.0, 2.0, 0.5)), ((12, 6), (1.5, 0.75))])
def test_fine_operator_is_the_oracle_matvec(O, MG, nels, sizes):
    prob = O.Cantilever(nels)
    x = rng(0).uniform(0.05, 1.0, prob.nel)
    orc = O.Oracle(prob, xmin=1e-3, p=3.0, h=sizes)
    orc.set_vars(x)
    E = np.array([O.rho_drho(v, 3.0, 1e-3)[0] for v in x])
    K, _ = MG.assemble(nels, 1.0 if sizes is None else sizes, E, node_mask(prob), orc.meandiag)
    idx = node_major(prob)
    v = rng(1).standard_normal(prob.ndof)
    y = orc.matvec(v)
    assert relerr(K @ v[idx], y[idx]) < 1e-13


@pytest.mark.parametrize("nels,h", [((8, 4, 4), 1.0), ((12, 8), 1.0), ((8, 4, 4), (1.0, 2.0, 0.5)),
                                    ((12, 8), (1.0, 0.5))])
def test_rediscretised_coarse_operator_is_galerkin_for_uniform_moduli(O, MG, nels, h):
    """Nested Q1 spaces (cubes or boxes): P^T K_h P = K_2h on the free DOFs
    when E is uniform."""
    prob = O.Cantilever(nels)
    mask = node_mask(prob)
    E = np.full(prob.nel, 0.7)
    Kf, fxf = MG.assemble(nels, h, E, mask, 1.0)
    nc = tuple(n // 2 for n in nels)
    cmask = MG.coarse_mask(nels, nc, mask)
    Kc, fxc = MG.assemble(nc, 2.0 * np.asarray(h), MG.coarsen_E(E, nels), cmask, 1.0)
    P = MG.prolongation(nels)
    G = (P.T @ Kf @ P).toarray()
    free = ~fxc
    assert np.max(np.abs(G[np.ix_(free, free)] - Kc.toarray()[np.ix_(free, free)])) < 1e-13
    # the clamped x = 0 face stays the only prescribed set on the coarse level
    dim = len(nels)
    NXc = nc[0] + 1
    nodes = np.arange(Kc.shape[0]) // dim
    assert np.array_equal(fxc, nodes % NXc == 0)


@pytest.mark.parametrize("nels", [(8, 4, 6), (10, 4)])
def test_prolongation_reproduces_linear_fields(MG, nels):
    dim = len(nels)
    P = MG.prolongation(nels)
    F = [n + 1 for n in nels]
    C = [n // 2 + 1 for n in nels]

    def coords(N, h):
        g = np.meshgrid(*[np.arange(n) * h for n in N], indexing="ij")
        return [a.transpose(*reversed(range(dim))).ravel() for a in g]  # x-fastest

    cf, cc = coords(F, 1.0), coords(C, 2.0)
    a = np.array([0.3, -1.2, 0.7][:dim])
    for d in range(dim):
        uc = np.zeros(P.shape[1])
        uc[d::dim] = 2.0 + sum(a[k] * cc[k] for k in range(dim))
        uf = P @ uc
        assert np.max(np.abs(uf[d::dim] - (2.0 + sum(a[k] * cf[k] for k in range(dim))))) < 1e-13


@pytest.mark.parametrize("nels", [(16, 8, 8), (42, 14, 14), (24, 8)])
def test_vcycle_is_symmetric_positive_definite(O, MG, nels):
    prob = O.Cantilever(nels)
    g = rng(5)
    E = 1e-3 + (1 - 1e-3) * g.uniform(0, 1, prob.nel) ** 3
    md = 1.0
    M = MG.MGOracle(nels, E, node_mask(prob), md)
    fx = M.fx[0]
    r1 = g.standard_normal(fx.size); r1[fx] = 0
    r2 = g.standard_normal(fx.size); r2[fx] = 0
    a, b = M.apply(r1) @ r2, r1 @ M.apply(r2)
    assert abs(a - b) <= 1e-12 * max(abs(a), 1e-300) + 1e-14
    assert M.apply(r1) @ r1 > 0 and M.apply(r2) @ r2 > 0


def test_vcycle_cuts_pcg_iterations(O, MG):
    """On a SIMP-like design (smooth random field cubed, xmin 1e-9) the
    preconditioned solve needs several times fewer iterations than Jacobi."""
    nels = (32, 16, 16)
    prob = O.Cantilever(nels, load="centre")
    g = rng(7)
    base = g.uniform(0, 1, (nels[2] // 4, nels[1] // 4, nels[0] // 4))
    field = np.kron(base, np.ones((4, 4, 4))).ravel()
    E = 1e-9 + (1 - 1e-9) * np.clip(field, 0.05, 1.0) ** 3
    M = MG.MGOracle(nels, E, node_mask(prob), 1.0)
    K = M.K[0]
    f = prob.f[node_major(prob)]
    it_mg, x = MG.pcg_iterations(K, f, M.apply)
    it_j, _ = MG.pcg_iterations(K, f, lambda r: M.dinv[0] * r)
    assert np.linalg.norm(f - K @ x) <= 1e-8 * np.linalg.norm(f)
    assert it_mg * 4 < it_j, (it_mg, it_j)


def test_jacobi_weight_bounds_every_design(O, MG):
    """The per-level weight keeps w * lambda_max(D^-1 K) < 2 for a
    checkerboard-like random design, where lambda_max exceeds that of a
    uniform one (a fixed weight of 0.6 fails this: 0.6 * 4.17 > 2)."""
    nels = (16, 8, 8)
    prob = O.Cantilever(nels)
    E = 1e-9 + rng(9).uniform(0.05, 1.0, prob.nel) ** 3
    K, fx = MG.assemble(nels, 1.0, E, node_mask(prob), 1.0)
    Dm = 1.0 / np.sqrt(K.diagonal())
    lam = np.linalg.eigvalsh((K.toarray() * Dm[:, None]) * Dm[None, :]).max()
    w = MG.jacobi_weight(3, 1.0)
    assert lam > 2.0 / 0.6 and w * lam < 2.0, (lam, w)
    assert abs(w - 1.9 / (1.01 * 5.318181818181818)) < 1e-12
