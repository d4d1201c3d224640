"""CPU tests: pin the oracle (oracle/oracle.c) before it is trusted as the checker.

Pins, in order of strength:
 * the reference's own known answers (force_dof numbers, plane-stress Q4
   closed form of test/topopt_problems/element_stiffness_matrix.jl:11-33 with
   plane-strain constants, PowerPenalty values of test/Utilities/test_penalties.jl);
 * golden fixtures from an independent numpy restatement (tests/golden);
 * mathematical invariants chosen so that plausible slips fail (rigid modes,
   patch test, symmetry, u'Ku = f'u, central finite differences).
"""
import math
import os

import numpy as np
import pytest

from conftest import relerr

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = ["cantilever2d_6x4", "cantilever3d_3x2x2", "cantilever3d_4x2x4"]


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


# ------------------------------------------------------------ mesh / DOFs --

def test_force_dof_known_answers(O):
    # test/topopt_problems/problems.jl:20 and :63
    assert O.Cantilever((160, 40)).force_dof == 161 * 21 * 2
    node_dofs, _, _ = O.numbering((60, 20))
    assert node_dofs[1, 61 * 20] == (61 * 20 + 2) * 2  # HalfMBB top-left node


def test_numbering_shape_facts(O):
    # test/ferrite_upgrade_behavior.jl:4-26
    nd, cd, cn = O.numbering((10, 6))
    assert cd.shape[0] == 8 and nd.shape == (2, 11 * 7)
    nd3, cd3, cn3 = O.numbering((4, 4, 2))
    assert cn3.shape[0] == 8 and cd3.shape == (24, 32)
    assert sorted(nd3.ravel()) == list(range(1, nd3.size + 1))


@pytest.mark.parametrize("name", CASES)
def test_numbering_matches_golden(O, name):
    d = gold(name)
    nd, cd, _ = O.numbering(tuple(d["nels"]))
    assert np.array_equal(nd, d["node_dofs"]) and np.array_equal(cd, d["cell_dofs"])


def test_load_node(O):
    p2 = O.Cantilever((8, 4))
    assert p2.load_node == 8 + 9 * 2
    p3 = O.Cantilever((4, 2, 4))
    assert p3.load_node == 4 + 5 * 1  # reference choice: z = 0
    p3c = O.Cantilever((4, 2, 4), load="centre")
    assert p3c.load_node == 4 + 5 * (1 + 3 * 2)
    with pytest.raises(ValueError):
        O.Cantilever((4, 3))


# ------------------------------------------------------------------- Ke ----

def test_q4_plane_strain_closed_form(O):
    E, nu = 1.0, 0.3
    Es, nus = E / (1 - nu ** 2), nu / (1 - nu)  # plane strain via plane-stress formula
    k = np.array([1 / 2 - nus / 6, 1 / 8 + nus / 8, -1 / 4 - nus / 12, -1 / 8 + 3 * nus / 8,
                  -1 / 4 + nus / 12, -1 / 8 - nus / 8, nus / 6, 1 / 8 - 3 * nus / 8])
    idx = np.array([[1, 2, 3, 4, 5, 6, 7, 8], [2, 1, 8, 7, 6, 5, 4, 3], [3, 8, 1, 6, 7, 4, 5, 2],
                    [4, 7, 6, 1, 8, 3, 2, 5], [5, 6, 7, 8, 1, 2, 3, 4], [6, 5, 4, 3, 2, 1, 8, 7],
                    [7, 4, 5, 2, 3, 8, 1, 6], [8, 3, 2, 5, 4, 7, 6, 1]]) - 1
    KE = Es / (1 - nus ** 2) * k[idx]
    assert np.max(np.abs(O.element_stiffness(2, E, nu) - KE)) < 1e-15


@pytest.mark.parametrize("dim", [2, 3])
def test_ke_invariants(O, dim):
    K = O.element_stiffness(dim)
    assert np.max(np.abs(K - K.T)) < 1e-15
    ev = np.linalg.eigvalsh(K)
    nrigid = 3 if dim == 2 else 6
    assert np.all(np.abs(ev[:nrigid]) < 1e-13) and ev[nrigid] > 1e-3
    for q in (2, 3):
        assert np.max(np.abs(O.element_stiffness(dim, quad_order=q) - K)) < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_ke_matches_golden_and_cell_size(O, dim):
    d = gold("cantilever2d_6x4" if dim == 2 else "cantilever3d_3x2x2")
    assert np.max(np.abs(O.element_stiffness(dim) - d["Ke"])) < 1e-14
    b = np.load(os.path.join(GOLD, "ke_boxes.npz"))
    h = [1.0, 2.0] if dim == 2 else [1.0, 2.0, 0.5]
    Kh = O.element_stiffness(dim, 2.0, 0.25, h=h)
    assert np.max(np.abs(Kh - b["ke%d_h" % dim])) < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_patch_test_uniaxial_strain(O, dim):
    """u_x = x (uniform strain e_xx = 1): the x-reaction on the x = 1 face of
    one unit cell equals sigma_xx * area = (lam + 2 mu) (3D Lame constants,
    plane strain in 2D)."""
    E, nu = 1.0, 0.3
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    K = O.element_stiffness(dim, E, nu)
    off = [[0, 0], [1, 0], [1, 1], [0, 1]] if dim == 2 else \
        [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]
    u = np.zeros(K.shape[0])
    for a, o in enumerate(off):
        u[dim * a] = o[0]
    fint = K @ u
    right = [a for a, o in enumerate(off) if o[0] == 1]
    fx = sum(fint[dim * a] for a in right)
    assert abs(fx - (lam + 2 * mu)) < 1e-14
    if dim == 3:  # lateral reaction on the y = 1 face: sigma_yy = lam
        top = [a for a, o in enumerate(off) if o[1] == 1]
        assert abs(sum(fint[dim * a + 1] for a in top) - lam) < 1e-14


def test_penalty_values(O):
    # test/Utilities/test_penalties.jl:17-54 (PowerPenaltyFun) through density()
    for p, x, val in [(3.0, 0.5, 0.125), (1.0, 0.5, 0.5), (2.0, 0.5, 0.25), (3.0, 0.0, 0.0)]:
        r, dr = O.rho_drho(x, p, 0.0)
        assert abs(r - val) < 1e-15
        assert abs(dr - (p * x ** (p - 1) if p != 1 else 1.0)) < 1e-15
    r, dr = O.rho_drho(0.5, 3.0, 1e-3)
    assert abs(r - (0.125 * (1 - 1e-3) + 1e-3)) < 1e-15
    r, dr = O.rho_drho(0.5, 3.0, 1e-3, pbi=False)  # (xmin + (1-xmin) x)^p
    d = 0.5 * (1 - 1e-3) + 1e-3
    assert abs(r - d ** 3) < 1e-15 and abs(dr - 3 * d ** 2 * (1 - 1e-3)) < 1e-15


POWER, RATIONAL, SINH = 0, 1, 2
NOPROJ, HEAVISIDE, SIGMOID = 0, 1, 2


def _pen(O, x, kind, p, proj=NOPROJ, beta=0.0):
    """P(x) and P'(x) alone: xmin = 0 with the penalty applied before the
    interpolation leaves rho = P(x)."""
    return O.rho_drho(x, p, 0.0, pbi=True, kind=kind, proj=proj, beta=beta)


def test_rational_and_sinh_known_answers(O):
    # test/Utilities/test_penalties.jl:117-179
    v, d = _pen(O, 0.5, RATIONAL, 3.0)
    assert abs(v - 0.5 / (1 + 3 * 0.5)) < 1e-15
    assert abs(d - (1 + 3.0) / (1 + 3.0 * 0.5) ** 2) < 1e-15  # ForwardDiff known answer
    assert _pen(O, 0.0, RATIONAL, 3.0)[0] == 0.0 and abs(_pen(O, 1.0, RATIONAL, 3.0)[0] - 1) < 1e-15
    v, d = _pen(O, 0.5, SINH, 3.0)
    assert abs(v - math.sinh(1.5) / math.sinh(3.0)) < 1e-15
    assert abs(d - 3.0 * math.cosh(1.5) / math.sinh(3.0)) < 1e-14
    assert _pen(O, 0.0, SINH, 3.0)[0] == 0.0 and abs(_pen(O, 1.0, SINH, 3.0)[0] - 1) < 1e-10
    # p = 1 power is the identity with unit slope everywhere (x = 0 included)
    assert _pen(O, 0.0, POWER, 1.0) == (0.0, 1.0)


def test_projection_known_answers(O):
    # test/Utilities/test_penalties.jl:181-268; PowerPenaltyFun(1) is the identity
    for beta in (5.0, 3.0):
        v, d = _pen(O, 0.5, POWER, 1.0, HEAVISIDE, beta)
        assert abs(v - (1 - math.exp(-beta * 0.5) + 0.5 * math.exp(-beta))) < 1e-15 and d > 0
    assert _pen(O, 0.0, POWER, 1.0, HEAVISIDE, 5.0)[0] < 0.01
    assert _pen(O, 1.0, POWER, 1.0, HEAVISIDE, 5.0)[0] > 0.99
    for beta in (4.0, 2.0):
        v, d = _pen(O, 0.5, POWER, 1.0, SIGMOID, beta)
        assert abs(v - 0.5) < 1e-15 and d > 0 and math.isfinite(d)
    assert _pen(O, 0.0, POWER, 1.0, SIGMOID, 4.0)[0] < 0.1
    assert _pen(O, 1.0, POWER, 1.0, SIGMOID, 4.0)[0] > 0.9
    # sigmoid slope at 0.5 is (beta + 1) / 4
    assert abs(_pen(O, 0.5, POWER, 1.0, SIGMOID, 4.0)[1] - 5.0 / 4) < 1e-15


def test_projected_penalty_known_answers(O):
    # test/Utilities/test_penalties.jl:270-317: penalty(proj(x))
    hp = lambda x: 1 - math.exp(-5.0 * x) + x * math.exp(-5.0)
    for x in (0.1, 0.5, 0.9):
        assert abs(_pen(O, x, POWER, 3.0, HEAVISIDE, 5.0)[0] - hp(x) ** 3) < 1e-15
    sg = lambda x: 1 / (1 + math.exp((3.0 + 1) * (-x + 0.5)))
    rp = lambda y: y / (1 + 2.0 * (1 - y))
    for x in (0.2, 0.6, 0.8):
        assert abs(_pen(O, x, RATIONAL, 2.0, SIGMOID, 3.0)[0] - rp(sg(x))) < 1e-15


@pytest.mark.parametrize("kind,p", [(POWER, 3.0), (POWER, 1.0), (RATIONAL, 3.0), (SINH, 2.5)])
@pytest.mark.parametrize("proj,beta", [(NOPROJ, 0.0), (HEAVISIDE, 4.0), (SIGMOID, 3.0)])
@pytest.mark.parametrize("pbi", [True, False])
def test_rho_derivative_central_differences(O, kind, p, proj, beta, pbi):
    """get_rho_drho's derivative against central differences of get_rho, with
    the interpolation density(x, xmin) in both orders."""
    xmin, h = 1e-3, 1e-6
    for x in (0.05, 0.3, 0.5, 0.77, 0.95):
        r, dr = O.rho_drho(x, p, xmin, pbi, kind, proj, beta)
        rp_, _ = O.rho_drho(x + h, p, xmin, pbi, kind, proj, beta)
        rm_, _ = O.rho_drho(x - h, p, xmin, pbi, kind, proj, beta)
        fd = (rp_ - rm_) / (2 * h)
        assert abs(dr - fd) <= 1e-7 * max(1.0, abs(dr)), (x, dr, fd)
        # interpolation order: rho = xmin + (1-xmin) P(x) or P(xmin + (1-xmin) x)
        if pbi:
            assert abs(r - (xmin + (1 - xmin) * _pen(O, x, kind, p, proj, beta)[0])) < 1e-15
        else:
            assert abs(r - _pen(O, xmin + (1 - xmin) * x, kind, p, proj, beta)[0]) < 1e-15


# ---------------------------------------------------------- operator ----

@pytest.mark.parametrize("name", CASES)
def test_matvec_matches_dense_golden(O, name):
    d = gold(name)
    prob = O.Cantilever(tuple(d["nels"]))
    assert np.array_equal(prob.fixed, d["fixed"]) and np.array_equal(prob.f, d["f"])
    orc = O.Oracle(prob, xmin=float(d["xmin"]), p=float(d["p"]))
    orc.set_vars(d["x"])
    assert abs(orc.meandiag - float(d["meandiag"])) < 1e-12 * float(d["meandiag"])
    assert relerr(orc.matvec(d["v"]), d["Kv"]) < 1e-14


def test_matvec_symmetry_and_rigid_mode(O):
    prob = O.Cantilever((7, 4, 4))
    orc = O.Oracle(prob, xmin=1e-3, p=3.0)
    g = np.random.default_rng(0)
    orc.set_vars(g.uniform(0.01, 1, prob.nel))
    a, b = g.standard_normal(prob.ndof), g.standard_normal(prob.ndof)
    assert abs(a @ orc.matvec(b) - b @ orc.matvec(a)) < 1e-12 * abs(a @ orc.matvec(b))
    # a rigid translation gives zero force on free DOFs whose cells are all free
    t = np.zeros(prob.ndof)
    t[prob.node_dofs[1] - 1] = 1.0
    t[prob.fixed.astype(bool)] = 0.0
    y = orc.matvec(t)
    NX = 8
    interior = [n for n in range(prob.node_dofs.shape[1]) if n % NX >= 2]
    assert np.max(np.abs(y[prob.node_dofs[:, interior] - 1])) < 1e-13


# ------------------------------------------------------------ solves ----

@pytest.mark.parametrize("name", CASES)
def test_direct_and_pcg_match_golden(O, name):
    d = gold(name)
    prob = O.Cantilever(tuple(d["nels"]))
    orc = O.Oracle(prob, xmin=float(d["xmin"]), p=float(d["p"]))
    orc.set_vars(d["x"])
    u, rc = orc.direct(prob.f)
    assert rc == 0 and relerr(u, d["u"]) < 1e-12
    up, it, res = orc.pcg(prob.f, abstol=0.0, reltol=1e-14, maxiter=10000)
    assert relerr(up, d["u"]) < 1e-10
    obj, ce, g = orc.element_energy(u)
    assert relerr(ce, d["ce"]) < 1e-12 and relerr(g, d["grad"]) < 1e-12
    assert abs(obj - float(d["obj"])) < 1e-12 * abs(obj)
    # compliance identity u'Ku = f'u (test/FEA/misc.jl:7-28)
    assert abs(obj - prob.f @ u) < 1e-10 * abs(obj)


def test_pcg_stopping_rule_and_warm_start(O):
    prob = O.Cantilever((16, 6))
    orc = O.Oracle(prob, xmin=1e-3, p=3.0)
    orc.set_vars(np.random.default_rng(3).uniform(0.1, 1, prob.nel))
    u, it, res = orc.pcg(prob.f, abstol=0.0, reltol=1e-6)
    assert res <= 1e-6 * np.linalg.norm(prob.f)
    u2, it2, res2 = orc.pcg(prob.f, abstol=0.0, reltol=1e-6, maxiter=it - 1)
    assert it2 == it - 1 and res2 > 1e-6 * np.linalg.norm(prob.f)
    u3, it3, _ = orc.pcg(prob.f, x0=u, abstol=1e-6 * np.linalg.norm(prob.f), reltol=0.0)
    assert it3 == 0


@pytest.mark.parametrize("p", [1.0, 2.0, 3.0])
def test_compliance_gradient_finite_differences(O, p):
    """central differences of compliance (direct solve), as in
    test/Functions/test_common_fns.jl:32-49 (xmin = 0.01)."""
    prob = O.Cantilever((4, 2))
    orc = O.Oracle(prob, xmin=0.01, p=p)
    g = np.random.default_rng(4)
    x = g.uniform(0.2, 0.9, prob.nel)

    def comp(xx):
        orc.set_vars(xx)
        u, _ = orc.direct(prob.f)
        return orc.element_energy(u)

    obj, ce, grad = comp(x)
    h = 1e-6
    for e in range(prob.nel):
        xp, xm = x.copy(), x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (comp(xp)[0] - comp(xm)[0]) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-5 * max(1.0, abs(grad[e]))


# ------------------------------------------------------------ filter ----

@pytest.mark.parametrize("name,r", [("cantilever2d_6x4", 1.5), ("cantilever2d_6x4", 2.0),
                                    ("cantilever2d_6x4", 3.0), ("cantilever3d_3x2x2", 1.5),
                                    ("cantilever3d_3x2x2", 2.0), ("cantilever3d_4x2x4", 1.5),
                                    ("cantilever3d_4x2x4", 2.5)])
def test_filter_bfs_matches_closed_form(O, name, r):
    d = gold(name)
    F = O.Filter(tuple(d["nels"]), r)
    key = str(r).replace(".", "p")
    assert np.max(np.abs(F.dense() - d["J_" + key])) < 1e-14
    assert relerr(F.apply(d["xf"]), d["Jx_" + key]) < 1e-14


def test_filter_properties(O):
    F = O.Filter((9, 6, 4), 2.0)
    n = F.nel
    np.testing.assert_allclose(F.apply(np.full(n, 0.5)), 0.5, rtol=1e-14)  # test_filters.jl:126-141
    g = np.random.default_rng(5)
    a, b = g.standard_normal(n), g.standard_normal(n)
    assert abs(b @ F.apply(a) - a @ F.apply_t(b)) < 1e-12 * abs(b @ F.apply(a))
    with pytest.raises(ValueError):  # ArgumentError for a too-small rmin (test_filters.jl:71-73)
        O.Filter((6, 4), 0.01)


def test_filter_stencil_weights_interior(O):
    """3D rmin = 1.5: only the 8 own nodes (d = sqrt(3)/2) contribute, so an
    interior cell averages its 27 neighbours weighted 8/4/2/1 over 64."""
    F = O.Filter((5, 5, 5), 1.5)
    J = F.dense()
    c = 2 + 5 * (2 + 5 * 2)
    row = J[c].reshape(5, 5, 5)  # [k, j, i]
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                shared = (2 - abs(di)) * (2 - abs(dj)) * (2 - abs(dk))
                assert abs(row[2 + dk, 2 + dj, 2 + di] - shared / 64) < 1e-15
    assert abs(J[c].sum() - 1) < 1e-15


# ---------------------------------------------------------------- OC ----

def test_oc_invariants(O):
    g = np.random.default_rng(6)
    F = O.Filter((12, 6), 2.0)
    x = g.uniform(0.1, 0.9, F.nel)
    dc = -g.uniform(0.0, 3.0, F.nel)
    xn, lam = O.oc_update(x, dc, F.colsum, 0.4, 0.2, 1e-13)
    assert abs(F.colsum @ xn / F.nel - 0.4) < 1e-9
    assert np.all(xn >= np.maximum(0, x - 0.2) - 1e-15) and np.all(xn <= np.minimum(1, x + 0.2) + 1e-15)
    # the volume is J x_new (sum of filtered densities)
    assert abs(F.apply(xn).mean() - 0.4) < 1e-9


def oc_closed_form_case(colsum, seed):
    """An OC step whose update stays strictly inside the move and box limits,
    built backwards from its answer.  With dv_e = colsum_e / nel, the
    unclamped update is x_new = x * sqrt(-dc / (lambda dv)); choosing a target
    t and lambda0, dc_e = -(t_e / x_e)^2 lambda0 dv_e makes t the update at
    lambda0, and the volume target sum_e colsum_e t_e / nel makes lambda0 the
    multiplier the bisection must find."""
    g = np.random.default_rng(seed)
    n = colsum.size
    x = g.uniform(0.3, 0.7, n)
    t = x * (1.0 + g.uniform(-0.25, 0.25, n))  # |t - x| <= 0.175 < move 0.2
    lam0 = 37.25
    dv = colsum / n
    dc = -(t / x) ** 2 * lam0 * dv
    volfrac = float(colsum @ t) / n
    return x, dc, volfrac, t, lam0


@pytest.mark.parametrize("nels,rmin", [((12, 6), 2.0), ((5, 4, 3), 1.5)])
def test_oc_step_closed_form(O, nels, rmin):
    """One OC step against its closed form: x_new = x sqrt(B_e) with
    B_e = -dc_e / (lambda dv_e), lambda fixed by the volume constraint.  A
    dropped square root, another damping exponent or a different dv (so a
    different lambda) fails here."""
    F = O.Filter(nels, rmin)
    x, dc, vf, t, lam0 = oc_closed_form_case(F.colsum, 3)
    xn, lam = O.oc_update(x, dc, F.colsum, vf, 0.2, 1e-14)
    assert abs(lam - lam0) <= 1e-12 * lam0
    assert np.max(np.abs(xn - t)) < 1e-12
    # the same data with an exponent of 1 (no square root) lands elsewhere
    assert np.max(np.abs(np.clip(x * (-dc / (lam0 * F.colsum / F.nel)), 0, 1) - t)) > 1e-2


def test_oc_converges_to_analytic_optimum(O):
    """OC on c(x) = sum_e a_e / x_e with V(x) = mean(x) <= vf (identity filter:
    colsum = 1, dv = 1/n).  Stationarity a_e / x_e^2 = lambda' gives the
    optimum x*_e = n vf sqrt(a_e) / sum sqrt(a); every OC step lands on
    sqrt(a / (lambda dv)) when the move limit allows, so the iteration reaches
    x* and stays; at x*, -dc = lambda dv on every cell."""
    g = np.random.default_rng(4)
    n, vf = 400, 0.45
    a = g.uniform(0.5, 2.0, n)
    xstar = n * vf * np.sqrt(a) / np.sqrt(a).sum()
    assert 0.05 < xstar.min() and xstar.max() < 0.95
    ones = np.ones(n)
    x = np.full(n, vf)
    for _ in range(8):
        x, lam = O.oc_update(x, -a / x ** 2, ones, vf, 0.2, 1e-15)
    assert np.max(np.abs(x - xstar)) < 1e-10
    np.testing.assert_allclose(a / x ** 2, lam / n, rtol=1e-9)


def test_simp_2d_runs_and_decreases(O):
    r = O.simp(O.Cantilever((30, 10)), 0.5, 2.0, 15, solver="direct")
    assert np.all(np.isfinite(r["obj"]))
    assert r["obj"][-1] < r["obj"][0]
    assert abs(r["xphys"].mean() - 0.5) < 1e-6


def test_projection_alone(O):
    # HeavisideProjectionFun / SigmoidProjectionFun (test/Utilities/test_penalties.jl:181-268)
    v, d = O.project([0.0, 0.5, 1.0], HEAVISIDE, 5.0)
    assert abs(v[1] - (1 - math.exp(-2.5) + 0.5 * math.exp(-5.0))) < 1e-15
    assert v[0] < 0.01 and v[2] > 0.99 and np.all(d > 0)
    v, d = O.project([0.5], SIGMOID, 4.0)
    assert abs(v[0] - 0.5) < 1e-15 and abs(d[0] - 5.0 / 4) < 1e-15
    v, d = O.project([0.3], NOPROJ, 0.0)
    assert v[0] == 0.3 and d[0] == 1.0


# ------------------------------------------------------------------ BESO ----

def test_beso_oracle_reference_invariants(O):
    """test/Algorithms/test_beso.jl and test/integration/test_end_to_end.jl:161-183:
    binary topology, fevals within maxiter, volume near the target once it is
    reached; the objective trace and change follow beso.jl:146-177."""
    prob = O.Cantilever((12, 6))
    r = O.beso(prob, 0.5, 2.0, maxiter=15, tol=0.05, p=1.0)
    t = r["topology"]
    assert np.all((t == 0) | (t == 1)) and 1 <= r["fevals"] <= 15
    assert abs(t.mean() - 0.5) < 0.1
    assert r["objval"] == r["obj"][-1]
    # volume target shrinks by er per iteration: after the first iteration the
    # bisection picks (within one cell) the count that meets 0.98 * V0
    r1 = O.beso(prob, 0.5, 2.0, maxiter=1, tol=0.05, p=1.0, x0=np.full(prob.nel, 0.8))
    assert r1["fevals"] == 1
    assert abs(r1["topology"].sum() - 0.98 * prob.nel) <= 1
    # black / white cells keep their values
    black = np.zeros(prob.nel, bool); black[:6] = True
    white = np.zeros(prob.nel, bool); white[-6:] = True
    r2 = O.beso(prob, 0.4, 2.0, maxiter=5, tol=0.05, p=1.0, black=black, white=white)
    assert np.all(r2["topology"][black] == 1) and np.all(r2["topology"][white] == 0)
    # change is |sum(first half) - sum(second half)| / sum(second half) of the last k objectives
    r3 = O.beso(prob, 0.5, 2.0, maxiter=12, tol=1e-9, p=1.0, k=4)
    tr = r3["obj"][-4:]
    assert abs(r3["change"] - abs(tr[:2].sum() - tr[2:].sum()) / tr[2:].sum()) < 1e-15


# ------------------------------------------------------ EnergyCriteria ----

def _energy_ratios(orc, f, x, energy):
    """(|x'Ax - energy| / x'Ax, |x'r| / x'Ax, x'Ax) with the true residual."""
    r = f - orc.matvec(x)
    xtr = float(x @ r)
    xAx = float(f @ x) - xtr
    return abs(xAx - energy) / xAx if xAx else math.inf, abs(xtr) / xAx if xAx else math.inf, xAx


def test_pcg_energy_criteria(O):
    """orc_pcg_energy follows orc_pcg's iterates (the same recurrence) and
    stops at the first iterate where both EnergyCriteria ratios, evaluated
    independently in numpy along the energy chain, fall below tol
    (src/FEA/convergence_criteria.jl:26-45)."""
    prob = O.Cantilever((6, 4, 4))
    orc = O.Oracle(prob, xmin=1e-3, p=3.0)
    g = np.random.default_rng(8)
    orc.set_vars(g.uniform(0.05, 1, prob.nel))
    f = prob.f
    tol = 1e-6
    x, k, energy, status, _ = orc.pcg_energy(f, tol)
    assert status == 1 and k > 5
    xk, kk, _ = orc.pcg(f, abstol=0.0, reltol=0.0, maxiter=k)
    assert kk == k and np.array_equal(xk, x)
    # independent evaluation of the stopping test at every iterate
    e = 0.0
    first = None
    for it in range(k + 1):
        xi = orc.pcg(f, abstol=0.0, reltol=0.0, maxiter=it)[0] if it else np.zeros(prob.ndof)
        r1, r2, xAx = _energy_ratios(orc, f, xi, e)
        e = xAx
        if it < k:  # not converged before k (with a rounding margin)
            assert not (r1 <= tol * (1 - 1e-6) and r2 <= tol * (1 - 1e-6)), it
        else:
            assert r1 <= tol * (1 + 1e-6) and r2 <= tol * (1 + 1e-6)
    assert energy == pytest.approx(float(x @ orc.matvec(x)), rel=1e-12)
    # the state carries over: restarting from the solution converges at once
    x2, k2, e2, s2, _ = orc.pcg_energy(f, tol, x0=x, energy=energy)
    assert (k2, s2) == (0, 1) and np.array_equal(x2, x)
    # maxiter short-circuits the test: energy is the last evaluated iterate's
    x3, k3, e3, s3, _ = orc.pcg_energy(f, tol, maxiter=k - 2)
    xm = orc.pcg(f, abstol=0.0, reltol=0.0, maxiter=k - 3)[0]
    assert (k3, s3) == (k - 2, 0)
    assert e3 == pytest.approx(float(xm @ orc.matvec(xm)), rel=1e-12)
    # a NaN energy is the reference's DomainError
    fb = f.copy()
    fb[0] = np.nan
    assert orc.pcg_energy(fb, tol)[3] == -1


# ------------------------------------------- assembled K (CGAssemblySolver) ----

def test_assembled_matrix_rule(O):
    """assemble! + apply!: free rows equal the matrix-free product (which
    skips prescribed columns), prescribed rows are mean|diag| times x, the
    matrix is symmetric, and CG on it reaches the direct solution."""
    prob = O.Cantilever((6, 4, 4))
    orc = O.Oracle(prob, xmin=1e-3, p=3.0)
    g = np.random.default_rng(5)
    orc.set_vars(g.uniform(0.05, 1, prob.nel))
    K, m = orc.assembled()
    assert abs(K - K.T).max() <= 1e-15 * abs(K).max()  # Ke symmetric to rounding
    x = g.standard_normal(prob.ndof)
    fx = prob.fixed.astype(bool)
    y = K @ x
    ymf = orc.matvec(x)
    assert relerr(y[~fx], ymf[~fx]) < 1e-14
    np.testing.assert_allclose(y[fx], m * x[fx], rtol=1e-15)
    # the mean |diag| of the unconstrained K, by brute force over the cells
    diag = np.zeros(prob.ndof)
    r = 1e-3 + (1 - 1e-3) * orc.vars ** 3
    for e in range(prob.nel):
        diag[prob.cell_dofs[:, e] - 1] += r[e] * np.diag(orc.Ke)
    assert m == pytest.approx(np.abs(diag).mean(), rel=1e-13)
    u, it, _ = orc.cg_assembled(K, np.where(fx, 0.0, prob.f), reltol=1e-12, jacobi=True)
    ud, rc = orc.direct(prob.f)
    assert rc == 0 and relerr(u, ud) < 1e-8


# ---------------------------------------------------------------- VolumeFun ----
@pytest.mark.parametrize("dim,h", [(2, (1.0, 1.0)), (2, (1.5, 0.75)), (3, (1.0, 1.0, 1.0)),
                                   (3, (1.0, 2.0, 0.5))])
@pytest.mark.parametrize("q", [2, 3, 4])
def test_cell_volume_closed_form(O, dim, h, q):
    """get_cell_volumes' quadrature of det J equals the box volume prod(h) at
    every order (det J is constant on an axis-aligned box)."""
    nels = (3,) * dim
    v = O.cell_volumes(nels, h, quad_order=q)
    np.testing.assert_allclose(v, np.prod(h), rtol=1e-15)


def test_volume_fun_pins(O):
    """VolumeFun (src/Functions/volume.jl:59-80): linear in x, V(1) = 1 as a
    fraction, V(e_k) = V_k / V_total (its constant gradient), absolute volume
    when fraction = false."""
    nels, h = (5, 4, 3), (1.0, 2.0, 0.5)
    cv = O.cell_volumes(nels, h)
    n = cv.size
    g = np.random.default_rng(8)
    a, b = g.uniform(0, 1, n), g.uniform(0, 1, n)
    assert abs(O.volume(2.0 * a - 3.0 * b, cv) - (2.0 * O.volume(a, cv) - 3.0 * O.volume(b, cv))) < 1e-14
    assert abs(O.volume(np.ones(n), cv) - 1.0) < 1e-15
    e = np.zeros(n)
    e[7] = 1.0
    assert abs(O.volume(e, cv) - cv[7] / cv.sum()) <= 1e-15 * cv[7] / cv.sum()
    assert abs(O.volume(a, cv, fraction=False) - a.sum() * np.prod(h)) < 1e-12


@pytest.mark.parametrize("nels,rmin,h", [((12, 6), 2.0, None), ((20, 8), 3.0, (1.0, 2.0)),
                                         ((5, 4, 3), 1.5, None), ((6, 5, 4), 2.0, None),
                                         ((6, 4, 4), 1.5, (1.0, 2.0, 0.5)), ((7, 6, 5), 2.5, None)])
def test_filter_closed_form_matches_bfs(O, nels, rmin, h):
    """The numpy closed form of the density filter (used where the sparse J
    is too large to build) equals the BFS-built J and J^T."""
    F = O.Filter(nels, rmin, h)
    g = np.random.default_rng(9)
    x = g.uniform(0, 1, F.nel)
    assert relerr(O.filter_closed_form(nels, rmin, x, h=h), F.apply(x)) < 1e-14
    d = g.standard_normal(F.nel)
    assert relerr(O.filter_closed_form(nels, rmin, d, transpose=True, h=h), F.apply_t(d)) < 1e-14
