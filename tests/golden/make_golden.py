"""Writes tests/golden/*.npz: small known-answer fixtures for the oracle.

Everything here is an independent numpy restatement (different code and a
different integration method from oracle/oracle.c), so agreement pins the
oracle against a plausible slip in either (a dropped term, a wrong sign or
index, a transposed operand):

* Ke by isoparametric Gauss-Legendre quadrature (numpy.polynomial.legendre,
  3 points) with an explicit Jacobian, Ferrite local node order, 3D Lame
  lambda in both 2D (plane strain) and 3D
  (TopOpt.jl src/TopOptProblems/matrices_and_vectors.jl:83-87, 126-177);
* dense global K with the matrix-free boundary rule of
  src/FEA/matrix_free_operator.jl:66-105 (rows/columns of prescribed DOFs
  removed from the element products, diagonal meandiag = sum_e tr(Ke));
* u by numpy.linalg.solve on the free DOFs; compliance, element energy and
  sensitivity (src/Functions/compute_element_energy.jl:18-38);
* the density-filter matrix from its node-based closed form (each node within
  rmin of the cell centre weighted by (rmin - d) and by the cells touching it,
  src/CheqFilters/CheqFilters.jl:66-118, density_filter.jl:49-108).

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import itertools
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def corners(dim):
    if dim == 2:
        return np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    return np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def ke_quadrature(dim, E=1.0, nu=0.3, h=None):
    h = np.ones(dim) if h is None else np.asarray(h, float)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((dim,) * 4)
    for i, j, k, l in itertools.product(range(dim), repeat=4):
        C[i, j, k, l] = lam * (i == j) * (k == l) + mu * ((i == k) * (j == l) + (i == l) * (j == k))
    xi, w = np.polynomial.legendre.leggauss(3)
    cn = corners(dim)
    s = 2 * cn - 1  # reference coordinates of the corners
    X = cn * h  # physical corner coordinates of one cell
    nb = cn.shape[0]
    K = np.zeros((dim * nb, dim * nb))
    for q in itertools.product(range(3), repeat=dim):
        p = xi[list(q)]
        wq = np.prod(w[list(q)])
        # dN/dxi
        dN = np.zeros((nb, dim))
        for a in range(nb):
            for d in range(dim):
                v = s[a, d] / 2 ** dim
                for e in range(dim):
                    if e != d:
                        v *= 1 + s[a, e] * p[e]
                dN[a, d] = v
        J = X.T @ dN  # dx/dxi
        G = dN @ np.linalg.inv(J)  # dN/dx
        dV = np.linalg.det(J) * wq
        for a in range(nb):
            for b in range(nb):
                blk = np.einsum("i,ijkl,l->jk", G[a], C, G[b])
                K[dim * a:dim * a + dim, dim * b:dim * b + dim] += blk * dV
    return K


def mesh(nels):
    """Ferrite first-touch numbering, restated independently (dict based)."""
    dim = len(nels)
    NX = nels[0] + 1
    NY = nels[1] + 1
    cn = corners(dim).astype(int)
    node_id = {}
    cells = []
    rng = [range(n) for n in nels[::-1]]
    for idx in itertools.product(*rng):
        ijk = idx[::-1]
        nodes = []
        for c in cn:
            pt = tuple(int(ijk[d] + c[d]) for d in range(dim))
            nid = pt[0] + NX * pt[1] + (NX * NY * pt[2] if dim == 3 else 0)
            if nid not in node_id:
                node_id[nid] = len(node_id)
            nodes.append(nid)
        cells.append(nodes)
    nn = len(node_id)
    node_dofs = np.zeros((dim, nn), np.int64)
    for nid, r in node_id.items():
        for d in range(dim):
            node_dofs[d, nid] = dim * r + d + 1
    cell_dofs = np.array([[node_dofs[d, n] for n in nodes for d in range(dim)] for nodes in cells],
                         np.int64).T
    cell_nodes = np.array(cells, np.int64).T
    return node_dofs, cell_dofs, cell_nodes


def cantilever(nels, load="reference"):
    dim = len(nels)
    node_dofs, cell_dofs, cell_nodes = mesh(nels)
    NX, NY = nels[0] + 1, nels[1] + 1
    nn = node_dofs.shape[1]
    fixed = np.zeros(node_dofs.size, bool)
    for n in range(nn):
        if n % NX == 0:
            fixed[node_dofs[:, n] - 1] = True
    k = nels[2] // 2 if (dim == 3 and load == "centre") else 0
    ln = nels[0] + NX * (nels[1] // 2 + NY * k)
    f = np.zeros(node_dofs.size)
    f[node_dofs[1, ln] - 1] = -1.0
    return node_dofs, cell_dofs, cell_nodes, fixed, f


def assemble(Ke, cell_dofs, fixed, px):
    nd = fixed.size
    K = np.zeros((nd, nd))
    for e in range(cell_dofs.shape[1]):
        dofs = cell_dofs[:, e] - 1
        mask = ~fixed[dofs]
        Kb = Ke * np.outer(mask, mask)
        K[np.ix_(dofs, dofs)] += px[e] * Kb
    md = cell_dofs.shape[1] * np.trace(Ke)
    K[fixed, :] = 0.0
    K[:, fixed] = 0.0
    K[fixed, fixed] = md
    return K, md


def filter_matrix(nels, rmin):
    dim = len(nels)
    NX, NY = nels[0] + 1, nels[1] + 1
    nel = int(np.prod(nels))
    cells = list(itertools.product(*[range(n) for n in nels[::-1]]))
    cells = [c[::-1] for c in cells]
    ncells_of = {}
    for ci, c in enumerate(cells):
        for off in itertools.product((0, 1), repeat=dim):
            node = tuple(c[d] + off[d] for d in range(dim))
            ncells_of.setdefault(node, []).append(ci)
    J = np.zeros((nel, nel))
    R = int(np.ceil(rmin)) + 1
    for i, c in enumerate(cells):
        ctr = np.array(c, float) + 0.5
        num = {}
        den = 0.0
        for off in itertools.product(range(-R, R + 2), repeat=dim):
            node = tuple(c[d] + off[d] for d in range(dim))
            if node not in ncells_of:
                continue
            dist = np.linalg.norm(np.array(node, float) - ctr)
            if dist < rmin:
                w = rmin - dist
                den += len(ncells_of[node]) * w
                for cj in ncells_of[node]:
                    num[cj] = num.get(cj, 0.0) + w
        for cj, v in num.items():
            J[i, cj] = v / den
    return J


def rho_drho(x, p, xmin):
    return xmin + (1 - xmin) * x ** p, (1 - xmin) * p * x ** (p - 1)


def case(nels, seed, rmins, p=3.0, xmin=1e-3):
    node_dofs, cell_dofs, cell_nodes, fixed, f = cantilever(nels)
    dim = len(nels)
    Ke = ke_quadrature(dim)
    g = np.random.default_rng(seed)
    nel = cell_dofs.shape[1]
    x = g.uniform(0.05, 1.0, nel)
    px, dpx = rho_drho(x, p, xmin)
    K, md = assemble(Ke, cell_dofs, fixed, px)
    v = g.standard_normal(fixed.size)
    y = K @ v
    free = ~fixed
    u = np.zeros(fixed.size)
    u[free] = np.linalg.solve(K[np.ix_(free, free)], f[free])
    ce = np.array([u[cell_dofs[:, e] - 1] @ Ke @ u[cell_dofs[:, e] - 1] for e in range(nel)])
    out = dict(nels=np.array(nels), Ke=Ke, node_dofs=node_dofs, cell_dofs=cell_dofs,
               fixed=fixed.astype(np.uint8), f=f, x=x, p=p, xmin=xmin, meandiag=md, v=v, Kv=y,
               u=u, compliance=f @ u, ce=ce, grad=-dpx * ce, obj=float(px @ ce))
    xf = g.uniform(0, 1, nel)
    out["xf"] = xf
    for r in rmins:
        Jm = filter_matrix(nels, r)
        key = str(r).replace(".", "p")
        out[f"J_{key}"] = Jm
        out[f"Jx_{key}"] = Jm @ xf
    return out


def main():
    cases = {
        "cantilever2d_6x4": case((6, 4), 0, [1.5, 2.0, 3.0]),
        "cantilever3d_3x2x2": case((3, 2, 2), 1, [1.5, 2.0]),
        "cantilever3d_4x2x4": case((4, 2, 4), 2, [1.5, 2.5]),
    }
    for name, d in cases.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print("wrote", name)
    # Ke of a non-cubic cell (h = 1, 2, 0.5) to pin the cell-size scaling
    np.savez_compressed(os.path.join(HERE, "ke_boxes.npz"),
                        ke2_h=ke_quadrature(2, 2.0, 0.25, [1.0, 2.0]),
                        ke3_h=ke_quadrature(3, 2.0, 0.25, [1.0, 2.0, 0.5]))


if __name__ == "__main__":
    main()
