"""TEST INFRASTRUCTURE ONLY — CPU oracle of the heat-conduction path.

numpy / scipy restatement of TopOpt.jl's HeatConductionProblem and its
solves, used only by tests/ as the checker:

* ``heat_ke``: the element conductivity matrix by Gauss quadrature of
  k grad(phi_a) . grad(phi_b) over the mapped reference cell
  (src/TopOptProblems/matrices_and_vectors.jl:451-496).
  Pinned: the unit-square closed form k/6 [4 -1 -2 -1; ...], symmetry,
  constant null space, quadrature-order independence.
* ``HeatProblem``: numbering by first touch over cells in order and Ferrite's
  corner order (src/TopOptProblems/metadata.jl:40-50 with one DOF per node);
  Dirichlet node sets left/right/bottom/top (src/TopOptProblems/grids.jl:
  111-114) and Tfix, applied in the reference's add! order
  (problem_types.jl:979-1022); the load Q = point sources (make_cload,
  matrices_and_vectors.jl:372-389) + facet flux integral by Gauss quadrature
  on each boundary face (_make_dloads, :288-338).  Pinned: total flux =
  q x face area; a 1D conduction problem with the exact linear solution.
* ``solve``: the direct solver's apply! semantics: T_p prescribed,
  K_ff T_f = Q_f - K_fp T_p (sparse LU).
* ``thermal_compliance``: J = Q'T and the adjoint gradient of
  src/Functions/thermal_compliance.jl (K_cond lam = -Q_cond, grad_e =
  drho_e lam_e' Ke T_e).  Pinned: the reference test facts (J = Q'T, J = 0
  without sources, T'KT > J + 1 with Tleft = 100) and central differences.
* ``temperature_pullback``: TemperatureFun's rrule (temperature.jl),
  pinned by central differences.

Facet-set names follow Ferrite's generate_grid (2D: left, right, bottom,
top; 3D: left, right, front, back, bottom, top); Ferrite is not vendored, so
the 3D names are parity unpinned.
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

CORNERS = {2: [(0, 0), (1, 0), (1, 1), (0, 1)],
           3: [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
               (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]}


def _gauss(n):
    return np.polynomial.legendre.leggauss(n)


def _shape_grads(dim, xi):
    """Gradients of the (bi/tri)linear shape functions at reference point xi
    in [-1, 1]^dim, shape (nc, dim)."""
    g = []
    for c in CORNERS[dim]:
        s = [2 * ci - 1 for ci in c]
        v = []
        for d in range(dim):
            t = 0.5 * s[d]
            for e in range(dim):
                if e != d:
                    t *= 0.5 * (1 + s[e] * xi[e])
            v.append(t)
        g.append(v)
    return np.array(g)


def heat_ke(dim, h, k=1.0, order=2):
    """Ke[a, b] = k int grad phi_a . grad phi_b over the box cell of sizes h."""
    pts, wts = _gauss(order)
    nc = 2 ** dim
    Ke = np.zeros((nc, nc))
    J = np.diag(np.asarray(h, float) / 2.0)  # reference -> physical
    Jinv = np.linalg.inv(J)
    detJ = np.linalg.det(J)
    for idx in itertools.product(range(order), repeat=dim):
        xi = [pts[i] for i in idx]
        w = np.prod([wts[i] for i in idx])
        G = _shape_grads(dim, xi) @ Jinv  # physical gradients
        Ke += k * w * detJ * (G @ G.T)
    return Ke


def _penalty(x, p, xmin, pbi=True):
    if pbi:
        return xmin + (1 - xmin) * x ** p, (1 - xmin) * p * x ** (p - 1)
    y = xmin + (1 - xmin) * x
    return y ** p, p * y ** (p - 1) * (1 - xmin)


class HeatProblem:
    def __init__(self, nels, sizes=None, k=1.0, Tleft=0.0, Tright=0.0, Ttop=None,
                 Tbottom=None, heatflux=None, cload=None, Tfix=None):
        self.nels = tuple(nels)
        dim = self.dim = len(nels)
        self.h = tuple(sizes) if sizes is not None else (1.0,) * dim
        self.k = k
        N = [n + 1 for n in nels]
        self.nnodes = int(np.prod(N))
        self.ncells = int(np.prod(nels))
        # cells x-fastest, corners in Ferrite order; first-touch numbering
        self.cells = []
        for c in range(self.ncells):
            i = c % nels[0]
            j = (c // nels[0]) % nels[1]
            kk = c // (nels[0] * nels[1]) if dim == 3 else 0
            nodes = []
            for o in CORNERS[dim]:
                n = (i + o[0]) + N[0] * ((j + o[1]) + (N[1] * (kk + o[2]) if dim == 3 else 0))
                nodes.append(n)
            self.cells.append(nodes)
        self.node_dof = -np.ones(self.nnodes, int)
        nxt = 0
        for nodes in self.cells:
            for n in nodes:
                if self.node_dof[n] < 0:
                    self.node_dof[n] = nxt
                    nxt += 1
        self.cell_dofs = np.array([[self.node_dof[n] for n in nodes] for nodes in self.cells])
        coords = np.array([[(n % N[0]), (n // N[0]) % N[1]] + ([n // (N[0] * N[1])] if dim == 3 else [])
                           for n in range(self.nnodes)], float)
        self.coords = coords * np.array(self.h)
        # Dirichlet (later conditions overwrite earlier ones)
        self.fixed = np.zeros(self.nnodes, bool)  # per DOF
        self.values = np.zeros(self.nnodes)
        L = [n * hh for n, hh in zip(nels, self.h)]
        sets = [(Tleft, np.isclose(self.coords[:, 0], 0.0)),
                (Tright, np.isclose(self.coords[:, 0], L[0])),
                (Ttop, np.isclose(self.coords[:, 1], L[1])),
                (Tbottom, np.isclose(self.coords[:, 1], 0.0))]
        for val, nodes in sets:
            if val is not None:
                d = self.node_dof[nodes]
                self.fixed[d] = True
                self.values[d] = val
        for node, val in (Tfix or {}).items():
            d = self.node_dof[node - 1]
            self.fixed[d] = True
            self.values[d] = val
        self.Ke = heat_ke(dim, self.h, k)
        self.Q = self._load(heatflux or {}, cload or {})

    def _load(self, heatflux, cload):
        Q = np.zeros(self.nnodes)
        for node, v in cload.items():
            if v != 0:
                Q[self.node_dof[node - 1]] += v
        if not heatflux:
            return Q
        dim = self.dim
        names = ({"left": (0, 0), "right": (0, 1), "bottom": (1, 0), "top": (1, 1)} if dim == 2
                 else {"left": (0, 0), "right": (0, 1), "front": (1, 0), "back": (1, 1),
                       "bottom": (2, 0), "top": (2, 1)})
        pts, wts = _gauss(2)
        for name, q in heatflux.items():
            ax, side = names[name]
            for c, nodes in enumerate(self.cells):
                idx = [c % self.nels[0], (c // self.nels[0]) % self.nels[1]] + \
                      ([c // (self.nels[0] * self.nels[1])] if dim == 3 else [])
                if idx[ax] != (self.nels[ax] - 1 if side else 0):
                    continue
                others = [d for d in range(dim) if d != ax]
                for qp in itertools.product(range(2), repeat=dim - 1):
                    xi = [0.0] * dim
                    xi[ax] = 1.0 if side else -1.0
                    w = 1.0
                    for d, qi in zip(others, qp):
                        xi[d] = pts[qi]
                        w *= wts[qi] * self.h[d] / 2.0
                    for a, o in enumerate(CORNERS[dim]):
                        phi = np.prod([0.5 * (1 + (2 * o[d] - 1) * xi[d]) for d in range(dim)])
                        Q[self.cell_dofs[c, a]] += phi * q * w
        return Q

    def assemble(self, rho):
        nc = 2 ** self.dim
        rows = np.repeat(self.cell_dofs, nc, axis=1).ravel()
        cols = np.tile(self.cell_dofs, (1, nc)).ravel()
        vals = (rho[:, None, None] * self.Ke[None]).reshape(-1)
        return sp.csr_matrix((vals, (rows, cols)), shape=(self.nnodes, self.nnodes))

    def _condensed_solve(self, K, rhs, values):
        f = self.fixed
        free = ~f
        T = np.zeros(self.nnodes)
        T[f] = values[f]
        b = rhs[free] - K[free][:, f] @ T[f]
        T[free] = spla.spsolve(K[free][:, free].tocsc(), b)
        return T

    def solve(self, x, p=1.0, xmin=1e-3, pbi=True):
        rho, _ = _penalty(np.asarray(x, float), p, xmin, pbi)
        K = self.assemble(rho)
        return self._condensed_solve(K, self.Q, self.values), K

    def thermal_compliance(self, x, p=1.0, xmin=1e-3, pbi=True):
        x = np.asarray(x, float)
        T, K = self.solve(x, p, xmin, pbi)
        J = float(self.Q @ T)
        lam = self._condensed_solve(K, -self.Q, np.zeros(self.nnodes))
        _, drho = _penalty(x, p, xmin, pbi)
        cell = np.einsum("ea,ab,eb->e", lam[self.cell_dofs], self.Ke, T[self.cell_dofs])
        return J, drho * cell, cell, T, K

    def temperature_pullback(self, x, D, p=1.0, xmin=1e-3, pbi=True):
        x = np.asarray(x, float)
        T, K = self.solve(x, p, xmin, pbi)
        Dz = np.asarray(D, float).copy()
        Dz[self.fixed] = 0.0
        lam = self._condensed_solve(K, Dz, np.zeros(self.nnodes))
        _, drho = _penalty(x, p, xmin, pbi)
        return -drho * np.einsum("ea,ab,eb->e", T[self.cell_dofs], self.Ke, lam[self.cell_dofs])
