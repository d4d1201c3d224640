"""Problem definition and FE metadata (host side, computed once).

Mirrors TopOpt.jl's ``PointLoadCantilever`` (src/TopOptProblems/problem_types.jl
:162-231, load node :398-402), ``RectilinearGrid``/``generate_grid``
(src/TopOptProblems/grids.jl:80-92) and ``Metadata`` (node_dofs, cell_dofs;
src/TopOptProblems/metadata.jl:40-50, :116-145).
"""
from __future__ import annotations

import numpy as np

# Ferrite local corner order (Quadrilateral / Hexahedron), as generate_grid
# builds cells: (i,j), (i+1,j), (i+1,j+1), (i,j+1) [, same at k+1].
_CORNERS = {
    2: np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.int64),
    3: np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.int64),
}


def cell_nodes(nels) -> np.ndarray:
    """0-based node ids (x-fastest) of every cell, shape (nel, 2**dim)."""
    dim = len(nels)
    ne = np.asarray(nels, np.int64)
    NX, NY = ne[0] + 1, ne[1] + 1
    idx = np.indices(tuple(ne[::-1])).reshape(dim, -1)[::-1]  # (dim, nel), x fastest
    corners = _CORNERS[dim]
    coords = idx[:, :, None] + corners.T[:, None, :]  # (dim, nel, nc)
    n = coords[0] + NX * coords[1]
    if dim == 3:
        n = n + NX * NY * coords[2]
    return n


def ferrite_numbering(nels, ncomp=None):
    """Cell-first-touch DOF numbering of Ferrite's DofHandler for a field
    with ``ncomp`` components per node (ip^dim by default, 1 for a scalar
    field such as the temperature).

    Returns (node_dofs[ncomp, nnodes], cell_dofs[ndpc, nel]) with 1-based DOF
    numbers, indexed like the Julia arrays (ncomp x nnodes, ndpc x nel)."""
    dim = len(nels) if ncomp is None else int(ncomp)
    cn = cell_nodes(nels)
    flat = cn.reshape(-1)
    nn = int(np.prod(np.asarray(nels) + 1))
    first = np.full(nn, np.iinfo(np.int64).max, np.int64)
    np.minimum.at(first, flat, np.arange(flat.size, dtype=np.int64))
    order = np.argsort(first, kind="stable")
    rank = np.empty(nn, np.int64)
    rank[order] = np.arange(nn, dtype=np.int64)
    node_dofs = dim * rank[None, :] + np.arange(dim, dtype=np.int64)[:, None] + 1
    cell_dofs = node_dofs[:, cn]  # (dim, nel, nc)
    cell_dofs = cell_dofs.transpose(2, 0, 1).reshape(dim * cn.shape[1], -1)  # local dim*a+d
    return node_dofs, cell_dofs


class PointLoadCantilever:
    """Cantilever clamped at x = 0 with a point load on the x = max face.

    PointLoadCantilever(nels, sizes, E=1.0, nu=0.3, force=1.0; load="reference")

    ``load="reference"`` places the load on the first node of the
    {x = max, y = mid} node set as TopOpt.jl does (in 3D that set is a line
    along z; its first node is at z = 0).  ``load="centre"`` places it at the
    centre of the x = max face.  The force is -force along y.
    """

    def __init__(self, nels, sizes=None, E=1.0, nu=0.3, force=1.0, load="reference"):
        nels = tuple(int(n) for n in nels)
        dim = len(nels)
        if dim not in (2, 3):
            raise ValueError("nels must have 2 or 3 entries")
        if nels[1] % 2 or (dim == 3 and nels[2] % 2):
            raise ValueError("Grid does not have an even number of elements along the y and/or z axes.")
        if load not in ("reference", "centre"):
            raise ValueError("load must be 'reference' or 'centre'")
        self.nels = nels
        self.dim = dim
        self.sizes = tuple(float(s) for s in (sizes if sizes is not None else (1.0,) * dim))
        self.E, self.nu, self.force = float(E), float(nu), float(force)
        self.load = load
        self.node_dofs, self.cell_dofs = ferrite_numbering(nels)
        self.nnodes = self.node_dofs.shape[1]
        self.ndofs = self.node_dofs.size
        self.ncells = int(np.prod(nels))
        NX, NY = nels[0] + 1, nels[1] + 1
        ids = np.arange(self.nnodes)
        fixed_nodes = ids[ids % NX == 0]  # Dirichlet on x = 0, all components
        self.fixed = np.zeros(self.ndofs, np.uint8)
        self.fixed[(self.node_dofs[:, fixed_nodes] - 1).ravel()] = 1
        self.prescribed_dofs = np.sort(self.node_dofs[:, fixed_nodes].ravel())
        k = nels[2] // 2 if (dim == 3 and load == "centre") else 0
        self.load_node = nels[0] + NX * (nels[1] // 2 + NY * k)  # 0-based node id
        self.force_dof = int(self.node_dofs[1, self.load_node])  # 1-based, y component
        self.f = np.zeros(self.ndofs)
        self.f[self.force_dof - 1] = -self.force

    def __repr__(self):
        return f"PointLoadCantilever(nels={self.nels}, load={self.load!r})"
