"""Design output: ``save_mesh`` (TopOpt.jl src/TopOptProblems/IO/VTK.jl:15-53).

Writes the VTK XML unstructured grid (.vtu) the reference writes through
WriteVTK: every node of the structured grid, and the cells whose density is
at least 0.5 as hexahedra (VTK type 12) or quadrilaterals (9), with Ferrite's
corner order (which is VTK's).  The file is plain XML with ASCII data arrays
(WriteVTK's default is compressed binary; readers accept both).  Optional
``point_data`` / ``cell_data`` (name -> array) are added for inspection, for
example the displacement of a solve.
"""
from __future__ import annotations

import os

import numpy as np

from .functions import PseudoDensities

_VTK_HEXAHEDRON = 12
_VTK_QUAD = 9
# Ferrite's Quadrilateral / Hexahedron local node order (src/TopOptProblems/grids.jl:80-92)
_OFF2 = [(0, 0), (1, 0), (1, 1), (0, 1)]
_OFF3 = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def _fmt(a):
    return " ".join(repr(float(v)) if isinstance(v, (float, np.floating)) else str(int(v))
                    for v in np.ravel(a))


def save_mesh(filename, problem, vars=None, *, point_data=None, cell_data=None):
    """save_mesh(filename, problem[, vars]): vars is the density vector (a
    solver's ``vars`` or a design); without it every cell is written.
    Returns the list of files written, like WriteVTK's vtk_save."""
    nel = problem.ncells
    if vars is None:
        rho = np.ones(nel)
    else:
        rho = np.asarray(vars.x if isinstance(vars, PseudoDensities) else vars, np.float64)
    if rho.size != nel:
        raise ValueError(f"Density vector ρ must have length equal to number of cells ({nel})")
    nels = problem.nels
    dim = len(nels)
    h = problem.sizes
    N = [n + 1 for n in nels]
    # node coordinates, node id = i + NX (j + NY k) (Ferrite's generate_grid order)
    axes = [np.arange(N[d]) * h[d] for d in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    coords = np.stack([g.transpose(*reversed(range(dim))).ravel() for g in grids], 1)
    if dim == 2:
        coords = np.hstack([coords, np.zeros((coords.shape[0], 1))])
    e = np.flatnonzero(rho >= 0.5)
    i = e % nels[0]
    j = (e // nels[0]) % nels[1]
    k = e // (nels[0] * nels[1]) if dim == 3 else np.zeros_like(e)
    off = _OFF3 if dim == 3 else _OFF2
    conn = np.stack([(i + o[0]) + N[0] * ((j + o[1]) + N[1] * ((k + o[2]) if dim == 3 else 0))
                     for o in off], 1)
    npc = len(off)
    ctype = _VTK_HEXAHEDRON if dim == 3 else _VTK_QUAD
    path = filename if filename.endswith(".vtu") else filename + ".vtu"
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian">',
             '  <UnstructuredGrid>',
             f'    <Piece NumberOfPoints="{coords.shape[0]}" NumberOfCells="{e.size}">',
             '      <Points>',
             '        <DataArray type="Float64" NumberOfComponents="3" format="ascii">',
             _fmt(coords), '        </DataArray>', '      </Points>', '      <Cells>',
             '        <DataArray type="Int64" Name="connectivity" format="ascii">', _fmt(conn),
             '        </DataArray>',
             '        <DataArray type="Int64" Name="offsets" format="ascii">',
             _fmt(npc * np.arange(1, e.size + 1)), '        </DataArray>',
             '        <DataArray type="UInt8" Name="types" format="ascii">',
             _fmt(np.full(e.size, ctype)), '        </DataArray>', '      </Cells>']
    if point_data:
        lines.append('      <PointData>')
        for name, arr in point_data.items():
            a = np.asarray(arr, np.float64).reshape(coords.shape[0], -1)
            lines += [f'        <DataArray type="Float64" Name="{name}" '
                      f'NumberOfComponents="{a.shape[1]}" format="ascii">', _fmt(a),
                      '        </DataArray>']
        lines.append('      </PointData>')
    if cell_data:
        lines.append('      <CellData>')
        for name, arr in cell_data.items():
            a = np.asarray(arr, np.float64).reshape(nel, -1)[e]
            lines += [f'        <DataArray type="Float64" Name="{name}" '
                      f'NumberOfComponents="{a.shape[1]}" format="ascii">', _fmt(a),
                      '        </DataArray>']
        lines.append('      </CellData>')
    lines += ['    </Piece>', '  </UnstructuredGrid>', '</VTKFile>']
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return [path]


def nodal_field(problem, u):
    """A DOF vector in the caller's (Ferrite) order as an (nnodes, dim) array
    in node order, for ``point_data``."""
    u = np.asarray(u, np.float64)
    return u[problem.node_dofs.T - 1]
