"""save_mesh (src/TopOptProblems/IO/VTK.jl:15-53): the .vtu the reference writes
through WriteVTK — all grid nodes, the cells with density >= 0.5 in Ferrite's
(= VTK's) corner order, VTK_HEXAHEDRON / VTK_QUAD.  Read back and checked
against the grid; host-only (no GPU)."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from conftest import rng


def read_vtu(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arr = {}
    for da in piece.iter("DataArray"):
        arr[da.get("Name") or "points"] = np.array(da.text.split(), float)
    return int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), arr


@pytest.mark.parametrize("nels,sizes", [((6, 4, 4), (1.0, 1.0, 1.0)), ((8, 4), (0.5, 2.0))])
def test_save_mesh_writes_solid_cells(T, tmp_path, nels, sizes):
    prob = T.PointLoadCantilever(nels, sizes=sizes)
    x = rng(3).uniform(0, 1, prob.ncells)
    files = T.save_mesh(str(tmp_path / "design"), prob, x)
    assert files == [str(tmp_path / "design.vtu")]
    npts, ncells, a = read_vtu(files[0])
    dim = len(nels)
    assert npts == prob.nnodes and ncells == int(np.sum(x >= 0.5))
    pts = a["points"].reshape(-1, 3)
    N = [n + 1 for n in nels]
    # node id i + NX (j + NY k) sits at (i h0, j h1, k h2)
    for n in (0, 1, N[0], N[0] * N[1] - 1, prob.nnodes - 1):
        i, j = n % N[0], (n // N[0]) % N[1]
        k = n // (N[0] * N[1]) if dim == 3 else 0
        assert np.allclose(pts[n, :dim], [c * s for c, s in zip((i, j, k), sizes)])
    conn = a["connectivity"].astype(int).reshape(ncells, 2 ** dim)
    first = int(np.flatnonzero(x >= 0.5)[0])
    ci, cj = first % nels[0], (first // nels[0]) % nels[1]
    ck = first // (nels[0] * nels[1]) if dim == 3 else 0
    base = ci + N[0] * (cj + N[1] * ck)
    want = [base, base + 1, base + 1 + N[0], base + N[0]]
    if dim == 3:
        want += [v + N[0] * N[1] for v in want]
    assert list(conn[0]) == want
    assert np.all(a["types"] == (12 if dim == 3 else 9))
    assert np.array_equal(a["offsets"], 2 ** dim * np.arange(1, ncells + 1))


def test_save_mesh_checks_length_and_fields(T, tmp_path):
    prob = T.PointLoadCantilever((4, 2, 2))
    with pytest.raises(ValueError, match="number of cells"):
        T.save_mesh(str(tmp_path / "bad"), prob, np.ones(3))
    files = T.save_mesh(str(tmp_path / "all.vtu"), prob)  # no design: every cell
    npts, ncells, _ = read_vtu(files[0])
    assert ncells == prob.ncells
    u = np.arange(prob.ndofs, dtype=float)
    un = T.nodal_field(prob, u)
    assert un.shape == (prob.nnodes, 3)
    assert un[5, 1] == u[prob.node_dofs[1, 5] - 1]
    files = T.save_mesh(str(tmp_path / "u"), prob, point_data={"u": un},
                        cell_data={"rho": np.ones(prob.ncells)})
    _, _, a = read_vtu(files[0])
    assert np.array_equal(a["u"].reshape(-1, 3), un)
