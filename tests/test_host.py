"""CPU tests of the product's host side: the C-ABI library loads and exports
every declared entry point, and the host mirror of TopOpt's problem setup
(Ferrite numbering, Dirichlet set, load) matches the oracle.  No kernel is
launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "topopt_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(tohip_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_abi():
    names = declared_functions()
    for must in ("tohip_create", "tohip_destroy", "tohip_matvec", "tohip_solve",
                 "tohip_compliance", "tohip_filter_setup", "tohip_filter_apply",
                 "tohip_oc_update", "tohip_simp_run", "tohip_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(T):
    lib = T.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    from topopt_jl_amd._lib import SIGNATURES
    assert set(SIGNATURES) == set(declared_functions())


def test_library_is_gfx950_and_missing_library_fails_loudly(T, tmp_path):
    lib = T.load_library()
    assert b"gfx950" in lib.tohip_version()
    with pytest.raises(T.TohipError, match="no CPU fallback"):
        T.load_library(str(tmp_path / "missing.so"))


@pytest.mark.parametrize("nels", [(6, 4), (160, 40), (4, 2, 2), (5, 4, 6)])
def test_host_numbering_matches_oracle(T, O, nels):
    nd, cd = T.ferrite_numbering(nels)
    nd2, cd2, _ = O.numbering(nels)
    assert np.array_equal(nd, nd2) and np.array_equal(cd, cd2)


@pytest.mark.parametrize("nels,load", [((160, 40), "reference"), ((6, 4, 4), "reference"),
                                       ((6, 4, 4), "centre")])
def test_host_problem_matches_oracle(T, O, nels, load):
    p = T.PointLoadCantilever(nels, load=load)
    q = O.Cantilever(nels, load=load)
    assert p.force_dof == q.force_dof
    assert np.array_equal(p.fixed, q.fixed) and np.array_equal(p.f, q.f)
    assert p.ndofs == q.ndof and p.ncells == q.nel


def test_host_problem_errors(T):
    with pytest.raises(ValueError, match="even number"):
        T.PointLoadCantilever((10, 5))
    with pytest.raises(ValueError):
        T.PointLoadCantilever((4, 2, 2), load="middle")


def test_feasolver_rejects_other_solver_types(T):
    class Other:
        pass
    with pytest.raises(TypeError):
        T.FEASolver(Other, T.PointLoadCantilever((4, 2)))


def test_simp_opts_struct_layout():
    import sys
    sys.path.insert(0, ROOT)
    import _pkg
    _pkg.load_package()
    from topopt_jl_amd._lib import SimpOpts
    # int64 + 5 doubles + int64 + 2 ints + double, natural alignment
    assert C.sizeof(SimpOpts) == 8 + 5 * 8 + 8 + 4 + 4 + 8


def test_julia_glue_binds_declared_symbols():
    """The Julia binding (no Julia runtime here) only ccalls entry points the
    header declares, and its SimpOpts mirrors tohip_simp_opts field by field."""
    text = open(os.path.join(ROOT, "topopt.jl_amd", "julia", "TopOptHIP.jl")).read()
    called = set(re.findall(r"ccall\(\s*\(:(tohip_\w+),", text))
    assert called and called <= set(declared_functions())
    for must in ("tohip_create", "tohip_solve", "tohip_compliance", "tohip_simp_run"):
        assert must in called
    hdr = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    c_fields = re.findall(r"\b(\w+);", re.search(r"struct tohip_simp_opts\s*\{(.*?)\}", hdr,
                                                  re.S).group(1))
    jl = re.search(r"struct SimpOpts\n(.*?)\nend", text, re.S).group(1)
    jl_fields = re.findall(r"^\s*(\w+)::", jl, flags=re.M)
    assert jl_fields == c_fields
