"""ctypes binding of libtopopt_hip (include/topopt_hip.h).

This is the Python stand-in for the Julia ``ccall`` layer: it passes exactly
what TopOpt.jl holds (1-based Int64 ``node_dofs`` in column-major order,
Float64 vectors in Ferrite DOF order, cell vectors in cell order).  There is no
CPU fallback: if the shared library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtopopt_hip.so")

TOHIP_OK = 0
TOHIP_NOT_CONVERGED = 1
TOHIP_BREAKDOWN = 2
TOHIP_ERR_ARG = -1
TOHIP_ERR_HIP = -2
TOHIP_ERR_STATE = -3
TOHIP_ERR_DOMAIN = -4
TOHIP_CONV_DEFAULT, TOHIP_CONV_ENERGY = 0, 1
TOHIP_FP64 = 0
TOHIP_MIXED = 1
TOHIP_MIXED_MG = 2
TOHIP_FP64_MG = 3
PRECISIONS = {"fp64": TOHIP_FP64, "mixed": TOHIP_MIXED, "mixed_mg": TOHIP_MIXED_MG,
              "fp64_mg": TOHIP_FP64_MG}
TOHIP_PEN_POWER, TOHIP_PEN_RATIONAL, TOHIP_PEN_SINH = 0, 1, 2
TOHIP_PROJ_NONE, TOHIP_PROJ_HEAVISIDE, TOHIP_PROJ_SIGMOID = 0, 1, 2

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class SimpOpts(C.Structure):
    """Mirror of ``tohip_simp_opts``."""
    _fields_ = [
        ("iterations", C.c_int64),
        ("volfrac", C.c_double),
        ("rmin", C.c_double),
        ("move", C.c_double),
        ("oc_tol", C.c_double),
        ("rel_tol", C.c_double),
        ("maxiter", C.c_int64),
        ("precision", C.c_int),
        ("warm_start", C.c_int),
        ("inner_rel_tol", C.c_double),
    ]


class MmaOpts(C.Structure):
    """struct tohip_mma_opts."""
    _fields_ = [("s_init", C.c_double), ("s_incr", C.c_double), ("s_decr", C.c_double)]


class BesoOpts(C.Structure):
    """Mirror of ``tohip_beso_opts``."""
    _fields_ = [
        ("maxiter", C.c_int64),
        ("vol_limit", C.c_double),
        ("rmin", C.c_double),
        ("er", C.c_double),
        ("tol", C.c_double),
        ("sens_tol", C.c_double),
        ("k", C.c_int),
        ("penalty_p", C.c_double),
        ("rel_tol", C.c_double),
        ("pcg_maxiter", C.c_int64),
        ("precision", C.c_int),
    ]


class BesoResultC(C.Structure):
    """Mirror of ``tohip_beso_result``."""
    _fields_ = [
        ("objval", C.c_double),
        ("change", C.c_double),
        ("converged", C.c_int),
        ("fevals", C.c_int64),
        ("volume", C.c_double),
    ]


# name -> (restype, argtypes); keep in step with include/topopt_hip.h
SIGNATURES = {
    "tohip_version": (C.c_char_p, []),
    "tohip_last_error": (C.c_char_p, [C.c_void_p]),
    "tohip_device_info": (C.c_int, [C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "tohip_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, _i64p, _dp, C.c_double,
                               C.c_double, C.c_double, C.c_double, C.c_int, _i64p, _u8p]),
    "tohip_destroy": (None, [C.c_void_p]),
    "tohip_sizes": (C.c_int, [C.c_void_p, _i64p, _i64p]),
    "tohip_memory": (C.c_int, [C.c_void_p, _i64p, _i64p, _i64p]),
    "tohip_set_penalty": (C.c_int, [C.c_void_p, C.c_double]),
    "tohip_set_penalty_fun": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double]),
    "tohip_set_convergence": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "tohip_csr_assemble": (C.c_int, [C.c_void_p, C.c_int]),
    "tohip_csr_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_double)]),
    "tohip_csr_download": (C.c_int, [C.c_void_p, _i64p, C.POINTER(C.c_int), _dp]),
    "tohip_csr_matvec": (C.c_int, [C.c_void_p, _dp, _dp]),
    "tohip_csr_solve": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_int64,
                                  C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_double)]),
    "tohip_simp_run_mma": (C.c_int, [C.c_void_p, C.POINTER(SimpOpts), C.POINTER(MmaOpts), _dp,
                                     _dp, _i64p, _dp]),
    "tohip_heat_last_error": (C.c_char_p, [C.c_void_p]),
    "tohip_heat_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, _i64p, _dp, C.c_double,
                                    C.c_double, C.c_double, C.c_int, _i64p, _u8p, _dp]),
    "tohip_heat_destroy": (None, [C.c_void_p]),
    "tohip_heat_set_penalty_fun": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int,
                                             C.c_double]),
    "tohip_heat_set_densities": (C.c_int, [C.c_void_p, _dp]),
    "tohip_heat_matvec": (C.c_int, [C.c_void_p, _dp, _dp]),
    "tohip_heat_solve": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_int64,
                                   C.c_int, C.c_int, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_double)]),
    "tohip_heat_cell_products": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, _dp, _dp]),
    "tohip_heat_element_matrix": (C.c_int, [C.c_void_p, _dp]),
    "tohip_heat_bench_matvec": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "tohip_get_convergence_energy": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "tohip_mg_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "tohip_mg_weights": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                   C.POINTER(C.c_int)]),
    "tohip_filter_apply_projected": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                               C.c_double, C.c_int, C.c_double]),
    "tohip_filter_pullback_projected": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.c_int, C.c_double, C.c_int,
                                                  C.c_double]),
    "tohip_set_densities": (C.c_int, [C.c_void_p, _dp]),
    "tohip_matvec": (C.c_int, [C.c_void_p, _dp, _dp]),
    "tohip_solve": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_int64, C.c_int,
                              C.c_int, _i64p, _dp]),
    "tohip_compliance": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    "tohip_compliance_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_void_p]),
    "tohip_filter_setup": (C.c_int, [C.c_void_p, C.c_double]),
    "tohip_filter_apply": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int]),
    "tohip_volume": (C.c_int, [C.c_void_p, _dp, _dp]),
    "tohip_oc_update": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp,
                                  _dp]),
    "tohip_set_load": (C.c_int, [C.c_void_p, _dp]),
    "tohip_simp_run": (C.c_int, [C.c_void_p, C.POINTER(SimpOpts), _dp, _dp, _i64p, _dp]),
    "tohip_beso_run": (C.c_int, [C.c_void_p, C.POINTER(BesoOpts), _dp, _u8p, _dp,
                                 C.POINTER(BesoResultC), _dp]),
    "tohip_get_u": (C.c_int, [C.c_void_p, _dp]),
    "tohip_bench_kernel": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "tohip_get_xphys": (C.c_int, [C.c_void_p, _dp]),
    "tohip_nn_last_error": (C.c_char_p, [C.c_void_p]),
    "tohip_nn_create": (C.c_int, [C.POINTER(C.c_void_p), _i64p, C.c_int, C.c_int, C.c_uint64,
                                  _u8p]),
    "tohip_nn_destroy": (None, [C.c_void_p]),
    "tohip_nn_num_params": (C.c_int64, [C.c_void_p]),
    "tohip_nn_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "tohip_nn_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "tohip_nn_add_sample": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "tohip_nn_clear_samples": (C.c_int, [C.c_void_p]),
    "tohip_nn_num_samples": (C.c_int, [C.c_void_p]),
    "tohip_nn_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                 C.c_uint64, _dp]),
    "tohip_nn_eval": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "tohip_nn_predict": (C.c_int, [C.c_void_p, _dp, _dp, C.POINTER(C.c_float)]),
    "tohip_nn_attach": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tohip_nn_calls": (C.c_int, [C.c_void_p, _i64p]),
    "tohip_nn_conv_check": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_int] + [C.POINTER(C.c_float)] * 8),
}

# tohip_bench_kernel ids
KERNELS = {"matvec64": 0, "matvec32": 1, "pcg64": 2, "pcg32": 3, "update64": 4, "update32": 5,
           "filter": 6, "filter_t": 7, "energy": 8, "oc": 9, "copy64": 10, "mg_pre": 11,
           "mg_restrict": 12, "mg_coarse": 13, "mg_post": 14, "vcycle": 15, "csr_spmv": 16}

_lib = None


class TohipError(RuntimeError):
    """A hard error (< 0) returned through the C ABI."""


class TohipDomainError(TohipError, ArithmeticError):
    """TOHIP_ERR_DOMAIN: EnergyCriteria found a NaN or negative x'Ax (the
    reference throws DomainError, src/FEA/convergence_criteria.jl:35-41)."""


def load_library(path: str | None = None):
    """Load libtopopt_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise TohipError(f"libtopopt_hip.so not found at {p}; build it with "
                         "`python topopt.jl_amd/build.py` (there is no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def device_info():
    """(count, name) of the visible GPUs; (0, "") when there is none."""
    lib = load_library()
    n = C.c_int(0)
    buf = C.create_string_buffer(256)
    rc = lib.tohip_device_info(C.byref(n), buf, 256)
    return (n.value, buf.value.decode()) if rc == 0 else (0, "")


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def i64ptr(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i64p)


def u8ptr(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u8p)


def check(rc: int, ctx=None, what: str = "") -> int:
    """Map C return codes: < 0 raises TohipError; TOHIP_BREAKDOWN warns (the
    displacement is NaN, as the reference's direct solver leaves it with a
    warning, src/FEA/solvers_api.jl:152-158); other codes >= 0 are returned."""
    if rc < 0:
        msg = load_library().tohip_last_error(ctx)
        msg = msg.decode() if msg else ""
        cls = TohipDomainError if rc == TOHIP_ERR_DOMAIN else TohipError
        raise cls(f"{what} failed ({rc}): {msg}")
    if rc == TOHIP_BREAKDOWN:
        import warnings
        msg = load_library().tohip_last_error(ctx)
        warnings.warn(f"{what}: {msg.decode() if msg else 'PCG breakdown'}", RuntimeWarning,
                      stacklevel=3)
    return rc


class Context:
    """Owns one tohip_ctx (released by close() or garbage collection)."""

    def __init__(self, dim, nels, h, E, nu, xmin, p, pbi, node_dofs, fixed_mask):
        self.lib = load_library()
        self.ptr = C.c_void_p()
        nels_a = np.ascontiguousarray(nels, np.int64)
        h_a = np.ascontiguousarray(h, np.float64)
        nd = None if node_dofs is None else np.ascontiguousarray(
            np.asarray(node_dofs, np.int64).T.reshape(-1))  # column-major dim x nnodes
        fm = np.ascontiguousarray(fixed_mask, np.uint8)
        rc = self.lib.tohip_create(C.byref(self.ptr), int(dim), i64ptr(nels_a), dptr(h_a),
                                   float(E), float(nu), float(xmin), float(p), int(bool(pbi)),
                                   None if nd is None else i64ptr(nd), u8ptr(fm))
        if rc < 0:
            msg = self.lib.tohip_last_error(None)
            raise TohipError(f"tohip_create failed ({rc}): {msg.decode() if msg else ''}")
        nd_ = C.c_int64()
        nc_ = C.c_int64()
        self.lib.tohip_sizes(self.ptr, C.byref(nd_), C.byref(nc_))
        self.ndofs, self.ncells = nd_.value, nc_.value

    def close(self):
        if self.ptr:
            self.lib.tohip_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
