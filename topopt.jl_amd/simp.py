"""Device-resident SIMP loop (filter -> solve -> compliance -> pullback ->
design update).

One call runs many iterations with every vector resident in HBM; only the
per-iteration scalars come back.  It is the fused form of the loop the
reference drives from Julia (docs/tutorials/simp.qmd:86-114).  The design
update is the OC step this build defines, or MMA87 (``optimizer=MMA87()``),
the optimizer of the reference's tutorials.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import SimpOpts, check, dptr


@dataclass
class MMA87:
    """Method of moving asymptotes (Svanberg 1987) as the SIMP design update:
    asymptotes from x -+ s_init for two iterations, then widened by s_incr or
    narrowed by s_decr as the steps keep or flip sign."""
    s_init: float = 0.5
    s_incr: float = 1.2
    s_decr: float = 0.7


@dataclass
class SIMPResult:
    x: np.ndarray          # design after the last OC update
    obj: np.ndarray        # compliance per iteration
    pcg_iters: np.ndarray  # PCG iterations per solve (inner iterations when mixed)
    change: np.ndarray     # max |x_new - x| per iteration
    converged: bool        # every solve reached its tolerance
    nn_calls: int = 0      # network predictions used (warm_start="learned")


def simp_run(solver, *, volfrac, rmin, iterations, x0=None, move=0.2, oc_tol=1e-12,
             rel_tol=1e-8, maxiter=200000, precision=None, warm_start=True,
             inner_rel_tol=None, optimizer="oc") -> SIMPResult:
    """Run ``iterations`` SIMP iterations on the GPU.

    The solve stops at ||f - K u|| <= rel_tol * ||f||.  ``warm_start=True``
    starts each solve from the previous displacement (the reference's
    ``initially_zero=false``), False from zero (its default), ``"learned"``
    from the point of span{previous displacement, prediction of the network
    attached with ``solver.attach_warm_start``} with the smallest residual,
    until two iterations in a row leave more than 1/4 of the previous
    displacement's starting residual, or a solve from a prediction takes over
    twice the previous solve's PCG iterations (then the previous displacement
    alone; ``nn_calls`` of the result counts the predictions used).
    """
    prob = solver.problem
    lib = solver.lib
    solver._sync_penalty()
    solver._sync_conv()
    resident = isinstance(x0, str) and x0 == "device"
    if not resident:
        f = np.ascontiguousarray(prob.f, np.float64)
        check(lib.tohip_set_load(solver.ctx.ptr, dptr(f)), solver.ctx.ptr, "tohip_set_load")
        x = np.full(prob.ncells, volfrac) if x0 is None else np.array(x0, np.float64)
        x = np.ascontiguousarray(x)
    prec = solver.precision if precision is None else precision
    if inner_rel_tol is None:  # FP64 refresh after this residual drop (mixed paths)
        inner_rel_tol = 1e-3 if prec == "mixed_mg" else 0.1
    o = SimpOpts(iterations=int(iterations), volfrac=float(volfrac), rmin=float(rmin),
                 move=float(move), oc_tol=float(oc_tol), rel_tol=float(rel_tol),
                 maxiter=int(maxiter),
                 precision=_lib.PRECISIONS[prec],
                 warm_start=2 if warm_start == "learned" else (1 if warm_start else 0),
                 inner_rel_tol=float(inner_rel_tol))
    n = max(int(iterations), 1)
    obj = np.zeros(n)
    its = np.zeros(n, np.int64)
    chg = np.zeros(n)
    if optimizer == "mma":
        optimizer = MMA87()
    if isinstance(optimizer, MMA87):
        m = _lib.MmaOpts(s_init=float(optimizer.s_init), s_incr=float(optimizer.s_incr),
                         s_decr=float(optimizer.s_decr))
        rc = lib.tohip_simp_run_mma(solver.ctx.ptr, C.byref(o), C.byref(m),
                                    None if resident else dptr(x), dptr(obj),
                                    its.ctypes.data_as(C.POINTER(C.c_int64)), dptr(chg))
    elif optimizer == "oc":
        rc = lib.tohip_simp_run(solver.ctx.ptr, C.byref(o), None if resident else dptr(x),
                                dptr(obj), its.ctypes.data_as(C.POINTER(C.c_int64)), dptr(chg))
    else:
        raise ValueError("optimizer must be 'oc', 'mma' or an MMA87(...)")
    solver._pull_conv()
    check(rc, solver.ctx.ptr, "tohip_simp_run")
    k = int(iterations)
    calls = C.c_int64(0)
    if warm_start == "learned":
        check(lib.tohip_nn_calls(solver.ctx.ptr, C.byref(calls)), solver.ctx.ptr,
              "tohip_nn_calls")
    if resident:  # leave everything on the device (timed callers)
        return SIMPResult(x=None, obj=obj[:k], pcg_iters=its[:k], change=chg[:k],
                          converged=rc == _lib.TOHIP_OK, nn_calls=calls.value)
    u = np.empty(prob.ndofs)
    check(lib.tohip_get_u(solver.ctx.ptr, dptr(u)), solver.ctx.ptr, "tohip_get_u")
    solver.u = u
    return SIMPResult(x=x, obj=obj[:k], pcg_iters=its[:k], change=chg[:k],
                      converged=rc == _lib.TOHIP_OK, nn_calls=calls.value)
