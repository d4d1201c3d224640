"""BESO (TopOpt.jl src/Algorithms/beso.jl) on the device.

``BESO(comp, vol, vol_limit, filter; maxiter, tol, p, er, sens_tol, k)`` and
its call ``beso(x0; black, white) -> BESOResult`` mirror the reference
(beso.jl:45-188).  The whole loop (filter, solve, compliance, pullback,
sensitivity averaging, threshold bisection) runs in ``tohip_beso_run`` with
the design resident on the GPU; only the objective crosses to the host each
iteration.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import BesoOpts, BesoResultC, check, dptr
from .functions import ComplianceFun, DensityFilterFun, VolumeFun


@dataclass
class BESOResult:
    """BESOResult (beso.jl:1-7) plus the final volume fraction."""
    topology: np.ndarray
    objval: float
    change: float
    converged: bool
    fevals: int
    volume: float
    obj_trace: np.ndarray  # objective of every iteration


class BESO:
    """BESO(comp, vol, vol_limit, filter; maxiter=200, tol=1e-4, p=3.0,
    er=0.02, sens_tol=tol/100, k=10) (beso.jl:45-87).  Setting p changes the
    solver's penalty exponent, as the reference's setpenalty! does."""

    def __init__(self, comp, vol, vol_limit, filter, *, maxiter=200, tol=1e-4, p=3.0, er=0.02,
                 sens_tol=None, k=10, rel_tol=1e-8, pcg_maxiter=200000, precision=None):
        if not isinstance(comp, ComplianceFun) or not isinstance(vol, VolumeFun):
            raise TypeError("BESO needs a ComplianceFun and a VolumeFun")
        if not isinstance(filter, DensityFilterFun) or filter.solver is not comp.solver:
            raise TypeError("BESO's filter must be a DensityFilterFun on the compliance's solver")
        self.comp, self.vol, self.filter = comp, vol, filter
        self.vol_limit = float(vol_limit)
        self.maxiter, self.tol, self.er, self.k = int(maxiter), float(tol), float(er), int(k)
        self.sens_tol = self.tol / 100 if sens_tol is None else float(sens_tol)
        self.rel_tol, self.pcg_maxiter = float(rel_tol), int(pcg_maxiter)
        self.precision = precision
        self.penalty = comp.solver.penalty
        self.setpenalty(p)
        n = comp.solver.problem.ncells
        self.topology = np.zeros(n)
        self.vars = np.zeros(n)

    def setpenalty(self, p):
        """setpenalty!(b::BESO, p) (beso.jl:89-92)."""
        self.penalty.p = float(p)
        return self

    def __call__(self, x0=None, black=None, white=None) -> BESOResult:
        s = self.comp.solver
        prob = s.problem
        n = prob.ncells
        x0 = np.array(s.vars if x0 is None else x0, np.float64)
        passive = None
        for name, m in (("black", black), ("white", white)):
            if m is not None and len(m) not in (0, n):
                raise ValueError(f"{name} must have length {n} (got {len(m)})")
        b = np.zeros(n, bool) if black is None or len(black) == 0 else np.asarray(black, bool)
        w = np.zeros(n, bool) if white is None or len(white) == 0 else np.asarray(white, bool)
        if np.any(b & w):
            raise ValueError("elements cannot be both black and white")
        if b.any() or w.any():
            passive = np.ascontiguousarray(b.astype(np.uint8) + 2 * w.astype(np.uint8))
        s._sync_penalty()
        s._sync_conv()
        f = np.ascontiguousarray(prob.f, np.float64)
        check(s.lib.tohip_set_load(s.ctx.ptr, dptr(f)), s.ctx.ptr, "tohip_set_load")
        prec = s.precision if self.precision is None else self.precision
        o = BesoOpts(maxiter=self.maxiter, vol_limit=self.vol_limit, rmin=self.filter.rmin,
                     er=self.er, tol=self.tol, sens_tol=self.sens_tol, k=self.k,
                     penalty_p=float(self.penalty.p), rel_tol=self.rel_tol,
                     pcg_maxiter=self.pcg_maxiter, precision=_lib.PRECISIONS[prec])
        topo = np.empty(n)
        res = BesoResultC()
        hist = np.zeros(max(self.maxiter, 1))
        rc = s.lib.tohip_beso_run(s.ctx.ptr, C.byref(o), dptr(x0),
                                  None if passive is None else passive.ctypes.data_as(_lib._u8p),
                                  dptr(topo), C.byref(res), dptr(hist))
        s._pull_conv()
        check(rc, s.ctx.ptr, "tohip_beso_run")
        self.topology[:] = topo
        self.vars[:] = topo
        s.vars = topo.copy()
        return BESOResult(topology=topo, objval=res.objval, change=res.change,
                          converged=bool(res.converged), fevals=int(res.fevals),
                          volume=res.volume, obj_trace=hist[:res.fevals])
