"""Differentiable functions of the SIMP loop, backed by libtopopt_hip.

Mirrors TopOpt.jl:
  PseudoDensities                 src/TopOpt.jl:30-76
  ComplianceFun (+ rrule)         src/Functions/compliance.jl:16-76
  VolumeFun (+ rrule)             src/Functions/volume.jl:13-80
  DensityFilterFun (+ rrule)      src/CheqFilters/density_filter.jl:14-47
  ProjectedDensityFilterFun       src/CheqFilters/density_filter.jl:130-169
The rrules are exposed as ``pullback(Δ)`` methods (the product with the
stored Jacobian), since there is no AD system on this side.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, dptr


class PseudoDensities:
    """Wrapper marking a vector as (filtered) pseudo-densities."""

    def __init__(self, x, filtered=False):
        self.x = np.ascontiguousarray(x, np.float64)
        self.filtered = filtered

    def __len__(self):
        return self.x.size


def _as_x(x):
    return x.x if isinstance(x, PseudoDensities) else np.ascontiguousarray(x, np.float64)


class ComplianceFun:
    """J = sum_e rho_e u_e^T Ke u_e with gradient -drho_e u_e^T Ke u_e."""

    def __init__(self, solver):
        self.solver = solver
        n = solver.problem.ncells
        self.cell_comp = np.zeros(n)
        self.grad = np.zeros(n)

    def __call__(self, x, initially_zero=True) -> float:
        """solver.vars .= x; solver(); compute_compliance (compliance.jl:58-67)
        with the displacement kept on the device between the solve and the
        element energies (no download and re-upload of u)."""
        s = self.solver
        s.vars = _as_x(x).copy()
        s.solve_on_device(initially_zero=initially_zero)
        obj = C.c_double()
        u = None if not s.assembly else dptr(np.ascontiguousarray(s.u))
        check(s.lib.tohip_compliance(s.ctx.ptr, u, C.byref(obj), dptr(self.cell_comp),
                                     dptr(self.grad)),
              s.ctx.ptr, "tohip_compliance")
        return obj.value

    def pullback(self, delta=1.0):
        """rrule(ComplianceFun) pullback: grad * Δ."""
        return self.grad * delta

    def getpenalty(self):
        """getpenalty(::ComplianceFun) (src/Functions/compliance.jl:23): the solver's penalty."""
        return self.solver.getpenalty()

    def setpenalty(self, p):
        """setpenalty!(::ComplianceFun, p) (src/Functions/compliance.jl:24)."""
        self.solver.setpenalty(p)
        return self


class VolumeFun:
    """V = sum_e x_e V_e (/ V_total when fraction=True)."""

    def __init__(self, solver, fraction=True):
        self.solver = solver
        self.fraction = fraction
        n = solver.problem.ncells
        cellvol = float(np.prod(solver.problem.sizes))
        self.total_volume = cellvol * n
        self.grad = np.full(n, cellvol / self.total_volume if fraction else cellvol)

    def __call__(self, x) -> float:
        s = self.solver
        v = C.c_double()
        xx = np.ascontiguousarray(_as_x(x))
        check(s.lib.tohip_volume(s.ctx.ptr, dptr(xx), C.byref(v)), s.ctx.ptr, "tohip_volume")
        return v.value if self.fraction else v.value * self.total_volume

    def pullback(self, delta=1.0):
        return self.grad * delta


class DensityFilterFun:
    """x~ = J x with J from the reference's neighbour search (rmin)."""

    def __init__(self, solver, rmin):
        self.solver = solver
        self.rmin = float(rmin)
        check(solver.lib.tohip_filter_setup(solver.ctx.ptr, self.rmin), solver.ctx.ptr,
              "DensityFilterFun")

    def _apply(self, x, transpose):
        s = self.solver
        check(s.lib.tohip_filter_setup(s.ctx.ptr, self.rmin), s.ctx.ptr, "DensityFilterFun")
        xx = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(xx)
        check(s.lib.tohip_filter_apply(s.ctx.ptr, dptr(xx), dptr(y), int(transpose)), s.ctx.ptr,
              "tohip_filter_apply")
        return y

    def __call__(self, x) -> PseudoDensities:
        return PseudoDensities(self._apply(_as_x(x), False), filtered=True)

    def pullback(self, delta):
        """rrule(DensityFilterFun) pullback: J^T Δ."""
        return self._apply(_as_x(delta), True)


def _projection_spec(p):
    from .solver import AbstractProjection
    from . import _lib
    if p is None:
        return _lib.TOHIP_PROJ_NONE, 0.0
    if isinstance(p, AbstractProjection):
        return p.code, float(p.beta)
    raise TypeError("pre/post projections run on the device: use None, HeavisideProjectionFun "
                    "or SigmoidProjectionFun")


class ProjectedDensityFilterFun:
    """post(J pre(x)) (src/CheqFilters/density_filter.jl:130-169).

    ``ProjectedDensityFilterFun(filter, preproj, postproj)`` wraps a
    DensityFilterFun; ``ProjectedDensityFilterFun(solver, rmin=..., projection=...)``
    is the reference's convenience constructor (a new filter, no pre-projection,
    ``projection`` after it, HeavisideProjectionFun(10) by default)."""

    def __init__(self, filt, preproj=None, postproj=None, *, rmin=None, projection="default"):
        from .solver import GenericFEASolver, HeavisideProjectionFun
        if isinstance(filt, GenericFEASolver):
            if rmin is None:
                raise TypeError("ProjectedDensityFilterFun(solver; rmin, projection) needs rmin")
            filt = DensityFilterFun(filt, rmin)
            preproj = None
            postproj = HeavisideProjectionFun(10.0) if projection == "default" else projection
        elif not isinstance(filt, DensityFilterFun):
            raise TypeError("ProjectedDensityFilterFun wraps a DensityFilterFun or a solver")
        self.filter = filt
        self.preproj = preproj
        self.postproj = postproj
        self._pre = _projection_spec(preproj)
        self._post = _projection_spec(postproj)
        self._x = None

    def getdim(self):
        return self.filter.solver.problem.ncells

    def __call__(self, x) -> PseudoDensities:
        s = self.filter.solver
        check(s.lib.tohip_filter_setup(s.ctx.ptr, self.filter.rmin), s.ctx.ptr,
              "ProjectedDensityFilterFun")
        xx = np.ascontiguousarray(_as_x(x), np.float64)
        y = np.empty_like(xx)
        check(s.lib.tohip_filter_apply_projected(s.ctx.ptr, dptr(xx), dptr(y), *self._pre,
                                                 *self._post),
              s.ctx.ptr, "tohip_filter_apply_projected")
        self._x = xx.copy()
        return PseudoDensities(y, filtered=True)

    def pullback(self, delta, x=None):
        """Cotangent of x: pre'(x) .* J^T (post'(J pre(x)) .* delta), at the
        point of the last call (or at ``x``)."""
        s = self.filter.solver
        xx = self._x if x is None else np.ascontiguousarray(_as_x(x), np.float64)
        if xx is None:
            raise ValueError("pullback before the forward call")
        check(s.lib.tohip_filter_setup(s.ctx.ptr, self.filter.rmin), s.ctx.ptr,
              "ProjectedDensityFilterFun")
        d = np.ascontiguousarray(_as_x(delta), np.float64)
        out = np.empty_like(d)
        check(s.lib.tohip_filter_pullback_projected(s.ctx.ptr, dptr(xx), dptr(d), dptr(out),
                                                    *self._pre, *self._post),
              s.ctx.ptr, "tohip_filter_pullback_projected")
        return out


def oc_update(solver, x, dc, volfrac, move=0.2, tol=1e-12):
    """Build-defined OC step on the device; returns (x_new, lambda)."""
    xx = np.ascontiguousarray(x, np.float64)
    d = np.ascontiguousarray(dc, np.float64)
    xn = np.empty_like(xx)
    lam = C.c_double()
    check(solver.lib.tohip_oc_update(solver.ctx.ptr, dptr(xx), dptr(d), float(volfrac),
                                     float(move), float(tol), dptr(xn), C.byref(lam)),
          solver.ctx.ptr, "tohip_oc_update")
    return xn, lam.value
