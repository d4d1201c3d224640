"""Heat conduction (HeatTransfer physics) on the GPU.

Mirrors TopOpt.jl's ``HeatConductionProblem`` / ``HeatTree``
(src/TopOptProblems/problem_types.jl:783-1075), the matrix-free solver for it
(``FEASolver(CGMatrixFreeSolver, problem)``, src/FEA/solvers_api.jl:468-574),
``ThermalComplianceFun`` (src/Functions/thermal_compliance.jl) and
``TemperatureFun`` (src/Functions/temperature.jl).  Linear cells only.

The device solves with the prescribed temperatures applied as the direct
solver's ``apply!`` does: T = the prescribed value at prescribed DOFs and
K_ff T_f = Q_f - K_fp T_p on the free ones.
"""
from __future__ import annotations

import copy
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import TohipError, dptr, i64ptr, load_library, u8ptr
from .functions import PseudoDensities
from .problems import cell_nodes, ferrite_numbering
from .solver import AbstractPenalty, PowerPenaltyFun, _penalty_spec

# Ferrite generate_grid facet sets (Quadrilateral / Hexahedron): name -> (axis, side)
_FACETS = {
    2: {"left": (0, 0), "right": (0, 1), "bottom": (1, 0), "top": (1, 1)},
    3: {"left": (0, 0), "right": (0, 1), "front": (1, 0), "back": (1, 1), "bottom": (2, 0),
        "top": (2, 1)},
}


def _check(lib, rc, ptr, what):
    if rc < 0:
        msg = lib.tohip_heat_last_error(ptr)
        raise TohipError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    if rc == _lib.TOHIP_BREAKDOWN:  # NaN temperatures, with a warning
        import warnings
        msg = lib.tohip_heat_last_error(ptr)
        warnings.warn(f"{what}: {msg.decode() if msg else 'PCG breakdown'}", RuntimeWarning,
                      stacklevel=3)
    return rc


class HeatConductionProblem:
    """HeatConductionProblem(nels, sizes, k=1.0; Tleft=0.0, Tright=0.0,
    Ttop=None, Tbottom=None, heatflux={}, cload={}, Tfix={}).

    Dirichlet node sets: left / right are the x = min / max nodes, bottom / top
    the y = min / max nodes (src/TopOptProblems/grids.jl:111-114), in 3D too;
    None leaves a side free.  ``heatflux`` maps Ferrite facet-set names (2D:
    left, right, bottom, top; 3D: left, right, front, back, bottom, top) to a
    flux into the domain; ``cload`` maps 1-based node indices to point heat
    sources; ``Tfix`` maps 1-based node indices to prescribed temperatures.
    Where sets overlap, the later condition (left, right, top, bottom, Tfix
    order) sets the value."""

    physics = "HeatTransfer"

    def __init__(self, nels, sizes=None, k=1.0, *, celltype="Linear", Tleft=0.0, Tright=0.0,
                 Ttop=None, Tbottom=None, heatflux=None, cload=None, Tfix=None):
        nels = tuple(int(n) for n in nels)
        dim = len(nels)
        if dim not in (2, 3):
            raise ValueError("nels must have 2 or 3 entries")
        if celltype != "Linear":
            raise NotImplementedError("the device path supports linear cells (celltype='Linear')")
        self.nels, self.dim = nels, dim
        self.sizes = tuple(float(s) for s in (sizes if sizes is not None else (1.0,) * dim))
        self.k = float(k)
        self.node_dofs, self.cell_dofs = ferrite_numbering(nels, ncomp=1)
        self.nnodes = self.node_dofs.shape[1]
        self.ndofs = self.nnodes
        self.ncells = int(np.prod(nels))
        self.heatfluxdict = {str(a): float(b) for a, b in (heatflux or {}).items()}
        self.cloaddict = {int(a): float(b) for a, b in (cload or {}).items()}
        self.Tfix = {int(a): float(b) for a, b in (Tfix or {}).items()}
        N = [n + 1 for n in nels]
        ids = np.arange(self.nnodes)
        ix, iy = ids % N[0], (ids // N[0]) % N[1]
        dofs = self.node_dofs[0] - 1  # node id -> 0-based DOF
        self.fixed = np.zeros(self.ndofs, np.uint8)
        self.fixed_values = np.zeros(self.ndofs)
        for val, sel in ((Tleft, ix == 0), (Tright, ix == nels[0]), (Ttop, iy == nels[1]),
                         (Tbottom, iy == 0)):
            if val is not None:
                self.fixed[dofs[sel]] = 1
                self.fixed_values[dofs[sel]] = float(val)
        for node, val in self.Tfix.items():
            if not 1 <= node <= self.nnodes:
                raise ValueError(f"Tfix node {node} outside 1..{self.nnodes}")
            self.fixed[dofs[node - 1]] = 1
            self.fixed_values[dofs[node - 1]] = val
        self.prescribed_dofs = np.flatnonzero(self.fixed) + 1
        self.fixedload = self._fixedload(dofs)

    def _fixedload(self, dofs):
        """Q: point sources (make_cload) plus the facet flux integral
        fe_i = int phi_i q dGamma (_make_dloads): on a linear face each of its
        2^(dim-1) corners receives q * area / 2^(dim-1)."""
        Q = np.zeros(self.ndofs)
        for node, v in self.cloaddict.items():
            if not 1 <= node <= self.nnodes:
                raise ValueError(f"cload node {node} outside 1..{self.nnodes}")
            if v != 0:
                Q[dofs[node - 1]] += v
        facets = _FACETS[self.dim]
        N = [n + 1 for n in self.nels]
        axes = np.indices(tuple(N[::-1])).reshape(self.dim, -1)[::-1]  # node coords, x fastest
        for name, q in self.heatfluxdict.items():
            if name not in facets:
                raise KeyError(name)
            ax, side = facets[name]
            area = float(np.prod([self.sizes[d] for d in range(self.dim) if d != ax]))
            on = axes[ax] == (self.nels[ax] if side else 0)
            # face nodes count once per adjacent boundary face: 1 (corner), 2 (edge), 4 (interior)
            mult = np.ones(self.nnodes)
            for d in range(self.dim):
                if d != ax:
                    mult *= np.where((axes[d] == 0) | (axes[d] == self.nels[d]), 1.0, 2.0)
            share = q * area / 2 ** (self.dim - 1)
            Q[dofs[on]] += share * mult[on]
        return Q

    def __repr__(self):
        return f"HeatConductionProblem(nels={self.nels})"


def HeatTree(nels, sizes=None, k=1.0, *, celltype="Linear", q=1.0):
    """Flux q through the top facets, the bottom (y = min) nodes at T = 0,
    the sides insulated (src/TopOptProblems/problem_types.jl:1048-1075)."""
    return HeatConductionProblem(nels, sizes, k, celltype=celltype, Tleft=None, Tright=None,
                                 Ttop=None, Tbottom=0.0, heatflux={"top": float(q)})


class HeatFEASolver:
    """GenericFEASolver{HeatTransfer} on the GPU: ``vars`` the densities, ``u``
    the temperature (caller DOF order), conductivity rho(x) k."""

    def __init__(self, problem, *, xmin=None, penalty=None, cg_max_iter=700, abstol=None,
                 reltol=None, penalty_before_interpolation=True):
        self.problem = problem
        self.lib = load_library()
        self.xmin = 1.0 / 1000 if xmin is None else float(xmin)
        if penalty is None:
            penalty = PowerPenaltyFun(1.0)
        elif not isinstance(penalty, AbstractPenalty):
            raise TypeError(f"Unsupported penalty value {penalty!r}.")
        self.penalty = penalty
        self.prev_penalty = copy.deepcopy(penalty)
        self.cg_max_iter = int(cg_max_iter)
        self.abstol = 1e-7 if abstol is None else float(abstol)
        self.reltol = math.sqrt(np.finfo(np.float64).eps) if reltol is None else float(reltol)
        self.pbi = bool(penalty_before_interpolation)
        self.ptr = C.c_void_p()
        nels = np.ascontiguousarray(problem.nels, np.int64)
        h = np.ascontiguousarray(problem.sizes, np.float64)
        nd = np.ascontiguousarray(problem.node_dofs.reshape(-1), np.int64)
        fm = np.ascontiguousarray(problem.fixed, np.uint8)
        fv = np.ascontiguousarray(problem.fixed_values, np.float64)
        rc = self.lib.tohip_heat_create(C.byref(self.ptr), problem.dim, i64ptr(nels), dptr(h),
                                        problem.k, self.xmin, float(self.penalty.p),
                                        int(self.pbi), i64ptr(nd), u8ptr(fm), dptr(fv))
        _check(self.lib, rc, None, "tohip_heat_create")
        self.vars = np.ones(problem.ncells)
        self.u = np.zeros(problem.ndofs)
        self.lhs = np.zeros(problem.ndofs)
        self.rhs = np.zeros(problem.ndofs)
        self.last_iters = 0
        self.converged = True
        self._uploaded = None
        self._sync_penalty()

    def getpenalty(self):
        return self.penalty

    def setpenalty(self, p):
        if isinstance(p, bool) or not isinstance(p, (AbstractPenalty, int, float, np.floating,
                                                     np.integer)):
            raise TypeError(f"Unsupported penalty value {p!r}.")
        self.prev_penalty = copy.deepcopy(self.penalty)
        if isinstance(p, AbstractPenalty):
            self.penalty = p
        else:
            self.penalty.p = float(p)
        self._sync_penalty()

    def _sync_penalty(self):
        spec = _penalty_spec(self.penalty)
        if spec != self._uploaded:
            _check(self.lib, self.lib.tohip_heat_set_penalty_fun(self.ptr, *spec), self.ptr,
                   "tohip_heat_set_penalty_fun")
            self._uploaded = spec

    def set_densities(self, x=None):
        if x is not None:
            self.vars = np.ascontiguousarray(x, np.float64).copy()
        self._sync_penalty()
        v = np.ascontiguousarray(self.vars, np.float64)
        _check(self.lib, self.lib.tohip_heat_set_densities(self.ptr, dptr(v)), self.ptr,
               "tohip_heat_set_densities")

    def matvec(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        _check(self.lib, self.lib.tohip_heat_matvec(self.ptr, dptr(x), dptr(y)), self.ptr,
               "tohip_heat_matvec")
        return y

    def element_matrix(self):
        nc = 2 ** self.problem.dim
        ke = np.empty(nc * nc)
        _check(self.lib, self.lib.tohip_heat_element_matrix(self.ptr, dptr(ke)), self.ptr,
               "tohip_heat_element_matrix")
        return ke.reshape(nc, nc)

    def _solve(self, rhs, out, dirichlet, initially_zero=True, abstol=None, reltol=None,
               maxiter=None):
        self.set_densities()
        f = np.ascontiguousarray(rhs, np.float64)
        T = np.ascontiguousarray(out, np.float64).copy()
        iters = C.c_int64()
        res = C.c_double()
        rc = self.lib.tohip_heat_solve(self.ptr, dptr(f), dptr(T),
                                       self.abstol if abstol is None else float(abstol),
                                       self.reltol if reltol is None else float(reltol),
                                       self.cg_max_iter if maxiter is None else int(maxiter),
                                       int(bool(initially_zero)), int(bool(dirichlet)),
                                       C.byref(iters), C.byref(res))
        _check(self.lib, rc, self.ptr, "tohip_heat_solve")
        self.last_iters = iters.value
        self.converged = rc == _lib.TOHIP_OK
        return T

    def __call__(self, *, assemble_f=True, rhs=None, initially_zero=True, abstol=None,
                 reltol=None, maxiter=None):
        """solver(): the forward solve with the problem's heat load and the
        prescribed temperatures into ``u``; ``assemble_f=False`` solves
        K lhs = rhs with homogeneous conditions (apply_zero!) into ``lhs``."""
        if assemble_f and rhs is None:
            self.u = self._solve(self.problem.fixedload, self.u, True, initially_zero, abstol,
                                 reltol, maxiter)
        else:
            r = np.asarray(self.rhs if rhs is None else rhs, np.float64).copy()
            r[self.problem.fixed.astype(bool)] = 0.0
            self.lhs = self._solve(r, self.lhs, False, True, abstol, reltol, maxiter)
        return None

    def cell_products(self, lam, T, sign):
        lam = np.ascontiguousarray(lam, np.float64)
        T = np.ascontiguousarray(T, np.float64)
        cell = np.empty(self.problem.ncells)
        grad = np.empty(self.problem.ncells)
        _check(self.lib, self.lib.tohip_heat_cell_products(self.ptr, dptr(lam), dptr(T),
                                                           float(sign), dptr(cell), dptr(grad)),
               self.ptr, "tohip_heat_cell_products")
        return cell, grad

    def close(self):
        if self.ptr:
            self.lib.tohip_heat_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _densities(x):
    return np.asarray(x.x if isinstance(x, PseudoDensities) else x, np.float64)


class ThermalComplianceFun:
    """J = Q'T with the adjoint gradient dJ/dx_e = drho_e lam_e' Ke T_e,
    K_cond lam = -Q_cond (src/Functions/thermal_compliance.jl).  ``grad``
    holds the gradient of the last evaluation, ``cell_comp`` lam_e' Ke T_e."""

    def __init__(self, solver):
        if not isinstance(solver, HeatFEASolver):
            raise ValueError("ThermalComplianceFun can only be used with HeatTransferTopOptProblem.")
        self.solver = solver
        n = solver.problem.ncells
        self.cell_comp = np.zeros(n)
        self.grad = np.zeros(n)

    def __call__(self, x):
        s = self.solver
        s.vars = _densities(x).copy()
        s()
        Q = s.problem.fixedload
        obj = float(Q @ s.u)
        # adjoint: K_cond lam = -Q with Q zeroed on the prescribed DOFs
        s(assemble_f=False, rhs=-Q)
        self.cell_comp, self.grad = s.cell_products(s.lhs, s.u, 1.0)
        return obj

    def pullback(self, dy=1.0):
        return self.grad * float(dy)

    def getdim(self):
        return 1


class TemperatureFun:
    """T = tf(x): the temperature of the forward solve; ``pullback(D)`` gives
    dT/dx' D = -drho_e T_e' Ke lam_e with K lam = D (src/Functions/temperature.jl)."""

    def __init__(self, solver):
        if not isinstance(solver, HeatFEASolver):
            raise ValueError("TemperatureFun can only be used with HeatTransferTopOptProblem.")
        self.solver = solver
        self.fevals = 0

    def __call__(self, x):
        s = self.solver
        self.fevals += 1
        s.vars = _densities(x).copy()
        s()
        return s.u.copy()

    def pullback(self, D):
        s = self.solver
        s(assemble_f=False, rhs=np.asarray(D, np.float64))
        _, g = s.cell_products(s.lhs, s.u, -1.0)
        return g


def cell_temperature(T, problem):
    """Mean of the nodal temperatures over each cell's nodes."""
    T = np.asarray(T, np.float64)
    cn = cell_nodes(problem.nels)  # (nel, nc) node ids
    return T[problem.node_dofs[0][cn] - 1].mean(axis=1)
