"""FEA solver object backed by libtopopt_hip.

Mirrors TopOpt.jl's solver plugin API for this path:
  FEASolver(Solver, problem; xmin, penalty, cg_max_iter, abstol, ...)
      src/FEA/solvers_api.jl:468-574, 599-603
  (s::GenericFEASolver)(; initially_zero=true)   call operator :285-374
  solve_system!(::Type{CGMatrixFreeSolver}, ...)  :253-282 (-> cg_solve! :177-220)
``HIPMatrixFreeSolver`` plays the role of the new ``AbstractLinearSolver``
subtype; unlike the reference's call operator it never assembles K on the host.
"""
from __future__ import annotations

import copy
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import Context, check, dptr, i64ptr


class HIPMatrixFreeSolver:
    """Solver tag: matrix-free Jacobi PCG on the GPU (FP64 or mixed)."""


class HIPCGAssemblySolver:
    """Solver tag: CGAssemblySolver (src/FEA/solvers_api.jl:223-250) on the
    GPU: K(vars) assembled in CSR with apply!'s boundary rule, then CG with
    the identity preconditioner (the reference's default) or Jacobi
    (``preconditioner="jacobi"``).  FP64.  The comparator of the matrix-free
    operator."""


class AbstractPenalty:
    """SIMP penalty applied to the pseudo densities (src/Utilities/penalties.jl:3-10).
    ``kind`` is the library's TOHIP_PEN_* code; the device evaluates the
    family itself, ``__call__`` is the host-side convenience of the reference."""
    kind = _lib.TOHIP_PEN_POWER

    def __init__(self, p=1.0):
        self.p = float(p)

    def __repr__(self):
        return f"{type(self).__name__}({self.p})"


class PowerPenaltyFun(AbstractPenalty):
    """x -> x**p (src/Utilities/penalties.jl:30-33)."""
    kind = _lib.TOHIP_PEN_POWER

    def __call__(self, x):
        return np.asarray(x, np.float64) ** self.p


class RationalPenaltyFun(AbstractPenalty):
    """x -> x / (1 + p (1 - x)) (src/Utilities/penalties.jl:35-44)."""
    kind = _lib.TOHIP_PEN_RATIONAL

    def __call__(self, x):
        x = np.asarray(x, np.float64)
        return x / (1 + self.p * (1 - x))


class SinhPenaltyFun(AbstractPenalty):
    """x -> sinh(p x) / sinh(p) (src/Utilities/penalties.jl:46-54)."""
    kind = _lib.TOHIP_PEN_SINH

    def __call__(self, x):
        return np.sinh(self.p * np.asarray(x, np.float64)) / math.sinh(self.p)


class AbstractProjection:
    """Projection pushing densities toward 0/1 (src/Utilities/penalties.jl:11-17)."""
    code = _lib.TOHIP_PROJ_NONE

    def __init__(self, beta):
        self.beta = float(beta)

    def __repr__(self):
        return f"{type(self).__name__}({self.beta})"


class HeavisideProjectionFun(AbstractProjection):
    """x -> 1 - exp(-b x) + x exp(-b) (src/Utilities/penalties.jl:76-85)."""
    code = _lib.TOHIP_PROJ_HEAVISIDE

    def __call__(self, x):
        x = np.asarray(x, np.float64)
        return 1 - np.exp(-self.beta * x) + x * math.exp(-self.beta)


class SigmoidProjectionFun(AbstractProjection):
    """x -> 1 / (1 + exp((b + 1)(0.5 - x))) (src/Utilities/penalties.jl:87-95)."""
    code = _lib.TOHIP_PROJ_SIGMOID

    def __call__(self, x):
        x = np.asarray(x, np.float64)
        return 1 / (1 + np.exp((self.beta + 1) * (-x + 0.5)))


class ProjectedPenaltyFun(AbstractPenalty):
    """penalty(proj(x)) (src/Utilities/penalties.jl:56-69); the default
    projection is HeavisideProjectionFun(10); ``p`` forwards to the penalty."""

    def __init__(self, penalty, proj=None):
        if not isinstance(penalty, AbstractPenalty) or isinstance(penalty, ProjectedPenaltyFun):
            raise TypeError("ProjectedPenaltyFun wraps a PowerPenaltyFun, RationalPenaltyFun "
                            "or SinhPenaltyFun")
        self.penalty = penalty
        self.proj = HeavisideProjectionFun(10.0) if proj is None else proj
        if not isinstance(self.proj, AbstractProjection):
            raise TypeError("proj must be a HeavisideProjectionFun or SigmoidProjectionFun")

    @property
    def kind(self):
        return self.penalty.kind

    @property
    def p(self):
        return self.penalty.p

    @p.setter
    def p(self, value):
        self.penalty.p = float(value)

    def __call__(self, x):
        return self.penalty(self.proj(x))

    def __repr__(self):
        return f"ProjectedPenaltyFun({self.penalty!r}, {self.proj!r})"


def _penalty_spec(pen):
    """(kind, p, proj, beta) as tohip_set_penalty_fun takes them."""
    if isinstance(pen, ProjectedPenaltyFun):
        return (pen.kind, float(pen.p), pen.proj.code, float(pen.proj.beta))
    return (pen.kind, float(pen.p), _lib.TOHIP_PROJ_NONE, 0.0)


class ConvergenceCriteria:
    """CG stopping rule (src/FEA/convergence_criteria.jl:7)."""


class DefaultCriteria(ConvergenceCriteria):
    """Residual test ||r|| <= max(reltol ||r0||, abstol)."""


class EnergyCriteria(ConvergenceCriteria):
    """EnergyCriteria(energy=0.0) (src/FEA/convergence_criteria.jl:20-45):
    stop when |x'Ax - energy| / x'Ax <= tol and |x'r| / x'Ax <= tol, checked
    before every CG iteration; ``energy`` is updated by every check and
    carries over between solves.  tol is the solve's reltol.  FP64 solver."""

    def __init__(self, energy=0.0):
        self.energy = float(energy)


class GenericFEASolver:
    """Device-backed solver state (fields named as in the reference struct)."""

    def __init__(self, problem, *, xmin=None, penalty=None, cg_max_iter=700, abstol=None,
                 reltol=None, precision="fp64", penalty_before_interpolation=True,
                 conv=None, assembly=False, preconditioner=None, safe=False):
        if precision not in _lib.PRECISIONS:
            raise ValueError("precision must be 'fp64', 'mixed', 'mixed_mg' or 'fp64_mg'")
        if conv is None:
            conv = DefaultCriteria()
        elif not isinstance(conv, ConvergenceCriteria):
            raise TypeError(f"Unsupported convergence criterion {conv!r}.")
        self.conv = conv
        if preconditioner not in (None, "identity", "jacobi"):
            raise ValueError("preconditioner must be None / 'identity' or 'jacobi'")
        self.assembly = bool(assembly)
        if self.assembly and (precision != "fp64" or isinstance(conv, EnergyCriteria)):
            raise ValueError("the assembled-matrix solver runs FP64 CG with the residual rule")
        self.jacobi = preconditioner == "jacobi"
        self.safe = bool(safe)
        self.problem = problem
        self.xmin = 1.0 / 1000 if xmin is None else float(xmin)
        if penalty is None:
            penalty = PowerPenaltyFun(1.0)
        elif not isinstance(penalty, AbstractPenalty):
            raise TypeError(f"Unsupported penalty value {penalty!r}.")
        self.penalty = penalty
        self.prev_penalty = copy.deepcopy(penalty)
        self.cg_max_iter = int(cg_max_iter)
        self.abstol = 1e-7 if abstol is None else float(abstol)
        # IterativeSolvers' default relative tolerance sqrt(eps(Float64)); the
        # reference does not forward one.
        self.reltol = math.sqrt(np.finfo(np.float64).eps) if reltol is None else float(reltol)
        self.precision = precision
        self.pbi = bool(penalty_before_interpolation)
        self.ctx = Context(problem.dim, problem.nels, problem.sizes, problem.E, problem.nu,
                           self.xmin, self.penalty.p, self.pbi, problem.node_dofs, problem.fixed)
        self.vars = np.ones(problem.ncells)
        self._u = np.zeros(problem.ndofs)
        self._u_stale = False        # the device holds a newer displacement than _u
        self._u_dev_current = False  # the device displacement equals the latest u
        self._f_dev = None           # load last sent with tohip_set_load
        self.last_iters = 0
        self.last_resid = 0.0
        self.converged = True
        self._uploaded = (_lib.TOHIP_PEN_POWER, float(self.penalty.p), _lib.TOHIP_PROJ_NONE, 0.0)
        self._sync_penalty()

    # --- reference-style helpers -------------------------------------------
    @property
    def lib(self):
        return self.ctx.lib

    @property
    def u(self):
        """The displacement of the last solve (solver.u / lhs).  After a
        device-resident solve (ComplianceFun) it is fetched on first access."""
        if self._u_stale:
            u = np.empty(self.problem.ndofs)
            check(self.lib.tohip_get_u(self.ctx.ptr, dptr(u)), self.ctx.ptr, "tohip_get_u")
            self._u = u
            self._u_stale = False
        return self._u

    @u.setter
    def u(self, v):
        self._u = np.ascontiguousarray(v, np.float64)
        self._u_stale = False
        self._u_dev_current = False

    def getpenalty(self):
        return self.penalty

    def getprevpenalty(self):
        return self.prev_penalty

    def setpenalty(self, p):
        """setpenalty!(solver, p) (src/FEA/solvers_api.jl:419-429): a penalty
        object replaces the current one, a number changes its p in place."""
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
            check(self.lib.tohip_set_penalty_fun(self.ctx.ptr, *spec), self.ctx.ptr,
                  "tohip_set_penalty_fun")
            self._uploaded = spec

    def _sync_conv(self):
        """Upload the stopping rule and EnergyCriteria.energy (which the user
        may have set) before a solve."""
        energy = isinstance(self.conv, EnergyCriteria)
        check(self.lib.tohip_set_convergence(
            self.ctx.ptr, _lib.TOHIP_CONV_ENERGY if energy else _lib.TOHIP_CONV_DEFAULT,
            float(self.conv.energy) if energy else 0.0), self.ctx.ptr, "tohip_set_convergence")

    def _pull_conv(self):
        """EnergyCriteria.energy after the device's last check."""
        if isinstance(self.conv, EnergyCriteria):
            e = C.c_double()
            check(self.lib.tohip_get_convergence_energy(self.ctx.ptr, C.byref(e)), self.ctx.ptr,
                  "tohip_get_convergence_energy")
            self.conv.energy = e.value

    def set_densities(self, x=None):
        if x is not None:
            self.vars = np.ascontiguousarray(x, np.float64).copy()
        self._sync_penalty()
        v = np.ascontiguousarray(self.vars, np.float64)
        check(self.lib.tohip_set_densities(self.ctx.ptr, dptr(v)), self.ctx.ptr,
              "tohip_set_densities")

    def matvec(self, x):
        """K(vars) x, the MatrixFreeOperator product."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        check(self.lib.tohip_matvec(self.ctx.ptr, dptr(x), dptr(y)), self.ctx.ptr, "tohip_matvec")
        return y

    def __call__(self, *, initially_zero=True, rhs=None, abstol=None, reltol=None,
                 maxiter=None, precision=None):
        """Solve K(vars) u = f (f = the problem's load unless rhs is given).

        Returns None like the reference; the result is in ``self.u``.  A solve
        that hits maxiter leaves ``converged = False`` (the reference's CG
        returns silently in that case)."""
        self.set_densities()
        f = np.ascontiguousarray(self.problem.f if rhs is None else rhs, np.float64)
        if rhs is not None:  # apply_zero! on the prescribed DOFs
            f = f.copy()
            f[self.problem.fixed.astype(bool)] = 0.0
        u = np.ascontiguousarray(self.u, np.float64).copy()
        iters = C.c_int64()
        res = C.c_double()
        prec = self.precision if precision is None else precision
        if self.assembly:
            rc = self.lib.tohip_csr_solve(self.ctx.ptr, dptr(f), dptr(u),
                                          self.abstol if abstol is None else float(abstol),
                                          self.reltol if reltol is None else float(reltol),
                                          self.cg_max_iter if maxiter is None else int(maxiter),
                                          int(bool(initially_zero)), int(self.jacobi),
                                          int(self.safe), C.byref(iters), C.byref(res))
            check(rc, self.ctx.ptr, "tohip_csr_solve")
            self.u = u
            self.last_iters = iters.value
            self.last_resid = res.value
            self.converged = rc == _lib.TOHIP_OK
            return None
        self._sync_conv()
        rc = self.lib.tohip_solve(self.ctx.ptr, dptr(f), dptr(u),
                                  self.abstol if abstol is None else float(abstol),
                                  self.reltol if reltol is None else float(reltol),
                                  self.cg_max_iter if maxiter is None else int(maxiter),
                                  int(bool(initially_zero)),
                                  _lib.PRECISIONS[prec],
                                  C.byref(iters), C.byref(res))
        self._pull_conv()
        check(rc, self.ctx.ptr, "tohip_solve")
        self.u = u
        self._u_dev_current = True
        self.last_iters = iters.value
        self.last_resid = res.value
        self.converged = rc == _lib.TOHIP_OK
        return None

    def solve_on_device(self, *, initially_zero=True):
        """solver() for the problem's own load with the displacement kept on
        the device: the load is sent once (tohip_set_load), no displacement
        crosses PCIe, and ``u`` is fetched only when read.  Warm starts
        (initially_zero=False) continue from the device displacement unless
        ``u`` was assigned on the host since.  What ComplianceFun uses."""
        if self.assembly:  # the CSR solver takes host vectors
            return self(initially_zero=initially_zero)
        self.set_densities()
        f = self.problem.f
        if self._f_dev is None or not np.array_equal(self._f_dev, f):
            check(self.lib.tohip_set_load(self.ctx.ptr, dptr(np.ascontiguousarray(f, np.float64))),
                  self.ctx.ptr, "tohip_set_load")
            self._f_dev = np.array(f, np.float64)
        iters = C.c_int64()
        res = C.c_double()
        host_start = not initially_zero and not self._u_dev_current
        u = np.ascontiguousarray(self.u, np.float64).copy() if host_start else None
        self._sync_conv()
        rc = self.lib.tohip_solve(self.ctx.ptr, None, None if u is None else dptr(u),
                                  self.abstol, self.reltol, self.cg_max_iter,
                                  int(bool(initially_zero)), _lib.PRECISIONS[self.precision],
                                  C.byref(iters), C.byref(res))
        self._pull_conv()
        check(rc, self.ctx.ptr, "tohip_solve")
        if u is not None:
            self.u = u
        else:
            self._u_stale = True
        self._u_dev_current = True
        self.last_iters = iters.value
        self.last_resid = res.value
        self.converged = rc == _lib.TOHIP_OK
        return None

    def assembled_matrix(self):
        """globalinfo.K after assemble! and apply! (scipy CSR, caller DOF
        order), assembled on the device at the current densities."""
        import scipy.sparse as sp
        self.set_densities()
        check(self.lib.tohip_csr_assemble(self.ctx.ptr, int(self.safe)), self.ctx.ptr,
              "tohip_csr_assemble")
        nnz = C.c_int64()
        check(self.lib.tohip_csr_info(self.ctx.ptr, C.byref(nnz), None, None), self.ctx.ptr,
              "tohip_csr_info")
        n = self.problem.ndofs
        rowptr = np.empty(n + 1, np.int64)
        cols = np.empty(nnz.value, np.int32)
        vals = np.empty(nnz.value)
        check(self.lib.tohip_csr_download(self.ctx.ptr, i64ptr(rowptr),
                                          cols.ctypes.data_as(C.POINTER(C.c_int)), dptr(vals)),
              self.ctx.ptr, "tohip_csr_download")
        # internal DOF dim*node + d -> caller DOF node_dofs[d, node] - 1
        perm = (np.asarray(self.problem.node_dofs, np.int64).T.reshape(-1) - 1)
        rows = np.repeat(np.arange(n), np.diff(rowptr))
        return sp.csr_matrix((vals, (perm[rows], perm[cols])), shape=(n, n))

    def mg_apply(self, r, refresh=True):
        """z = M r with the multigrid preconditioner of precision "mixed_mg"
        (FP32 V-cycle at the current densities).  Each density upload
        refreshes the operator and the smoother weights' estimate;
        refresh=False applies the preconditioner as it stands (the same M for
        several vectors)."""
        if refresh:
            self.set_densities()
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty_like(r)
        check(self.lib.tohip_mg_apply(self.ctx.ptr, dptr(r), dptr(z)), self.ctx.ptr,
              "tohip_mg_apply")
        return z

    def mg_weights(self):
        """Damped-Jacobi weight and lambda_max(D^-1 K) estimate of every
        multigrid level (empty before the first multigrid solve)."""
        nl = C.c_int(0)
        w = np.zeros(16)
        lam = np.zeros(16)
        check(self.lib.tohip_mg_weights(self.ctx.ptr, dptr(w), dptr(lam), 16, C.byref(nl)),
              self.ctx.ptr, "tohip_mg_weights")
        return w[:nl.value].copy(), lam[:nl.value].copy()

    def memory(self):
        """Device bytes held by this solver (and its network) and the device's
        used / total bytes now."""
        v = [C.c_int64() for _ in range(3)]
        check(self.lib.tohip_memory(self.ctx.ptr, *[C.byref(a) for a in v]), self.ctx.ptr,
              "tohip_memory")
        return {"solver_bytes": v[0].value, "device_used_bytes": v[1].value,
                "device_total_bytes": v[2].value}

    def attach_warm_start(self, net):
        """Use a trained WarmStartNet on this solver's grid for
        ``simp_run(..., warm_start="learned")``; None detaches it."""
        check(self.lib.tohip_nn_attach(self.ctx.ptr, None if net is None else net.ptr),
              self.ctx.ptr, "tohip_nn_attach")
        self._net = net  # the context holds a raw pointer: keep the network alive

    def close(self):
        self.ctx.close()


def FEASolver(solver_type, problem, **kwargs) -> GenericFEASolver:
    """FEASolver(HIPMatrixFreeSolver, problem; xmin, penalty, cg_max_iter, abstol,
    reltol, precision, conv) -> GenericFEASolver."""
    if solver_type is HIPCGAssemblySolver:
        if getattr(problem, "physics", "LinearElasticity") != "LinearElasticity":
            raise ValueError("HIPCGAssemblySolver supports the elasticity problems")
        return GenericFEASolver(problem, assembly=True, **kwargs)
    if solver_type is not HIPMatrixFreeSolver:
        raise TypeError("solver_type must be HIPMatrixFreeSolver or HIPCGAssemblySolver")
    if getattr(problem, "physics", "LinearElasticity") == "HeatTransfer":
        # physics_type(::HeatTransferTopOptProblem) = HeatTransfer (solvers_api.jl:452)
        from .heat import HeatFEASolver
        if kwargs.pop("precision", "fp64") != "fp64":
            raise ValueError("the heat-conduction solver runs in FP64")
        if not isinstance(kwargs.pop("conv", None), (type(None), DefaultCriteria)):
            raise ValueError("the heat-conduction solver uses the residual stopping rule")
        return HeatFEASolver(problem, **kwargs)
    return GenericFEASolver(problem, **kwargs)
