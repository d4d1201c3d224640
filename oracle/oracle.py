"""TEST INFRASTRUCTURE ONLY — Python front of the oracle.

ctypes bindings to ``liboracle.so`` (the plain-C restatement in ``oracle.c``)
plus the SIMP driver glue.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg import this module, as the checker and the
CPU baseline.  The product package never imports it.

Reference mapping (TopOpt.jl v0.14.0):
  Cantilever            PointLoadCantilever  src/TopOptProblems/problem_types.jl:162-231
  Oracle.matvec         mul!(y, ::MatrixFreeOperator, x)  src/FEA/matrix_free_operator.jl:66-105
  Oracle.pcg            cg_solve! / IterativeSolvers.cg!  src/FEA/solvers_api.jl:177-220
  Oracle.element_energy compute_element_energy  src/Functions/compute_element_energy.jl:18-38
  Oracle.filter_*       DensityFilterFun  src/CheqFilters/density_filter.jl:14-108
  Oracle.oc_update      build-defined OC (the reference has none)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

_lib = None


class _Op(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("ndpc", C.c_int),
        ("nel", C.c_int64), ("ndof", C.c_int64),
        ("cell_dofs", C.c_void_p), ("Ke", C.c_void_p), ("fixed", C.c_void_p),
        ("xmin", C.c_double), ("p", C.c_double), ("pbi", C.c_int),
        ("meandiag", C.c_double),
        ("dc_off", C.c_void_p), ("dc_val", C.c_void_p),
        ("xes", C.c_void_p), ("px", C.c_void_p),
        ("kind", C.c_int), ("proj", C.c_int), ("beta", C.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile oracle.c into liboracle.so (gcc, OpenMP, no FMA contraction)."""
    src = os.path.join(_HERE, "oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        import subprocess
        cmd = ["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
               src, "-o", LIB_PATH + ".tmp", "-lm"]
        subprocess.run(cmd, check=True)
        os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_nnodes.restype = C.c_int64
        L.orc_nnodes.argtypes = [C.c_int, _ip]
        L.orc_ncells.restype = C.c_int64
        L.orc_ncells.argtypes = [C.c_int, _ip]
        L.orc_numbering.restype = C.c_int64
        L.orc_numbering.argtypes = [C.c_int, _ip, _ip, _ip, _ip]
        L.orc_element_stiffness.restype = C.c_int
        L.orc_element_stiffness.argtypes = [C.c_int, C.c_double, C.c_double, _dp, C.c_int, _dp]
        L.orc_rho_drho.restype = None
        L.orc_rho_drho.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_rho_drho_fun.restype = None
        L.orc_rho_drho_fun.argtypes = [C.c_double, C.c_int, C.c_double, C.c_int, C.c_double,
                                       C.c_double, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        L.orc_project.restype = None
        L.orc_project.argtypes = [C.c_double, C.c_int, C.c_double, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
        L.orc_op_init.restype = C.c_int
        L.orc_op_init.argtypes = [C.POINTER(_Op)]
        L.orc_op_free.restype = None
        L.orc_op_free.argtypes = [C.POINTER(_Op)]
        L.orc_op_set_vars.restype = None
        L.orc_op_set_vars.argtypes = [C.POINTER(_Op), _dp]
        L.orc_matvec.restype = None
        L.orc_matvec.argtypes = [C.POINTER(_Op), _dp, _dp]
        L.orc_jacobi_diag.restype = None
        L.orc_jacobi_diag.argtypes = [C.POINTER(_Op), _dp]
        L.orc_pcg.restype = C.c_int64
        L.orc_pcg.argtypes = [C.POINTER(_Op), _dp, _dp, C.c_double, C.c_double, C.c_int64,
                              C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.orc_pcg_energy.restype = C.c_int64
        L.orc_pcg_energy.argtypes = [C.POINTER(_Op), _dp, _dp, C.c_double, C.c_int64, C.c_int,
                                     C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]
        L.orc_element_energy.restype = C.c_double
        L.orc_element_energy.argtypes = [C.POINTER(_Op), _dp, _dp, _dp, _dp]
        L.orc_filter_build.restype = C.c_int64
        L.orc_filter_build.argtypes = [C.c_int, _ip, _dp, C.c_double, C.c_void_p, C.c_void_p,
                                       C.c_void_p]
        L.orc_csr_mv.restype = None
        L.orc_csr_mv.argtypes = [C.c_int64, _ip, _ip, _dp, _dp, _dp]
        L.orc_csr_mtv.restype = None
        L.orc_csr_mtv.argtypes = [C.c_int64, C.c_int64, _ip, _ip, _dp, _dp, _dp]
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_cell_volume.restype = C.c_double
        L.orc_cell_volume.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_volume.restype = C.c_double
        L.orc_volume.argtypes = [C.c_int64, _dp, _dp, C.c_int]
        L.orc_oc_update.restype = C.c_double
        L.orc_oc_update.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double,
                                    C.c_double, _dp]
        L.orc_direct_solve.restype = C.c_int
        L.orc_direct_solve.argtypes = [C.POINTER(_Op), _dp, _dp]
        _lib = L
    return _lib


# ---------------------------------------------------------------- mesh ----

def numbering(nels):
    """Ferrite cell-first-touch numbering: (node_dofs dim x nnodes, cell_dofs
    ndpc x nel, cell_nodes nnpc x nel), all 1-based, as Julia arrays."""
    dim = len(nels)
    ne = np.asarray(nels, np.int64)
    nn = int(np.prod(ne + 1))
    nel = int(np.prod(ne))
    nnpc = 1 << dim
    node_dofs = np.zeros(dim * nn, np.int64)
    cell_dofs = np.zeros(dim * nnpc * nel, np.int64)
    cell_nodes = np.zeros(nnpc * nel, np.int64)
    lib().orc_numbering(dim, ne, node_dofs, cell_dofs, cell_nodes)
    return (node_dofs.reshape(nn, dim).T, cell_dofs.reshape(nel, dim * nnpc).T,
            cell_nodes.reshape(nel, nnpc).T)


def element_stiffness(dim, E=1.0, nu=0.3, h=None, quad_order=4):
    h = np.ones(dim) if h is None else np.asarray(h, np.float64)
    n = dim * (1 << dim)
    Ke = np.zeros(n * n)
    rc = lib().orc_element_stiffness(dim, E, nu, np.ascontiguousarray(h, np.float64),
                                     quad_order, Ke)
    if rc != 0:
        raise ValueError("bad element stiffness arguments")
    return Ke.reshape(n, n)


def rho_drho(x, p, xmin, pbi=True, kind=0, proj=0, beta=0.0):
    """get_rho_drho: rho and d rho / d x for penalty family ``kind`` (0 power,
    1 rational, 2 sinh) after projection ``proj`` (0 none, 1 Heaviside,
    2 sigmoid) of steepness ``beta``."""
    r, d = C.c_double(), C.c_double()
    lib().orc_rho_drho_fun(float(x), int(kind), float(p), int(proj), float(beta), float(xmin),
                           int(pbi), C.byref(r), C.byref(d))
    return r.value, d.value


def project(x, proj, beta):
    """Elementwise projection (0 none, 1 Heaviside, 2 sigmoid): values and
    derivatives (dual-number evaluation in oracle.c)."""
    x = np.asarray(x, np.float64)
    v, d = np.empty_like(x), np.empty_like(x)
    a, b = C.c_double(), C.c_double()
    for i, xi in enumerate(x.ravel()):
        lib().orc_project(float(xi), int(proj), float(beta), C.byref(a), C.byref(b))
        v.flat[i], d.flat[i] = a.value, b.value
    return v, d


@dataclass
class Cantilever:
    """PointLoadCantilever restated (problem_types.jl:162-231, :398-402).

    ``load`` = "reference": the first node of the {x=max, y=mid} node set
    (smallest node id, i.e. z=0 in 3D); "centre": the centre of the x=max face.
    """
    nels: tuple
    E: float = 1.0
    nu: float = 0.3
    force: float = 1.0
    load: str = "reference"

    def __post_init__(self):
        dim = len(self.nels)
        if self.nels[1] % 2 or (dim == 3 and self.nels[2] % 2):
            raise ValueError("Grid does not have an even number of elements along the y and/or z axes.")
        self.dim = dim
        self.node_dofs, self.cell_dofs, self.cell_nodes = numbering(self.nels)
        self.ndof = self.node_dofs.size
        self.nel = self.cell_dofs.shape[1]
        NX, NY = self.nels[0] + 1, self.nels[1] + 1
        nn = self.node_dofs.shape[1]
        ids = np.arange(nn)
        i = ids % NX
        fixed_nodes = ids[i == 0]
        self.fixed = np.zeros(self.ndof, np.uint8)
        self.fixed[(self.node_dofs[:, fixed_nodes] - 1).ravel()] = 1
        j = self.nels[1] // 2
        k = 0
        if dim == 3 and self.load == "centre":
            k = self.nels[2] // 2
        self.load_node = (self.nels[0]) + NX * (j + NY * k)
        self.force_dof = int(self.node_dofs[1, self.load_node])  # 1-based
        self.f = np.zeros(self.ndof)
        self.f[self.force_dof - 1] = -self.force


class Oracle:
    """Holds an orc_op for a Cantilever and the reference-side operations."""

    def __init__(self, prob: Cantilever, xmin=1e-3, p=1.0, pbi=True, quad_order=4, kind=0,
                 proj=0, beta=0.0, h=None):
        """``h``: cell edge lengths (unit cells by default), as the sizes of
        RectilinearGrid (src/TopOptProblems/grids.jl:80-92)."""
        self.prob = prob
        self.dim = prob.dim
        self.Ke = np.ascontiguousarray(element_stiffness(prob.dim, prob.E, prob.nu, h=h,
                                                          quad_order=quad_order))
        self.cell_dofs_c = np.ascontiguousarray(prob.cell_dofs.T.reshape(-1))
        self.fixed = np.ascontiguousarray(prob.fixed)
        op = _Op()
        op.dim = prob.dim
        op.ndpc = self.Ke.shape[0]
        op.nel = prob.nel
        op.ndof = prob.ndof
        op.cell_dofs = self.cell_dofs_c.ctypes.data
        op.Ke = self.Ke.ctypes.data
        op.fixed = self.fixed.ctypes.data
        op.xmin = xmin
        op.p = p
        op.pbi = int(pbi)
        op.kind = int(kind)
        op.proj = int(proj)
        op.beta = float(beta)
        if lib().orc_op_init(C.byref(op)) != 0:
            raise MemoryError("oracle op init failed")
        self.op = op
        self.vars = np.ones(prob.nel)

    def __del__(self):
        try:
            lib().orc_op_free(C.byref(self.op))
        except Exception:
            pass

    @property
    def meandiag(self):
        return self.op.meandiag

    def set_vars(self, x):
        self.vars = np.ascontiguousarray(x, np.float64).copy()
        lib().orc_op_set_vars(C.byref(self.op), self.vars)

    def matvec(self, x):
        y = np.empty(self.prob.ndof)
        lib().orc_matvec(C.byref(self.op), np.ascontiguousarray(x, np.float64), y)
        return y

    def jacobi_diag(self):
        d = np.empty(self.prob.ndof)
        lib().orc_jacobi_diag(C.byref(self.op), d)
        return d

    def pcg(self, b, x0=None, abstol=0.0, reltol=1e-8, maxiter=100000, jacobi=True):
        x = np.zeros(self.prob.ndof) if x0 is None else np.array(x0, np.float64)
        res = C.c_double()
        it = lib().orc_pcg(C.byref(self.op), np.ascontiguousarray(b, np.float64), x, abstol,
                           reltol, maxiter, int(x0 is None), int(jacobi), C.byref(res))
        return x, int(it), res.value

    def pcg_energy(self, b, tol, x0=None, energy=0.0, maxiter=100000, jacobi=True):
        """Jacobi PCG stopped by EnergyCriteria; returns (x, iterations,
        energy, status, ||r||) with status 1 converged, 0 maxiter, -1 NaN,
        -2 negative x'Ax."""
        x = np.zeros(self.prob.ndof) if x0 is None else np.array(x0, np.float64)
        e = C.c_double(energy)
        st = C.c_int()
        res = C.c_double()
        it = lib().orc_pcg_energy(C.byref(self.op), np.ascontiguousarray(b, np.float64), x,
                                  float(tol), int(maxiter), int(x0 is None), int(jacobi),
                                  C.byref(e), C.byref(st), C.byref(res))
        return x, int(it), e.value, st.value, res.value

    def assembled(self, safe=False):
        """globalinfo.K after assemble! (src/TopOptProblems/assemble.jl:28-90:
        sum_e rho_e Ke on cell_dofs, unconstrained Ke) and Ferrite's apply!
        with zero prescribed values: prescribed rows and columns zeroed, their
        diagonal the mean |diag| of the assembled K; ``safe`` then sets the
        mean |diag| of that constrained matrix on a zero diagonal
        (src/FEA/solvers_api.jl:232-239).  scipy CSR
        in the caller's DOF order."""
        import scipy.sparse as sp
        r = np.array([rho_drho(v, self.op.p, self.op.xmin, bool(self.op.pbi), self.op.kind,
                               self.op.proj, self.op.beta)[0] for v in self.vars])
        cd = self.prob.cell_dofs - 1  # (ndpc, nel)
        nd = cd.shape[0]
        rows = np.repeat(cd.T, nd, axis=1).ravel()
        cols = np.tile(cd.T, (1, nd)).ravel()
        vals = (r[:, None, None] * self.Ke[None]).ravel()
        K = sp.csr_matrix((vals, (rows, cols)), shape=(self.prob.ndof,) * 2).tolil()
        m = np.abs(K.diagonal()).mean()
        fx = np.flatnonzero(self.prob.fixed)
        K[fx, :] = 0.0
        K[:, fx] = 0.0
        K = K.tocsr()
        d = K.diagonal()
        d[fx] = m
        if safe:  # meandiag of the constrained K (solve_system! runs it after apply!)
            d[d == 0.0] = np.abs(d).mean()
        K.setdiag(d)
        return K, m

    def cg_assembled(self, K, b, reltol=1e-8, abstol=0.0, maxiter=100000, jacobi=False):
        """IterativeSolvers' (P)CG on an assembled matrix (the order of
        orc_pcg): returns (x, iterations, ||r||)."""
        n = b.size
        x = np.zeros(n)
        r = b.copy()
        dinv = 1.0 / K.diagonal() if jacobi else np.ones(n)
        res = np.linalg.norm(r)
        tol = max(reltol * res, abstol)
        u = np.zeros(n)
        rho = 1.0
        it = 0
        while it < maxiter and not res <= tol:
            c = dinv * r
            rho_prev, rho = rho, c @ r
            u = c + (rho / rho_prev) * u
            c = K @ u
            alpha = rho / (u @ c)
            x += alpha * u
            r -= alpha * c
            res = np.linalg.norm(r)
            it += 1
        return x, it, res

    def direct(self, f):
        u = np.empty(self.prob.ndof)
        rc = lib().orc_direct_solve(C.byref(self.op), np.ascontiguousarray(f, np.float64), u)
        return u, rc

    def element_energy(self, u):
        ce = np.empty(self.prob.nel)
        g = np.empty(self.prob.nel)
        obj = lib().orc_element_energy(C.byref(self.op), np.ascontiguousarray(u, np.float64),
                                       self.vars, ce, g)
        return obj, ce, g


class Filter:
    """DensityFilterFun restated: sparse J built by the reference's BFS."""

    def __init__(self, nels, rmin, h=None):
        dim = len(nels)
        ne = np.asarray(nels, np.int64)
        hh = np.ones(dim) if h is None else np.asarray(h, np.float64)
        nnz = lib().orc_filter_build(dim, ne, hh, rmin, None, None, None)
        if nnz < 0:
            raise ValueError("DensityFilterFun: no neighbouring nodes within rmin (ArgumentError)")
        self.nel = int(np.prod(ne))
        self.rowptr = np.zeros(self.nel + 1, np.int64)
        self.col = np.zeros(nnz, np.int64)
        self.val = np.zeros(nnz)
        lib().orc_filter_build(dim, ne, hh, rmin, self.rowptr.ctypes.data, self.col.ctypes.data,
                               self.val.ctypes.data)
        self.colsum = self.apply_t(np.ones(self.nel))

    def apply(self, x):
        y = np.empty(self.nel)
        lib().orc_csr_mv(self.nel, self.rowptr, self.col, self.val,
                         np.ascontiguousarray(x, np.float64), y)
        return y

    def apply_t(self, x):
        y = np.empty(self.nel)
        lib().orc_csr_mtv(self.nel, self.nel, self.rowptr, self.col, self.val,
                          np.ascontiguousarray(x, np.float64), y)
        return y

    def dense(self):
        J = np.zeros((self.nel, self.nel))
        for i in range(self.nel):
            J[i, self.col[self.rowptr[i]:self.rowptr[i + 1]]] = self.val[self.rowptr[i]:self.rowptr[i + 1]]
        return J


def filter_stencil(rmin, h):
    """Nodes within rmin of a cell centre, as offsets (a, b[, c]) from the
    cell's lower corner node, with the weights rmin - distance: the nodes the
    reference's neighbour search pushes (src/CheqFilters/CheqFilters.jl:66-118)."""
    dim = len(h)
    R = int(np.ceil(rmin / min(h))) + 1
    rng = np.arange(-R, R + 2)
    grids = np.meshgrid(*([rng] * dim), indexing="ij")
    offs = np.stack([g.ravel() for g in grids], 1)
    d = np.sqrt((((offs - 0.5) * np.asarray(h)) ** 2).sum(1))
    keep = d < rmin
    return offs[keep], rmin - d[keep]


def filter_closed_form(nels, rmin, x, transpose=False, h=None):
    """DensityFilterFun's J x (or J^T x) from the closed form of its
    construction on a structured grid (src/CheqFilters/density_filter.jl
    :49-108): the neighbour search pushes node n once per adjacent cell, so
    (J x)_i = sum_n w_n S(n) / sum_n v_n w_n with S(n) = sum of the cells
    around n and v_n their count.  Plain numpy on shifted arrays, for grids
    far too large for the sparse J of Filter; pinned against Filter on small
    grids (unit and box cells)."""
    dim = len(nels)
    h = np.ones(dim) if h is None else np.asarray(h, np.float64)
    offs, w = filter_stencil(rmin, h)
    if offs.shape[0] == 0:
        raise ValueError("DensityFilterFun: no neighbouring nodes within rmin (ArgumentError)")
    R = int(np.abs(offs).max()) + 1
    cshape = tuple(int(n) for n in nels[::-1])       # (nz, ny, nx) / (ny, nx)
    nshape = tuple(n + 1 for n in cshape)

    def node_sum(c):  # cell array -> node array: sum of the 2^dim cells around each node
        p = np.pad(c, 1)
        out = np.zeros(nshape)
        for corner in np.ndindex(*([2] * dim)):
            out += p[tuple(slice(k, k + n) for k, n in zip(corner, nshape))]
        return out

    def gather(nodes, scale):  # node array -> cell array: sum_s w_s nodes(n_i + off_s)
        p = np.pad(nodes, R)
        out = np.zeros(cshape)
        for o, ws in zip(offs, w):
            sl = tuple(slice(R + int(o[dim - 1 - a]), R + int(o[dim - 1 - a]) + cshape[a])
                       for a in range(dim))
            out += ws * p[sl]
        return out * scale

    def scatter(cells):  # cell array -> node array: sum_s w_s cells(n - off_s)
        p = np.pad(cells, R + 1)
        out = np.zeros(nshape)
        for o, ws in zip(offs, w):
            sl = tuple(slice(R + 1 - int(o[dim - 1 - a]), R + 1 - int(o[dim - 1 - a]) + nshape[a])
                       for a in range(dim))
            out += ws * p[sl]
        return out

    den = gather(node_sum(np.ones(cshape)), 1.0)
    X = np.asarray(x, np.float64).reshape(cshape)
    if not transpose:
        return gather(node_sum(X), 1.0 / den).ravel()
    T = scatter(X / den)
    out = np.zeros(cshape)
    for corner in np.ndindex(*([2] * dim)):
        out += T[tuple(slice(k, k + n) for k, n in zip(corner, cshape))]
    return out.ravel()


def set_threads(n=0):
    """OpenMP threads of the oracle (n <= 0: unchanged); returns the count."""
    return int(lib().orc_set_threads(int(n)))


def cell_volumes(nels, h=None, quad_order=4):
    """get_cell_volumes (src/TopOptProblems/elementinfo.jl:180-192) of the
    structured grid: every cell has the same quadrature volume."""
    dim = len(nels)
    hh = np.ones(dim) if h is None else np.asarray(h, np.float64)
    v = lib().orc_cell_volume(dim, np.ascontiguousarray(hh), quad_order)
    if v < 0:
        raise ValueError("bad cell volume arguments")
    return np.full(int(np.prod(nels)), v)


def volume(x, cellvol, fraction=True):
    """VolumeFun value (src/Functions/volume.jl:59-62); its gradient is the
    constant cellvol (/ total volume when fraction)."""
    x = np.ascontiguousarray(x, np.float64)
    return lib().orc_volume(x.size, x, np.ascontiguousarray(cellvol, np.float64), int(fraction))


def oc_update(x, dc, colsum, volfrac, move=0.2, tol=1e-12):
    xnew = np.empty_like(x)
    lam = lib().orc_oc_update(x.size, np.ascontiguousarray(x), np.ascontiguousarray(dc),
                              np.ascontiguousarray(colsum), volfrac, move, tol, xnew)
    return xnew, lam


def simp(prob: Cantilever, volfrac, rmin, iters, p=3.0, xmin=1e-9, x0=None, solver="pcg",
         reltol=1e-8, warm=True, move=0.2, oc_tol=1e-12, pbi=True, maxiter=100000, kind=0,
         proj=0, beta=0.0, optimizer="oc", mma=(0.5, 1.2, 0.7)):
    """Oracle SIMP loop: x~ = J x; solve; compliance; dc = J^T grad; OC, or
    MMA87 (``optimizer="mma"``, oracle/mma_oracle.py, (s_init, s_incr,
    s_decr) = ``mma``) on min c s.t. V(J x) <= volfrac.
    ``kind``/``proj``/``beta`` select the penalty family (orc_rho_drho_fun).

    Returns dict(x, xphys, obj history, pcg iteration counts, u)."""
    orc = Oracle(prob, xmin=xmin, p=p, pbi=pbi, kind=kind, proj=proj, beta=beta)
    flt = Filter(prob.nels, rmin)
    x = np.full(prob.nel, volfrac if x0 is None else x0, dtype=np.float64) if np.isscalar(
        x0) or x0 is None else np.array(x0, np.float64)
    u = None
    objs, its = [], []
    normf = np.linalg.norm(prob.f)
    if optimizer == "mma":
        from .mma_oracle import MMA87
        opt = MMA87(*mma)
    for _ in range(iters):
        xt = flt.apply(x)
        orc.set_vars(xt)
        if solver == "direct":
            u, rc = orc.direct(prob.f)
            it = 0
        else:
            if warm and u is not None:
                u, it, _ = orc.pcg(prob.f, x0=u, abstol=reltol * normf, reltol=0.0, maxiter=maxiter)
            else:
                u, it, _ = orc.pcg(prob.f, abstol=reltol * normf, reltol=0.0, maxiter=maxiter)
        obj, ce, g = orc.element_energy(u)
        dc = flt.apply_t(g)
        if optimizer == "mma":
            dg = flt.colsum / prob.nel
            x, _, _ = opt.step(x, dc, dg, float(dg @ x) - volfrac)
        else:
            x, _ = oc_update(x, dc, flt.colsum, volfrac, move, oc_tol)
        objs.append(obj)
        its.append(it)
    return dict(x=x, xphys=flt.apply(x), obj=np.array(objs), iters=np.array(its), u=u)


def beso(prob: Cantilever, vol_limit, rmin, maxiter=200, tol=1e-4, p=3.0, er=0.02,
         sens_tol=None, k=10, xmin=1e-3, x0=None, black=None, white=None, solver="direct",
         reltol=1e-12):
    """(b::BESO)(x0; black, white) restated (src/Algorithms/beso.jl:93-188):
    objective trace of length k, sensitivities averaged over two iterations,
    volume target shrunk by er per iteration, threshold bisection to sens_tol."""
    sens_tol = tol / 100 if sens_tol is None else sens_tol
    orc = Oracle(prob, xmin=xmin, p=p)
    flt = Filter(prob.nels, rmin)
    n = prob.nel
    x0 = np.full(n, 0.5) if x0 is None else np.asarray(x0, np.float64)
    topology = np.round(x0)  # ties to even, as Julia's round
    free = np.ones(n, bool)
    if black is not None and len(black):
        topology[np.asarray(black, bool)] = 1.0
        free &= ~np.asarray(black, bool)
    if white is not None and len(white):
        topology[np.asarray(white, bool)] = 0.0
        free &= ~np.asarray(white, bool)
    cv = 1.0
    total = float(n) * cv
    vol = true_vol = topology.sum() * cv / total
    change, it = 1.0, 0
    trace = np.zeros(k)
    hist = []
    sens = None
    normf = np.linalg.norm(prob.f)
    while (change > tol or true_vol > vol_limit) and it < maxiter:
        it += 1
        old = None if sens is None else sens.copy()
        vol = max(vol * (1 - er), vol_limit)
        for j in range(max(2, k - it + 2), k + 1):
            trace[j - 2] = trace[j - 1]
        orc.set_vars(flt.apply(topology))
        if solver == "direct":
            u, _ = orc.direct(prob.f)
        else:
            u, _, _ = orc.pcg(prob.f, abstol=reltol * normf, reltol=0.0)
        obj, _, g = orc.element_energy(u)
        trace[k - 1] = obj
        hist.append(obj)
        sens = -flt.apply_t(g)
        if it > 1:
            sens = (sens + old) / 2
        l1, l2 = sens.min(), sens.max()
        while (l2 - l1) / l2 > sens_tol:
            th = (l1 + l2) / 2
            topology[free] = (sens[free] > th).astype(np.float64)
            if topology.sum() * cv - vol * total > 0:
                l1 = th
            else:
                l2 = th
        true_vol = topology.sum() * cv / total
        if it >= k:
            lo, hi = trace[:k // 2].sum(), trace[k // 2:].sum()
            change = abs(lo - hi) / hi
    return dict(topology=topology, objval=trace[k - 1], change=change, converged=change <= tol,
                fevals=it, volume=true_vol, obj=np.array(hist))
