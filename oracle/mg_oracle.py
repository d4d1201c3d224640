"""TEST INFRASTRUCTURE ONLY — the oracle of the multigrid preconditioner.

A numpy/scipy restatement (FP64) of the V(1,1) cycle that libtopopt_hip uses
to precondition its mixed-precision PCG (``TOHIP_MIXED_MG``).  TopOpt.jl has no
multigrid (its matrix-free solver is Jacobi PCG, src/FEA/solvers_api.jl:177-220),
so this is pinned by mathematics rather than reference vectors: the assembled
level operators agree with the pinned oracle matvec (oracle.c) on the fine
level, the cycle is a symmetric positive definite operator, the prolongation
reproduces linear fields exactly, and two-grid PCG converges in far fewer
iterations than Jacobi PCG (tests/test_mg_oracle.py).  Only ``tests/`` import
this module.

Conventions (node-major, as the device): node n = i + NX (j + NY k),
DOF = dim n + d.  Level l + 1 halves every cell count of level l; coarse cell
moduli are the means of their children; the coarse element matrix is the Q1
matrix of the 2^l h box; coarse DOF (N, d) is prescribed when any fine node
(2I + a, 2J + b, 2K + c), a, b, c in {-1, 0, 1}, has d prescribed.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import oracle as O

COARSE_MAX_NODES = 600
COARSE_MAX_CELLS = 400

OFF3 = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
OFF2 = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]


def _nodes(nels):
    return tuple(n + 1 for n in nels)


def _fits(nels):
    nn = int(np.prod(_nodes(nels)))
    return nn <= COARSE_MAX_NODES and int(np.prod(nels)) <= COARSE_MAX_CELLS


def level_grids(nels):
    """Halve while every count is even, until the coarsest fits the
    one-workgroup coarse solver (at least one coarse level)."""
    gs = [tuple(nels)]
    while len(gs) < 2 or not _fits(gs[-1]):
        f = gs[-1]
        if any(n % 2 for n in f):
            break
        gs.append(tuple(n // 2 for n in f))
    if len(gs) < 2:
        raise ValueError("multigrid: every cell count must be even")
    return gs


def _hvec(h, dim, scale=1):
    """Cell sizes per axis: a scalar h is the cube of side h."""
    hv = [float(v) for v in h] if np.ndim(h) else [float(h)] * dim
    if len(hv) != dim:
        raise ValueError("one cell size per axis")
    return [v * scale for v in hv]


def assemble(nels, h, E, fixed_node_mask, meandiag):
    """K with the reference boundary rule: prescribed rows/columns removed,
    meandiag on their diagonal (matrix_free_operator.jl:66-105)."""
    dim = len(nels)
    N = _nodes(nels)
    NX, NY = N[0], N[1]
    nx, ny = nels[0], nels[1]
    nz = nels[2] if dim == 3 else 1
    e = np.arange(int(np.prod(nels)))
    ci, cj, ck = e % nx, (e // nx) % ny, e // (nx * ny)
    off = OFF3 if dim == 3 else OFF2
    nodes = np.stack([(ci + a) + NX * ((cj + b) + NY * (ck + c)) for a, b, c in off], 1)
    ndpc = dim * len(off)
    dofs = (dim * nodes[:, :, None] + np.arange(dim)[None, None, :]).reshape(-1, ndpc)
    Ke = O.element_stiffness(dim, h=_hvec(h, dim))
    rows = np.repeat(dofs, ndpc, axis=1).ravel()
    cols = np.tile(dofs, (1, ndpc)).ravel()
    vals = (E[:, None] * Ke.ravel()[None, :]).ravel()
    n = dim * int(np.prod(N))
    K = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    fx = np.array([(fixed_node_mask[i // dim] >> (i % dim)) & 1 for i in range(n)], bool)
    D = sp.diags((~fx).astype(float))
    return (D @ K @ D + sp.diags(fx * meandiag)).tocsr(), fx


def coarse_mask(nf, nc, mask):
    dim = len(nf)
    F, C = _nodes(nf), _nodes(nc)
    NZf = F[2] if dim == 3 else 1
    NZc = C[2] if dim == 3 else 1
    out = np.zeros(int(np.prod(C)), np.uint8)
    for K in range(NZc):
        for J in range(C[1]):
            for I in range(C[0]):
                m = 0
                for dk in ((-1, 0, 1) if dim == 3 else (0,)):
                    for dj in (-1, 0, 1):
                        for di in (-1, 0, 1):
                            i, j, k = 2 * I + di, 2 * J + dj, 2 * K + dk
                            if 0 <= i < F[0] and 0 <= j < F[1] and 0 <= k < NZf:
                                m |= int(mask[i + F[0] * (j + F[1] * k)])
                out[I + C[0] * (J + C[1] * K)] = m
    return out


def _p1(n):
    nf, nc = n + 1, n // 2 + 1
    r, c, v = [], [], []
    for i in range(nf):
        if i % 2 == 0:
            r.append(i); c.append(i // 2); v.append(1.0)
        else:
            r += [i, i]; c += [(i - 1) // 2, (i + 1) // 2]; v += [0.5, 0.5]
    return sp.csr_matrix((v, (r, c)), shape=(nf, nc))


def prolongation(nf):
    """Trilinear (bilinear) P from the coarse to the fine nodes, per component."""
    dim = len(nf)
    Pn = _p1(nf[0])
    Pn = sp.kron(_p1(nf[1]), Pn)
    if dim == 3:
        Pn = sp.kron(_p1(nf[2]), Pn)
    return sp.kron(Pn, sp.eye(dim)).tocsr()


def lambda_bound(dim, h):
    """1.01 lambda_max(diag(Ke)^-1/2 Ke diag(Ke)^-1/2): bounds
    lambda_max(D^-1 K) for any cell moduli (x^T K x = sum_e E_e x_e^T Ke x_e)."""
    Ke = O.element_stiffness(dim, h=_hvec(h, dim))
    d = np.sqrt(np.diag(Ke))
    return 1.01 * np.linalg.eigvalsh(Ke / np.outer(d, d)).max()


def jacobi_weight(dim, h):
    """1.9 / lambda_bound: damped Jacobi with this weight converges and the
    V-cycle is SPD whatever the design."""
    return 1.9 / lambda_bound(dim, h)


def chebyshev(lmax, kappa, iters):
    """Three-term Chebyshev coefficients on [lmax / kappa, lmax]:
    x_1 = c0 D^-1 b; x_{k+1} = x_k + a_k (x_k - x_{k-1}) + g_k D^-1 (b - K x_k)."""
    lo = lmax / kappa
    theta, delta = 0.5 * (lmax + lo), 0.5 * (lmax - lo)
    sigma = theta / delta
    rho = 1.0 / sigma
    a, g = [0.0], [0.0]
    for _ in range(1, iters):
        rn = 1.0 / (2.0 * sigma - rho)
        a.append(rn * rho)
        g.append(2.0 * rn / delta)
        rho = rn
    return 1.0 / theta, a, g


def coarsen_E(E, nels):
    dim = len(nels)
    if dim == 3:
        e = E.reshape(nels[2], nels[1], nels[0])
        s = sum(e[c::2, b::2, a::2] for a in (0, 1) for b in (0, 1) for c in (0, 1))
        return (s / 8.0).ravel()
    e = E.reshape(nels[1], nels[0])
    return (sum(e[b::2, a::2] for a in (0, 1) for b in (0, 1)) / 4.0).ravel()


class MGOracle:
    """V(1,1) cycle of the device preconditioner, in FP64."""

    def __init__(self, nels, E, fixed_node_mask, meandiag, omega=None, sweeps=30, h=1.0,
                 cheb_iters=32, cheb_kappa=1000.0):
        self.sweeps = sweeps
        self.cheb = None
        self.grids = level_grids(nels)
        dim = len(nels)
        self.omega = [omega if omega is not None else jacobi_weight(dim, _hvec(h, dim, 2 ** l))
                      for l in range(len(self.grids))]
        self.K, self.fx, self.dinv, self.P = [], [], [], []
        mask = np.asarray(fixed_node_mask, np.uint8)
        El = np.asarray(E, np.float64)
        for l, g in enumerate(self.grids):
            if l > 0:
                mask = coarse_mask(self.grids[l - 1], g, mask)
                El = coarsen_E(El, self.grids[l - 1])
            K, fx = assemble(g, _hvec(h, dim, 2 ** l), El, mask, meandiag)
            self.K.append(K)
            self.fx.append(fx)
            self.dinv.append(1.0 / K.diagonal())
            if l > 0:
                self.P.append(prolongation(self.grids[l - 1]))
        self.coarse_fits = _fits(self.grids[-1])
        if self.coarse_fits:
            lc = len(self.grids) - 1
            self.cheb = chebyshev(lambda_bound(dim, _hvec(h, dim, 2 ** lc)), cheb_kappa, cheb_iters)

    def _coarse(self, b):
        K, di, fx, w = self.K[-1], self.dinv[-1], self.fx[-1], self.omega[-1]
        if self.cheb is not None:  # the one-workgroup kernel: Chebyshev iteration
            c0, a, g = self.cheb
            dz = np.where(fx, 0.0, di)
            xo = np.zeros_like(b)
            x = c0 * dz * b
            for k in range(1, len(a)):
                x, xo = x + a[k] * (x - xo) + g[k] * dz * (b - K @ x), x
            return x
        # too large for that kernel: damped Jacobi sweeps
        x = np.where(fx, 0.0, w * di * b)
        for _ in range(self.sweeps - 1):
            x = np.where(fx, 0.0, x + w * di * (b - K @ x))
        return x

    def apply(self, r, l=0):
        if l == len(self.K) - 1:
            return self._coarse(r)
        K, di, fx, w = self.K[l], self.dinv[l], self.fx[l], self.omega[l]
        x = np.where(fx, 0.0, w * di * r)
        res = np.where(fx, 0.0, r - K @ x)
        bc = self.P[l].T @ res
        bc[self.fx[l + 1]] = 0.0
        xc = self.apply(bc, l + 1)
        x = np.where(fx, 0.0, x + self.P[l] @ xc)
        return np.where(fx, 0.0, x + w * di * (r - K @ x))


def pcg_iterations(K, b, M, tol=1e-8, maxiter=5000):
    """Preconditioned CG iteration count to ||r|| <= tol ||b||."""
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    p = z.copy()
    rz = r @ z
    t = tol * np.linalg.norm(b)
    for it in range(1, maxiter + 1):
        q = K @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= t:
            return it, x
        z = M(r)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return maxiter, x
