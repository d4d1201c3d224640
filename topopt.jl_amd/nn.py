"""Learned warm start: network handle, dataset generation, training, I/O.

The network and its training run entirely in libtopopt_hip (BF16 implicit-GEMM
convolutions on MFMA, FP32 accumulation, Adam on the device); this module is
the host driver.  Vectors here are node-major (dof = 3*node + d), the order
of a solver context created without Ferrite numbering.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import Context, SimpOpts, TohipError, check, dptr, load_library

_fp = C.POINTER(C.c_float)


def _fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_fp)


class WarmStartNet:
    """Encoder-decoder warm-start network on an nx x ny x nz cell grid."""

    def __init__(self, nels, max_batch=1, max_samples=0, seed=0, fixed=None):
        self.lib = load_library()
        self.nels = tuple(int(n) for n in nels)
        nx, ny, nz = self.nels
        self.ncells = nx * ny * nz
        self.nnodes = (nx + 1) * (ny + 1) * (nz + 1)
        self.ndofs = 3 * self.nnodes
        if fixed is None:  # cantilever: the x = 0 face clamped
            fixed = np.zeros(self.ndofs, np.uint8)
            nodes = np.arange(self.nnodes)
            fixed.reshape(-1, 3)[nodes % (nx + 1) == 0] = 1
        self.fixed = np.ascontiguousarray(fixed, np.uint8)
        self.ptr = C.c_void_p()
        ne = np.asarray(self.nels, np.int64)
        rc = self.lib.tohip_nn_create(C.byref(self.ptr), _lib.i64ptr(ne), int(max_batch),
                                      int(max_samples), C.c_uint64(seed),
                                      _lib.u8ptr(self.fixed))
        if rc < 0:
            raise TohipError(f"tohip_nn_create failed: {self.lib.tohip_nn_last_error(None).decode()}")
        self.nparams = int(self.lib.tohip_nn_num_params(self.ptr))

    @classmethod
    def for_problem(cls, problem, weights=None, max_batch=1, seed=0):
        """Network on a 3D problem's grid (its clamped DOFs mapped to the
        node-major order the solver uses inside), loaded from ``weights``
        (.npz written by save() or the training driver) when given."""
        nd = np.asarray(problem.node_dofs, np.int64)  # (3, nnodes), 1-based caller DOFs
        fixed = np.ascontiguousarray(problem.fixed[(nd - 1).T.reshape(-1)], np.uint8)
        net = cls(problem.nels, max_batch=max_batch, seed=seed, fixed=fixed)
        if weights is not None:
            net.load(weights)
        return net

    def _check(self, rc, what):
        if rc < 0:
            raise TohipError(f"{what} failed ({rc}): {self.lib.tohip_nn_last_error(self.ptr).decode()}")
        return rc

    def close(self):
        """Release the device network.  Solver contexts it is attached to
        are detached by the library (a later learned warm start there fails
        with a state error instead of using freed memory)."""
        if self.ptr:
            self.lib.tohip_nn_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters ----
    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        self._check(self.lib.tohip_nn_get_params(self.ptr, _fptr(p)), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {p.size}")
        self._check(self.lib.tohip_nn_set_params(self.ptr, _fptr(p)), "set_params")

    def save(self, path, **meta):
        np.savez_compressed(path, params=self.get_params(), nels=np.array(self.nels), **meta)

    def load(self, path):
        d = np.load(path, allow_pickle=False)
        self.set_params(d["params"])
        return d

    # ---- data / training ----
    def add_sample(self, xphys, f, u):
        self._check(self.lib.tohip_nn_add_sample(
            self.ptr, dptr(np.ascontiguousarray(xphys, np.float64)),
            dptr(np.ascontiguousarray(f, np.float64)), dptr(np.ascontiguousarray(u, np.float64))),
            "add_sample")

    @property
    def nsamples(self):
        return int(self.lib.tohip_nn_num_samples(self.ptr))

    def clear_samples(self):
        self._check(self.lib.tohip_nn_clear_samples(self.ptr), "clear_samples")

    def train(self, steps, batch, lr, seed=0, first=0, count=None):
        count = self.nsamples - first if count is None else count
        hist = np.zeros(max(int(steps), 1))
        self._check(self.lib.tohip_nn_train(self.ptr, int(first), int(count), int(steps),
                                            int(batch), float(lr), C.c_uint64(seed),
                                            dptr(hist)), "train")
        return hist[:steps]

    def eval(self, first, count):
        loss = C.c_double()
        self._check(self.lib.tohip_nn_eval(self.ptr, int(first), int(count), C.byref(loss)),
                    "eval")
        return loss.value

    def predict(self, xphys, f):
        out = np.empty(self.ndofs, np.float32)
        self._check(self.lib.tohip_nn_predict(
            self.ptr, dptr(np.ascontiguousarray(xphys, np.float64)),
            dptr(np.ascontiguousarray(f, np.float64)), _fptr(out)), "predict")
        return out


# (cin, cout, kernel, level) of the 15 convolutions, as built by tohip_nn_create:
# level l runs on the grid coarsened 2^l times; the last layer maps cells to nodes.
LAYERS = [(6, 16, 3, 0), (16, 16, 3, 0), (16, 32, 3, 1), (32, 32, 3, 1), (32, 64, 3, 2),
          (64, 64, 3, 2), (64, 128, 3, 3), (128, 128, 3, 3), (192, 64, 3, 2), (64, 64, 3, 2),
          (96, 32, 3, 1), (32, 32, 3, 1), (48, 16, 3, 0), (16, 16, 3, 0), (16, 3, 2, -1)]


def conv_flops(nels, batch=1):
    """Algorithmic FLOPs of one forward pass (multiply-add = 2; the used
    input/output channels only, not the padding to 8 / 32)."""
    nx, ny, nz = nels
    total = 0
    for cin, cout, k, lev in LAYERS:
        pos = ((nx + 1) * (ny + 1) * (nz + 1) if lev < 0 else
               (nx >> lev) * (ny >> lev) * (nz >> lev))
        total += 2 * pos * k ** 3 * cin * cout
    return batch * total


def conv_check(x, w, b, dy, k, pad, act):
    """One BF16 conv layer forward/backward on the GPU (see tohip_nn_conv_check)."""
    lib = load_library()
    B, D, H, W, cin = x.shape
    cout = w.shape[0]
    dims = (C.c_int * 3)(D, H, W)
    x = np.ascontiguousarray(x, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    dy = np.ascontiguousarray(dy, np.float32)
    y = np.empty(dy.shape, np.float32)
    dx = np.empty(x.shape, np.float32)
    dw = np.empty(w.shape, np.float32)
    db = np.empty(b.shape, np.float32)
    rc = lib.tohip_nn_conv_check(B, dims, cin, cout, k, pad, act, _fptr(x), _fptr(w), _fptr(b),
                                 _fptr(dy), _fptr(y), _fptr(dx), _fptr(dw), _fptr(db))
    if rc < 0:
        raise TohipError(f"conv_check failed: {lib.tohip_nn_last_error(None).decode()}")
    return y, dx, dw, db


# ------------------------------------------------------------ datasets ----

def cantilever_load(nels, j, k, force=1.0):
    """Node-major load vector: -force along y at node (nx, j, k)."""
    nx, ny, nz = nels
    f = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    n = nx + (nx + 1) * (j + (ny + 1) * k)
    f[3 * n + 1] = -force
    return f


@dataclass
class SimpRunSpec:
    volfrac: float
    load_j: int
    load_k: int
    seed: int


def make_context(nels, xmin=1e-9, p=3.0):
    """Solver context on the node-major cantilever (x = 0 face clamped)."""
    nx, ny, nz = nels
    nnodes = (nx + 1) * (ny + 1) * (nz + 1)
    fixed = np.zeros(3 * nnodes, np.uint8)
    fixed.reshape(-1, 3)[np.arange(nnodes) % (nx + 1) == 0] = 1
    ctx = Context(3, nels, (1.0, 1.0, 1.0), 1.0, 0.3, xmin, p, True, None, fixed)
    return ctx, fixed


def simp_pairs(ctx, nels, spec: SimpRunSpec, iters, keep_every, rmin=1.5, rel_tol=1e-8,
               precision=_lib.TOHIP_MIXED):
    """Run one SIMP run on the device and yield (xphys, f, u) every
    `keep_every` iterations (the warm-start network's training pairs)."""
    lib = ctx.lib
    f = cantilever_load(nels, spec.load_j, spec.load_k)
    check(lib.tohip_set_load(ctx.ptr, dptr(f)), ctx.ptr, "tohip_set_load")
    x = np.full(ctx.ncells, spec.volfrac)
    obj = np.zeros(1)
    its = np.zeros(1, np.int64)
    chg = np.zeros(1)
    xp = np.empty(ctx.ncells)
    u = np.empty(ctx.ndofs)
    for it in range(iters):
        o = SimpOpts(iterations=1, volfrac=spec.volfrac, rmin=rmin, move=0.2, oc_tol=1e-12,
                     rel_tol=rel_tol, maxiter=200000, precision=precision,
                     warm_start=1 if it > 0 else 0,
                     inner_rel_tol=1e-3 if precision == _lib.TOHIP_MIXED_MG else 0.1)
        check(lib.tohip_simp_run(ctx.ptr, C.byref(o), dptr(x) if it == 0 else None, dptr(obj),
                                 its.ctypes.data_as(C.POINTER(C.c_int64)), dptr(chg)),
              ctx.ptr, "tohip_simp_run")
        if it % keep_every == keep_every - 1 or it == 0:
            check(lib.tohip_get_xphys(ctx.ptr, dptr(xp)), ctx.ptr, "tohip_get_xphys")
            check(lib.tohip_get_u(ctx.ptr, dptr(u)), ctx.ptr, "tohip_get_u")
            yield xp.copy(), f, u.copy(), int(its[0])


def compare_warm_starts(ctx, net, nels, spec: SimpRunSpec, iters, rmin=1.5, rel_tol=1e-8,
                        precision=_lib.TOHIP_MIXED):
    """PCG iterations per SIMP iteration of one run started three ways: every
    solve from zero ("zero"), from the previous displacement ("previous"),
    and from the network's guess combined with the previous displacement
    ("learned": the point of their span with the smallest residual; the first
    solve uses the prediction alone).  Returns {name: (pcg_iters, obj)}."""
    lib = ctx.lib
    check(lib.tohip_nn_attach(ctx.ptr, net.ptr), ctx.ptr, "tohip_nn_attach")
    try:
        f = cantilever_load(nels, spec.load_j, spec.load_k)
        check(lib.tohip_set_load(ctx.ptr, dptr(f)), ctx.ptr, "tohip_set_load")
        out = {}
        for name, ws in (("zero", 0), ("previous", 1), ("learned", 2)):
            x = np.full(ctx.ncells, spec.volfrac)
            obj = np.zeros(iters)
            its = np.zeros(iters, np.int64)
            o = SimpOpts(iterations=iters, volfrac=spec.volfrac, rmin=rmin, move=0.2,
                         oc_tol=1e-12, rel_tol=rel_tol, maxiter=200000, precision=precision,
                         warm_start=ws,
                         inner_rel_tol=1e-3 if precision == _lib.TOHIP_MIXED_MG else 0.1)
            check(lib.tohip_simp_run(ctx.ptr, C.byref(o), dptr(x), dptr(obj),
                                     its.ctypes.data_as(C.POINTER(C.c_int64)), None),
                  ctx.ptr, "tohip_simp_run")
            out[name] = (its, obj)
        return out
    finally:  # the context does not keep the caller's network
        lib.tohip_nn_attach(ctx.ptr, None)


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                               "warmstart_64x32x32.npz")
