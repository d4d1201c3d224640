"""GPU tests of the learned warm start: every BF16 MFMA convolution shape the
network uses is checked (forward, input gradient, weight and bias gradients)
against the CPU oracle's emulation of the same BF16 storage points; the whole
network trains (loss falls) and its prediction feeds the solver's warm start.
"""
import os

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

# (cin, cout, k, pad, act, spatial) — the layer shapes of the network
SHAPES = [
    (8, 16, 3, 1, 1, (6, 4, 8)),     # first layer (6 used input channels, padded to 8)
    (16, 32, 3, 1, 1, (4, 4, 4)),
    (48, 16, 3, 1, 1, (4, 2, 4)),    # concatenation input, Cin not a multiple of 32
    (192, 64, 3, 1, 1, (2, 2, 2)),
    (128, 128, 3, 1, 1, (2, 2, 2)),
    (16, 8, 2, 1, 0, (4, 2, 4)),     # cells -> nodes output layer
    (8, 16, 2, 0, 0, (5, 3, 5)),     # its input-gradient shape
    # 65536 output positions: one wave per 16 positions walks all K-steps
    # (the small shapes above split K over the 4 waves of a block)
    (8, 16, 3, 1, 1, (32, 32, 32)),  # 4 taps per MFMA K-step
    (48, 16, 3, 1, 1, (32, 32, 32)),  # 2 taps per K-step in the input gradient
]


@pytest.mark.parametrize("cin,cout,k,pad,act,sp", SHAPES)
def test_conv_layer_matches_oracle(T, gpu, cin, cout, k, pad, act, sp):
    from oracle import nn_oracle as N
    import topopt_jl_amd.nn as NN
    g = np.random.default_rng(cin * 7 + cout)
    B = 2
    x = g.standard_normal((B, *sp, cin)).astype(np.float32)
    w = (g.standard_normal((cout, k ** 3, cin)) / np.sqrt(k ** 3 * cin)).astype(np.float32)
    b = g.standard_normal(cout).astype(np.float32) * 0.1
    osp = tuple(s + 2 * pad - k + 1 for s in sp)
    dy = g.standard_normal((B, *osp, cout)).astype(np.float32)
    y, dx, dw, db = NN.conv_check(x, w, b, dy, k, pad, act)
    ye, dxe, dwe, dbe = N.conv_layer_device_emulation(x, w, b, dy, k, pad, act)
    # FP32 accumulation vs float64 on identical BF16 operands
    assert relerr(y, ye) < 2e-5
    assert relerr(dx, dxe) < 1e-3  # dZ is BF16: rare tie flips of its rounding
    assert relerr(dw, dwe) < 1e-3
    assert relerr(db, dbe) < 1e-3


@pytest.mark.parametrize("nels", [(16, 8, 8), (48, 16, 16)])
def test_network_forward_matches_oracle(T, gpu, nels):
    """The whole encoder-decoder prediction (features, 15 convolutions, pools,
    upsampling into the concatenations, node output layer) against the
    oracle's float64 forward with the device's BF16 storage points.  The
    device accumulates in FP32, so a BF16 store can round the other way; the
    differences stay at the BF16 resolution (2^-8) of the activations."""
    from oracle import nn_oracle as N
    import topopt_jl_amd.nn as NN
    net = NN.WarmStartNet(nels, max_batch=1, seed=7)
    nx, ny, nz = nels
    g = np.random.default_rng(11)
    x = g.uniform(0.05, 1.0, net.ncells)
    f = NN.cantilever_load(nels, ny // 2, nz // 2) * 50.0
    f = f + 0.1 * g.standard_normal(f.size)
    u = net.predict(x, f)
    ue = N.network_forward(net.get_params(), nels, x, f, net.fixed)
    assert np.all(np.isfinite(u)) and np.abs(ue).max() > 0
    assert relerr(u.reshape(-1, 3), ue) < 2e-2


def test_network_trains_and_predicts(T, gpu):
    import topopt_jl_amd.nn as NN
    nels = (16, 8, 8)
    net = NN.WarmStartNet(nels, max_batch=4, max_samples=8, seed=1)
    assert net.nparams > 1_000_000  # channels 16/32/64/128, 3x3x3 kernels
    g = np.random.default_rng(3)
    # smooth cantilever-like targets: u_y ~ -(x/L)^2, u_x ~ -(x/L)(y - H/2)/H
    nx, ny, nz = nels
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1),
                             indexing="ij")
    xi, yj = ii.ravel() / nx, jj.ravel() / ny
    for i in range(8):
        x = g.uniform(0.1, 1.0, net.ncells)
        f = NN.cantilever_load(nels, 4, i % 8)
        u = np.zeros((net.nnodes, 3))
        u[:, 1] = -(xi ** 2) * (1 + 0.1 * i)
        u[:, 0] = -xi * (yj - 0.5)
        u[ii.ravel() == 0] = 0
        net.add_sample(x, f, u.ravel())
    l0 = net.eval(0, 8)
    hist = net.train(steps=100, batch=4, lr=3e-3, seed=0)
    l1 = net.eval(0, 8)
    assert np.all(np.isfinite(hist)) and l1 < 0.5 * l0
    p = net.get_params()
    net2 = NN.WarmStartNet(nels, max_batch=1, seed=99)
    net2.set_params(p)
    x = g.uniform(0.1, 1.0, net.ncells)
    f = NN.cantilever_load(nels, 4, 4)
    np.testing.assert_array_equal(net.predict(x, f), net2.predict(x, f))


def test_learned_warm_start_in_simp_loop(T, gpu):
    """warm_start = 2 runs the network inside the SIMP loop; the minimal-
    residual combination with the previous displacement never starts CG with a
    larger residual than the previous iterate, and the solves still converge to
    the same tolerance."""
    import ctypes as C
    import topopt_jl_amd.nn as NN
    from topopt_jl_amd._lib import SimpOpts, check, dptr
    nels = (16, 8, 8)
    ctx, fixed = NN.make_context(nels)
    net = NN.WarmStartNet(nels, max_batch=1, seed=5, fixed=fixed)
    assert ctx.lib.tohip_nn_attach(ctx.ptr, net.ptr) == 0
    f = NN.cantilever_load(nels, 4, 4)
    check(ctx.lib.tohip_set_load(ctx.ptr, dptr(f)), ctx.ptr, "load")
    res = {}
    for ws in (1, 2):
        x = np.full(ctx.ncells, 0.4)
        it = np.zeros(6, np.int64)
        obj = np.zeros(6)
        o = SimpOpts(iterations=6, volfrac=0.4, rmin=1.5, move=0.2, oc_tol=1e-12, rel_tol=1e-8,
                     maxiter=100000, precision=0, warm_start=ws, inner_rel_tol=0.1)
        ctx.lib.tohip_nn_attach(ctx.ptr, net.ptr)
        rc = ctx.lib.tohip_simp_run(ctx.ptr, C.byref(o), dptr(x), dptr(obj),
                                    it.ctypes.data_as(C.POINTER(C.c_int64)), None)
        assert rc == 0
        calls = C.c_int64(-1)
        assert ctx.lib.tohip_nn_calls(ctx.ptr, C.byref(calls)) == 0
        res[ws] = (obj, it, calls.value)
    np.testing.assert_allclose(res[1][0], res[2][0], rtol=1e-6)
    # an untrained network does not pay: after the first (unrated) prediction,
    # two rated misses in a row (or one solve that blows up) switch it off
    # for the rest of the run
    print(f"untrained network: {res[2][2]} predictions in 6 iterations", flush=True)
    assert res[1][2] == 0 and res[2][2] in (2, 3)
