"""TEST INFRASTRUCTURE ONLY — CPU oracle of the warm-start network layers.

numpy (float64) restatement of one stride-1, zero-padded 3D convolution
layer in NDHWC layout with its backward pass, plus SiLU, 2x2x2 average
pooling and nearest 2x upsampling, and the whole encoder-decoder forward
pass (`network_forward`).  `conv_layer_device_emulation` applies the
BF16 roundings at exactly the points where the device kernels store BF16
(inputs, weights, incoming gradient, pre-activation, dZ), so the GPU result
can be compared at FP32-accumulation accuracy.  The layer maths is pinned by
finite differences in tests/test_nn_oracle.py.  There is no reference
counterpart (TopOpt.jl has no convolutional network).
"""
from __future__ import annotations

import itertools

import numpy as np


def bf16(a):
    """Round float values to bfloat16 (nearest, ties to even), as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def taps(k):
    return list(itertools.product(range(k), repeat=3))  # (dz, dy, dx), dx fastest


def conv_fwd(x, w, b, k, pad):
    """x [B,D,H,W,Cin], w [Cout, k^3, Cin] -> z [B,OD,OH,OW,Cout]."""
    B, D, H, W, cin = x.shape
    cout = w.shape[0]
    xp = np.zeros((B, D + 2 * pad, H + 2 * pad, W + 2 * pad, cin))
    xp[:, pad:pad + D, pad:pad + H, pad:pad + W] = x
    OD, OH, OW = D + 2 * pad - k + 1, H + 2 * pad - k + 1, W + 2 * pad - k + 1
    z = np.zeros((B, OD, OH, OW, cout)) + np.asarray(b, np.float64)
    for t, (dz, dy, dx) in enumerate(taps(k)):
        z += xp[:, dz:dz + OD, dy:dy + OH, dx:dx + OW] @ w[:, t, :].T
    return z


def conv_bwd(x, w, dz, k, pad):
    """Gradients of conv_fwd w.r.t. x, w, b for output gradient dz."""
    B, D, H, W, cin = x.shape
    OD, OH, OW = dz.shape[1:4]
    xp = np.zeros((B, D + 2 * pad, H + 2 * pad, W + 2 * pad, cin))
    xp[:, pad:pad + D, pad:pad + H, pad:pad + W] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w, dtype=np.float64)
    for t, (ddz, ddy, ddx) in enumerate(taps(k)):
        sl = xp[:, ddz:ddz + OD, ddy:ddy + OH, ddx:ddx + OW]
        dw[:, t, :] = np.einsum("bzyxo,bzyxi->oi", dz, sl)
        dxp[:, ddz:ddz + OD, ddy:ddy + OH, ddx:ddx + OW] += dz @ w[:, t, :]
    db = dz.sum(axis=(0, 1, 2, 3))
    return dxp[:, pad:pad + D, pad:pad + H, pad:pad + W], dw, db


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def avgpool(x):
    B, D, H, W, C = x.shape
    return x.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def upsample(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def conv_layer_device_emulation(x, w, b, dy, k, pad, act):
    """Forward + backward of one layer with the device's BF16 storage points."""
    xb, wb, dyb = bf16(x), bf16(w), bf16(dy)
    z = conv_fwd(xb, wb, b, k, pad)
    y = silu(z) if act else z
    if act:
        dz = bf16(dyb * silu_grad(bf16(z)))
    else:
        dz = dyb
    dx, dw, db = conv_bwd(xb, wb, dz, k, pad)
    return y, dx, dw, db


# Layer table of the device network (cin, cout, k, pad); input channels are
# padded to 8 (6 features used), the node output layer has 8 channels (3 used).
NET_LAYERS = [(8, 16, 3, 1), (16, 16, 3, 1), (16, 32, 3, 1), (32, 32, 3, 1), (32, 64, 3, 1),
              (64, 64, 3, 1), (64, 128, 3, 1), (128, 128, 3, 1), (192, 64, 3, 1), (64, 64, 3, 1),
              (96, 32, 3, 1), (32, 32, 3, 1), (48, 16, 3, 1), (16, 16, 3, 1), (16, 8, 2, 1)]


def network_features(nels, xphys, f, fixed_dofs):
    """Cell features [1, nz, ny, nx, 8]: density, the load summed over the
    cell's 8 nodes (float32, corner order), 1 if a node is clamped, 1, 0, 0."""
    nx, ny, nz = nels
    NX, NY = nx + 1, ny + 1
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f3 = np.asarray(f, np.float64).reshape(-1, 3).astype(np.float32)
    clampn = np.asarray(fixed_dofs).reshape(-1, 3).any(axis=1)
    fs = np.zeros((nz, ny, nx, 3), np.float32)
    cl = np.zeros((nz, ny, nx), bool)
    for c in range(8):
        n = (i + (c & 1)) + NX * ((j + ((c >> 1) & 1)) + NY * (k + ((c >> 2) & 1)))
        fs += f3[n]
        cl |= clampn[n]
    feat = np.zeros((1, nz, ny, nx, 8))
    feat[0, ..., 0] = np.asarray(xphys, np.float64).reshape(nz, ny, nx).astype(np.float32)
    feat[0, ..., 1:4] = fs
    feat[0, ..., 4] = cl
    feat[0, ..., 5] = 1.0
    return bf16(feat)


def network_forward(params, nels, xphys, f, fixed_dofs):
    """The encoder-decoder forward pass with BF16 storage of every activation
    (FP32 parameters rounded to BF16), float64 accumulation; returns the
    node output [nnodes, 3] (FP32 on the device)."""
    params = np.asarray(params, np.float32)
    W, Bs, off = [], [], 0
    for cin, cout, k, _ in NET_LAYERS:
        n = cout * k ** 3 * cin
        W.append(bf16(params[off:off + n].reshape(cout, k ** 3, cin)))
        off += n
        Bs.append(params[off:off + cout].astype(np.float64))
        off += cout
    assert off == params.size

    def conv(li, x, act=True):
        _, _, k, pad = NET_LAYERS[li]
        z = conv_fwd(x, W[li], Bs[li], k, pad)
        # the device stores the BF16 output of the FP32 result
        z = z.astype(np.float32).astype(np.float64)
        return bf16(silu(z)) if act else z

    pool = lambda x: bf16(avgpool(x))
    x = network_features(nels, xphys, f, fixed_dofs)
    a00 = conv(0, x)
    s0 = conv(1, a00)
    a10 = conv(2, pool(s0))
    s1 = conv(3, a10)
    a20 = conv(4, pool(s1))
    s2 = conv(5, a20)
    e3 = conv(7, conv(6, pool(s2)))
    d2 = conv(9, conv(8, np.concatenate([upsample(e3), s2], axis=-1)))
    d1 = conv(11, conv(10, np.concatenate([upsample(d2), s1], axis=-1)))
    d0 = conv(13, conv(12, np.concatenate([upsample(d1), s0], axis=-1)))
    out = conv(14, d0, act=False)
    return out[0, ..., :3].reshape(-1, 3)
