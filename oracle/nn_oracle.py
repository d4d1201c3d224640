"""TEST INFRASTRUCTURE ONLY — CPU oracle of the warm-start network layers.

numpy (float64) restatement of one stride-1, zero-padded 3D convolution
layer in NDHWC layout with its backward pass, plus SiLU, 2x2x2 average
pooling and nearest 2x upsampling.  `conv_layer_device_emulation` applies the
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
