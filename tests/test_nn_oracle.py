"""CPU tests pinning the network-layer oracle (oracle/nn_oracle.py): the
convolution against a direct loop, its backward pass against central finite
differences, SiLU' against differences, and pooling/upsampling adjointness."""
import numpy as np
import pytest

from oracle import nn_oracle as N


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159265, 0.0, 65504.0])
    r = N.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie 1 + 2^-8 rounds to even (1.0)
    assert r[2] == 1.0078125            # above the tie rounds up
    assert abs(r[3] - (-3.140625)) < 1e-12
    assert r[4] == 0.0


@pytest.mark.parametrize("k,pad", [(3, 1), (2, 1), (2, 0), (1, 0)])
def test_conv_fwd_matches_direct_loop(k, pad):
    g = np.random.default_rng(0)
    x = g.standard_normal((2, 3, 4, 3, 2))
    w = g.standard_normal((3, k ** 3, 2))
    b = g.standard_normal(3)
    z = N.conv_fwd(x, w, b, k, pad)
    B, D, H, W, _ = x.shape
    ref = np.zeros_like(z)
    for bb in range(B):
        for oz in range(z.shape[1]):
            for oy in range(z.shape[2]):
                for ox in range(z.shape[3]):
                    s = b.copy()
                    for t, (dz, dy, dx) in enumerate(N.taps(k)):
                        iz, iy, ix = oz + dz - pad, oy + dy - pad, ox + dx - pad
                        if 0 <= iz < D and 0 <= iy < H and 0 <= ix < W:
                            s += w[:, t, :] @ x[bb, iz, iy, ix]
                    ref[bb, oz, oy, ox] = s
    assert np.max(np.abs(z - ref)) < 1e-12


@pytest.mark.parametrize("k,pad", [(3, 1), (2, 1)])
def test_conv_bwd_finite_differences(k, pad):
    g = np.random.default_rng(1)
    x = g.standard_normal((1, 3, 3, 4, 2))
    w = g.standard_normal((2, k ** 3, 2))
    b = g.standard_normal(2)
    dz = g.standard_normal(N.conv_fwd(x, w, b, k, pad).shape)
    L = lambda xx, ww, bb: np.sum(N.conv_fwd(xx, ww, bb, k, pad) * dz)
    dx, dw, db = N.conv_bwd(x, w, dz, k, pad)
    h = 1e-6
    for arr, grad, idxs in [(x, dx, [(0, 1, 2, 3, 1), (0, 0, 0, 0, 0)]),
                            (w, dw, [(1, 5 % k ** 3, 0), (0, 0, 1)])]:
        for idx in idxs:
            a1, a2 = arr.copy(), arr.copy()
            a1[idx] += h
            a2[idx] -= h
            if arr is x:
                fd = (L(a1, w, b) - L(a2, w, b)) / (2 * h)
            else:
                fd = (L(x, a1, b) - L(x, a2, b)) / (2 * h)
            assert abs(fd - grad[idx]) < 1e-6 * max(1, abs(fd))
    b1 = b.copy()
    b1[1] += h
    b2 = b.copy()
    b2[1] -= h
    assert abs((L(x, w, b1) - L(x, w, b2)) / (2 * h) - db[1]) < 1e-6


def test_silu_grad_and_resampling_adjoints():
    z = np.linspace(-6, 6, 41)
    fd = (N.silu(z + 1e-6) - N.silu(z - 1e-6)) / 2e-6
    assert np.max(np.abs(fd - N.silu_grad(z))) < 1e-8
    g = np.random.default_rng(2)
    a = g.standard_normal((2, 4, 2, 6, 3))
    c = g.standard_normal((2, 2, 1, 3, 3))
    # <pool(a), c> = <a, pool^T(c)> with pool^T = upsample / 8
    assert abs(np.sum(N.avgpool(a) * c) - np.sum(a * N.upsample(c) / 8)) < 1e-12


def test_network_features_direct_loop():
    """Cell features against a per-cell loop over the 8 corner nodes."""
    nels = (3, 2, 2)
    nx, ny, nz = nels
    g = np.random.default_rng(4)
    nn_ = (nx + 1) * (ny + 1) * (nz + 1)
    x = g.uniform(0, 1, nx * ny * nz)
    f = g.standard_normal(3 * nn_)
    fixed = np.zeros(3 * nn_, np.uint8)
    fixed[3 * 5 + 1] = 1  # node 5 (i = 1, j = 1, k = 0) clamped in y
    feat = N.network_features(nels, x, f, fixed)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                e = i + nx * (j + ny * k)
                nodes = [(i + a) + (nx + 1) * ((j + b) + (ny + 1) * (k + c))
                         for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                assert feat[0, k, j, i, 0] == N.bf16(x[e])
                np.testing.assert_allclose(feat[0, k, j, i, 1:4],
                                           N.bf16(f.reshape(-1, 3)[nodes].sum(0)), rtol=1e-2)
                assert feat[0, k, j, i, 4] == (1.0 if 5 in nodes else 0.0)
                assert feat[0, k, j, i, 5] == 1.0 and not feat[0, k, j, i, 6:].any()
