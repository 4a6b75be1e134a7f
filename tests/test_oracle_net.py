"""Pins of the network oracle (oracle/convnet_ref.py): convolution against a
direct loop, prolongation against its 1-D matrices, adjoint identities, and
the float64 backward against central finite differences."""

import numpy as np
import pytest

from oracle import convnet_ref as N


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("stride", [1, 2])
def test_conv3_direct_loop(stride):
    x = _rand((2, 5, 4, 6, 3), 0)            # B, Z, Y, X, C
    W = _rand((27, 3, 2), 1)
    b = _rand(2, 2)
    dims_in = (6, 4, 5)
    out_dims = dims_in if stride == 1 else tuple((n + 1) // 2 for n in dims_in)
    y = N.conv3(x, W, b, stride, out_dims)
    ref = np.zeros_like(y)
    Xo, Yo, Zo = out_dims
    for bb in range(2):
        for z in range(Zo):
            for yy in range(Yo):
                for xx in range(Xo):
                    acc = b.copy()
                    for dz in range(3):
                        for dy in range(3):
                            for dx in range(3):
                                sz, sy, sx = stride * z + dz - 1, stride * yy + dy - 1, stride * xx + dx - 1
                                if 0 <= sz < 5 and 0 <= sy < 4 and 0 <= sx < 6:
                                    acc = acc + x[bb, sz, sy, sx] @ W[dx + 3 * (dy + 3 * dz)]
                    ref[bb, z, yy, xx] = acc
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("stride", [1, 2])
def test_conv3_adjoints(stride):
    x = _rand((2, 7, 5, 6, 4), 3)
    W = _rand((27, 4, 3), 4)
    dims_in = (6, 5, 7)
    out_dims = dims_in if stride == 1 else tuple((n + 1) // 2 for n in dims_in)
    g = _rand((2, out_dims[2], out_dims[1], out_dims[0], 3), 5)
    lhs = np.sum(N.conv3(x, W, 0.0, stride, out_dims) * g)
    assert np.isclose(lhs, np.sum(x * N.conv3_adj_data(g, W, stride, dims_in)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(W * N.conv3_adj_weight(x, g, stride)), rtol=1e-12)


def test_up_matrices_and_adjoint():
    for n in (5, 6, 9):
        P = N.interp_matrix(n)
        # reproduces linear functions where both neighbours exist, injection at even nodes
        xc = np.arange(P.shape[1], dtype=float)
        xf = P @ (2 * xc)
        full = [i for i in range(n) if i % 2 == 0 or (i + 1) // 2 < P.shape[1]]
        assert np.allclose(xf[full], np.arange(n)[full])
    c = _rand((1, 3, 4, 5, 2), 6)
    fine = (9, 7, 6)       # X, Y, Z with coarse (5, 4, 3)
    f = N.up(c, fine)
    Px, Py, Pz = (N.interp_matrix(k) for k in fine)
    K = np.kron(Pz, np.kron(Py, Px))                 # x fastest
    ref = (K @ c[0].reshape(-1, 2)).reshape(f.shape[1:])
    assert np.allclose(f[0], ref)
    g = _rand(f.shape, 7)
    assert np.isclose(np.sum(f * g), np.sum(c * N.up_adj(g, (5, 4, 3))), rtol=1e-12)


def test_bf16_rounding():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], np.float32)
    r = N.bf16(a)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1 + 2 ** -6) and r[3] == -2.5
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)


def test_unet_backward_finite_differences():
    T = N.layer_table(c0=16, c1=16, c2=16, c3=16, cin=4, cout=3)
    net = N.UNetRef(T)
    rng = np.random.default_rng(8)
    params = N.init_params(T, seed=9, out_scale=1.0).astype(np.float64)
    params += 0.01 * rng.standard_normal(params.size)    # non-zero biases
    x = rng.standard_normal((1, 9, 9, 9, 4))
    dout = rng.standard_normal((1, 9, 9, 9, 3))
    net.forward(params, x, bf=False)
    g = net.backward(dout)
    sizes = [27 * ci * co + co for _, _, ci, co, _, _ in T]
    starts = np.cumsum([0] + sizes[:-1])
    idx = [int(s + k) for s in starts for k in (0, 5)] + [int(starts[-1] + sizes[-1] - 1)]
    eps = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (np.sum(net.forward(p1, x, bf=False) * dout) -
              np.sum(net.forward(p2, x, bf=False) * dout)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])
