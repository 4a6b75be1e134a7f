"""GPU tests of the corrector network kernels against the CPU oracle
(oracle/convnet_ref.py, which emulates the same BF16 rounding points)."""

import numpy as np
import pytest

from oracle import convnet_ref as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("chans,dims,B", [((16, 32, 64), (9, 7, 6), 2),
                                          ((16, 16, 16), (11, 10, 12), 1),
                                          ((32, 64, 64), (8, 9, 5), 1)])
def test_net_forward_backward_vs_oracle(torch_cuda, chans, dims, B):
    torch = torch_cuda
    from core.learned import CorrectorNet
    print(f"net fwd/bwd chans={chans} dims={dims} B={B}", flush=True)
    net = CorrectorNet(*chans, seed=1, out_scale=1.0)
    T = R.layer_table(chans[0], chans[1], chans[2], chans[2], 16, 16)
    rng = np.random.default_rng(2)
    p = net.params().cpu().numpy()
    p[np.arange(p.size) % 7 == 3] += 0.05          # non-zero biases everywhere
    net.set_params(torch.from_numpy(p).cuda())
    X, Y, Z = dims
    x = rng.standard_normal((B, Z, Y, X, 16)).astype(np.float32)
    x[..., 5:] = 0.0
    dout = rng.standard_normal((B, Z, Y, X, 16)).astype(np.float32)
    dout[..., 3:] = 0.0
    ref = R.UNetRef(T)
    y_ref = ref.forward(p, x, bf=True)
    g_ref = ref.backward(dout)
    y = net.forward(torch.from_numpy(x).cuda())
    g = net.backward(torch.from_numpy(dout).cuda()).cpu().numpy()
    y = y.cpu().numpy()
    print(f"  forward rel err {_rel(y, y_ref):.2e}, gradient rel err {_rel(g, g_ref):.2e}")
    assert _rel(y, y_ref) < 1e-2, _rel(y, y_ref)
    # per-layer weight and bias gradients
    o = 0
    for name, _, ci, co, _, _ in T:
        nW = 27 * ci * co
        for part, n in (("W", nW), ("b", co)):
            e = _rel(g[o:o + n], g_ref[o:o + n])
            assert e < 3e-2, (name, part, e)
            o += n


def test_adam_step(torch_cuda):
    torch = torch_cuda
    from core.learned import CorrectorNet
    net = CorrectorNet(16, 16, 16, seed=3)
    p0 = net.params().cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(4)
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    p = p0.copy()
    for step in (1, 2):
        g = rng.standard_normal(p0.size).astype(np.float32)
        net.grad.copy_(torch.from_numpy(g))
        net.adam(lr=1e-2)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g.astype(np.float64) ** 2
        p -= 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    assert np.allclose(net.params().cpu().numpy(), p, rtol=1e-5, atol=1e-6)
