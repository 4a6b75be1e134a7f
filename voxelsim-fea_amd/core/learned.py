"""Learned coarse-grid corrector: the network handle (C ABI ``vxn_*``).

The network is a 4-level 3D U-Net on a node grid with BF16 activations and
FP32 accumulation, run and trained by the library's own HIP kernels.  It has
no counterpart in the reference, whose V-cycle (core/multigrid.py:426-436)
has no learned part.  Tensors are torch device tensors ``[B, Z, Y, X, 16]``
float32 (voxel = x + X*(y + Y*z), channels innermost).
"""

import ctypes

import numpy as np

from . import _native

# (name, level, cin, cout, stride, relu) -- the library's layer order
def layer_table(c0=16, c1=32, c2=64):
    return [
        ("enc0a", 0, 16, c0, 1, True), ("enc0b", 0, c0, c0, 1, True),
        ("down1", 1, c0, c1, 2, True), ("enc1", 1, c1, c1, 1, True),
        ("down2", 2, c1, c2, 2, True), ("enc2", 2, c2, c2, 1, True),
        ("down3", 3, c2, c2, 2, True), ("enc3", 3, c2, c2, 1, True),
        ("dec2", 2, c2, c2, 1, True), ("dec2c", 2, c2, c1, 1, True),
        ("dec1", 1, c1, c1, 1, True), ("dec1c", 1, c1, c0, 1, True),
        ("dec0", 0, c0, c0, 1, True), ("out", 0, c0, 16, 1, False),
    ]


def he_init(table, seed=0, out_scale=0.1, n_in=16):
    """He-normal weights (fan-in 27*Cin; the first layer counts its n_in used
    input channels), zero biases, the linear output layer scaled down."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, (_, _, ci, co, _, relu) in enumerate(table):
        fan = 27 * (n_in if i == 0 else ci)
        std = np.sqrt(2.0 / fan) * (1.0 if relu else out_scale)
        parts.append(rng.standard_normal(27 * ci * co) * std)
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


class CorrectorNet:
    """Owns a ``vxn_handle``: parameters, Adam moments and activations."""

    def __init__(self, c0=16, c1=32, c2=64, seed=0, n_in=16, out_scale=0.1):
        import torch
        self._lib = _native.load()
        self.channels = (c0, c1, c2)
        self.table = layer_table(c0, c1, c2)
        h = ctypes.c_void_p()
        rc = self._lib.vxn_create(c0, c1, c2, ctypes.byref(h))
        if rc != _native.VX_OK:
            raise _native.NativeError(rc, "vxn_create")
        self._h = h
        self.n_params = int(self._lib.vxn_num_params(h))
        self._bind()
        self.grad = torch.zeros(self.n_params, dtype=torch.float32, device="cuda")
        self.step = 0
        self.set_params(torch.from_numpy(he_init(self.table, seed, out_scale, n_in)).cuda())

    def _bind(self):
        import torch
        self._lib.vxn_set_stream(self._h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    def _check(self, rc):
        return _native.check_net(rc, self._h)

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.vxn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_params(self, p):
        assert p.numel() == self.n_params and p.dtype.is_floating_point
        self._bind()
        self._check(self._lib.vxn_set_params(self._h, _native.ptr(p.float().contiguous())))
        self.step = 0

    def params(self):
        import torch
        out = torch.empty(self.n_params, dtype=torch.float32, device="cuda")
        self._bind()
        self._check(self._lib.vxn_get_params(self._h, _native.ptr(out)))
        return out

    def forward(self, x):
        """x [B, Z, Y, X, 16] float32 (device) -> y of the same shape."""
        import torch
        assert x.dim() == 5 and x.shape[4] == 16 and x.dtype == torch.float32 and x.is_cuda
        x = x.contiguous()
        y = torch.empty_like(x)
        B, Z, Y, X, _ = x.shape
        self._bind()
        self._check(self._lib.vxn_forward(self._h, B, X, Y, Z, _native.ptr(x), _native.ptr(y)))
        return y

    def backward(self, dout):
        """Parameter gradient of sum(y * dout) for the last forward (into self.grad)."""
        self._bind()
        self._check(self._lib.vxn_backward(self._h, _native.ptr(dout.contiguous()),
                                           _native.ptr(self.grad)))
        return self.grad

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        self._bind()
        self._check(self._lib.vxn_adam_step(self._h, _native.ptr(self.grad), lr, beta1, beta2,
                                            eps, self.step))


def train_corrector(mg, f_dev, level, net=None, n_samples=64, batch=4, steps=600, lr=3e-4,
                    seed=0, tol=1e-6, channels=(16, 32, 64), every=None, out_scale=0.1):
    """Train a corrector for ``mg`` (precision "mixed", density already set)
    on residual -> correction pairs from its own plain MG-PCG:

    1. a plain solve of ``f_dev`` (corrector off) counts the V-cycles;
    2. a second plain solve records ``n_samples`` level residuals t spread
       over the whole solve (every k-th V-cycle);
    3. the plain V-cycle's coarse-grid correction v of each t (the pair) and
       the residual it leaves, r = t - A v;
    4. Adam, on the GPU, on the energy error of the correction the V-cycle
       applies (the best combination of v and the network's y).

    The trained net is attached (enabled) to ``mg`` with the plain solve's
    iteration count as its budget; returns (net, report)."""
    import time
    import torch
    net = net or CorrectorNet(*channels, seed=seed, out_scale=out_scale)
    mg.detach_corrector()
    u = torch.empty_like(f_dev)
    t0 = time.perf_counter()
    mg.solve_device(f_dev, u, tol=tol)
    plain_iters = mg.last_iters
    if every is None:
        every = max(1, plain_iters // n_samples)
    T = mg.capture(level, n_samples, every)
    mg.solve_device(f_dev, u, tol=tol)
    mg.stop_capture()
    n = mg.capture_count()
    T = T[:n].contiguous()
    t1 = time.perf_counter()
    mg.attach_corrector(net, level, enabled=False)
    V, R, vt = mg.corrector_pairs(T)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    loss = mg.corrector_train(T, V, R, vt, batch=batch, steps=steps, lr=lr, seed=seed + 1)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    mg.attach_corrector(net, level, enabled=True, iter_budget=plain_iters)
    report = {"level": level, "samples": n, "every": every, "plain_iters": plain_iters,
              "loss_first": float(loss[:10].mean()), "loss_last": float(loss[-20:].mean()),
              "capture_s": t1 - t0, "pairs_s": t2 - t1, "train_s": t3 - t2,
              "steps": steps, "batch": batch, "lr": lr}
    return net, report
