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
