"""Compute-backend selection, MI355X edition (API of core/backend.py:14-45 and
core/backend_gpu.py:20-108).

``GPU_BUILD`` is True.  ``gpu_usable()`` loads the HIP library, counts devices
and runs a self-test on the device -- the operator of a 2x1x1 grid against
the library's host copy of the element operator, the analogue of the
reference's matmul/bincount CPU-vs-GPU check (core/backend_gpu.py:40-60).
There is no NumPy fallback for the solve: when no GPU is usable the solver
raises instead of computing somewhere else.
"""

import numpy as np

GPU_BUILD = True

_GPU_OK = None
_SELFTEST_ERR = ""


def is_gpu_build():
    return GPU_BUILD


def _lib():
    try:
        from . import _native
        return _native.load()
    except Exception:
        return None


def gpu_device_count():
    lib = _lib()
    if lib is None:
        return 0
    try:
        return int(lib.vx_device_count())
    except Exception:
        return 0


def _self_test():
    global _SELFTEST_ERR
    if gpu_device_count() < 1:
        _SELFTEST_ERR = "no device"
        return False
    try:
        import ctypes
        import torch
        from . import _native
        lib = _native.load()
        nx, ny, nz, nu = 2, 1, 1, 0.3
        h = ctypes.c_void_p()
        opt = _native.default_options(max_levels=1)
        _native.check(lib.vx_create(nx, ny, nz, nu, ctypes.byref(opt), ctypes.byref(h)))
        try:
            lib.vx_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
            _native.check(lib.vx_set_fixed(h, None, 0), h)
            E = torch.tensor([1.0, 2.0], dtype=torch.float64, device="cuda")
            _native.check(lib.vx_set_density(h, _native.ptr(E), None, 0.0), h)
            rng = np.random.default_rng(0)
            u = rng.standard_normal(3 * 12)
            ud = torch.tensor(u, device="cuda")
            yd = torch.empty_like(ud)
            _native.check(lib.vx_matvec(h, 0, _native.VX_FP64, _native.ptr(ud), _native.ptr(yd)), h)
            y = yd.cpu().numpy()
        finally:
            lib.vx_destroy(h)
        # host assembly from the library's own host element operator
        ref = np.zeros_like(u)
        offs = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
        for ex, Ee in ((0, 1.0), (1, 2.0)):
            dof = []
            for ox, oy, oz in offs:
                n = (ex + ox) + 3 * (oy + 2 * oz)
                dof += [3 * n, 3 * n + 1, 3 * n + 2]
            ue = np.ascontiguousarray(u[dof])
            ye = np.zeros(24)
            lib.vx_element_apply_host(nu, ue.ctypes.data, ye.ctypes.data)
            np.add.at(ref, dof, Ee * ye)
        ok = bool(np.allclose(y, ref, rtol=1e-12, atol=1e-12))
        if not ok:
            _SELFTEST_ERR = f"operator mismatch: max diff {np.abs(y - ref).max():.3e}"
        return ok
    except Exception as exc:
        _SELFTEST_ERR = f"{type(exc).__name__}: {exc}"
        return False


def gpu_usable():
    """Cached: library present, a device visible and the self-test passed."""
    global _GPU_OK
    if _GPU_OK is None:
        _GPU_OK = _self_test()
    return _GPU_OK


class _TorchXP:
    """Minimal array-module shim over torch device tensors (float64)."""

    float64 = np.float64

    @staticmethod
    def asarray(a, dtype=None):
        import torch
        t = torch.as_tensor(np.asarray(a) if not hasattr(a, "device") else a, device="cuda")
        return t.to(torch.float64) if dtype in (float, np.float64) else t

    @staticmethod
    def zeros(n, dtype=float):
        import torch
        return torch.zeros(n, dtype=torch.float64, device="cuda")

    @staticmethod
    def zeros_like(a):
        import torch
        return torch.zeros_like(a)

    @staticmethod
    def where(c, a, b):
        import torch
        return torch.where(c, a, b)

    @staticmethod
    def asnumpy(a):
        return a.detach().cpu().numpy()


torch_xp = _TorchXP()


def get_xp(use_gpu=False):
    """torch-backed shim when the GPU is requested and usable, else numpy."""
    if use_gpu and gpu_usable():
        return torch_xp
    return np


def asnumpy(a):
    """Bring an array back to host numpy (torch tensors or numpy)."""
    if type(a).__module__.split(".")[0] == "torch":
        return a.detach().cpu().numpy()
    return np.asarray(a)


def gpu_status():
    if _lib() is None:
        return "GPU: libvxfea.so not built"
    if gpu_device_count() < 1:
        return "GPU: no HIP device visible"
    if gpu_usable():
        return "GPU: MI355X HIP solver ready (validated)"
    return f"GPU: self-test failed ({_SELFTEST_ERR})"
