"""ctypes binding of ``libvxfea.so`` (the HIP solver's C ABI, include/vxfea.h).

The library is built in-tree by ``voxelsim-fea_amd/build.py`` into
``voxelsim-fea_amd/lib/libvxfea.so``.  There is no fallback: if the library
is missing or no GPU is visible, the solver raises.
"""

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "lib", "libvxfea.so")

VX_OK = 0
VX_ERR_ARG = -1
VX_ERR_NOMEM = -2
VX_ERR_HIP = -3
VX_ERR_STALL = -4
VX_ERR_NOT_CONVERGED = -5
VX_ERR_NO_DEVICE = -6

VX_FP64, VX_MIXED = 0, 1
VX_SMOOTH_JACOBI, VX_SMOOTH_CHEBYSHEV = 0, 1
VX_COARSE_SWEEPS, VX_COARSE_DIRECT = 0, 1

_STATUS = {
    VX_ERR_ARG: "bad argument",
    VX_ERR_NOMEM: "device out of memory",
    VX_ERR_HIP: "HIP runtime error",
    VX_ERR_STALL: "MG-PCG stall guard",
    VX_ERR_NOT_CONVERGED: "not converged",
    VX_ERR_NO_DEVICE: "no GPU device",
}


class vx_options(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_int32),
        ("max_levels", ctypes.c_int32),
        ("min_elems", ctypes.c_int32),
        ("n_smooth", ctypes.c_int32),
        ("omega", ctypes.c_double),
        ("coarse_sweeps", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("smoother", ctypes.c_int32),
        ("coarse_mode", ctypes.c_int32),
        ("cheb_ratio_power", ctypes.c_int32),
        ("cheb_ratio", ctypes.c_double),
        ("inner_tol", ctypes.c_double),
        ("max_outer", ctypes.c_int32),
        ("z_chunk", ctypes.c_int32),
        ("e_coarsen", ctypes.c_int32),
        ("max_batch", ctypes.c_int32),
        ("flexible", ctypes.c_int32),
        ("level_smooth", ctypes.c_uint8 * 16),
        ("eager", ctypes.c_int32),
        ("block_steps", ctypes.c_int32),
        ("block_edge", ctypes.c_int32),
        ("keep_direction", ctypes.c_int32),
        ("keep_drift", ctypes.c_double),
    ]


class vx_info(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("outer_iters", ctypes.c_int32),
        ("n_levels", ctypes.c_int32),
        ("resid_ratio", ctypes.c_double),
        ("true_resid_ratio", ctypes.c_double),
        ("solve_ms", ctypes.c_double),
        ("corrector", ctypes.c_int32),
        ("coarse_direct", ctypes.c_int32),
    ]


PROGRESS_CB = ctypes.CFUNCTYPE(None, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                               ctypes.c_void_p)

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double

# name -> (restype, argtypes); every symbol include/vxfea.h declares
SIGNATURES = {
    "vx_default_options": (None, [ctypes.POINTER(vx_options)]),
    "vx_device_count": (ctypes.c_int, []),
    "vx_create": (ctypes.c_int, [_I32, _I32, _I32, _D, ctypes.POINTER(vx_options),
                                 ctypes.POINTER(_P)]),
    "vx_destroy": (None, [_P]),
    "vx_last_error": (ctypes.c_char_p, [_P]),
    "vx_set_stream": (ctypes.c_int, [_P, _P]),
    "vx_num_levels": (ctypes.c_int, [_P]),
    "vx_level_dims": (ctypes.c_int, [_P, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32),
                                     ctypes.POINTER(_I32)]),
    "vx_set_fixed": (ctypes.c_int, [_P, _P, _I64]),
    "vx_set_density": (ctypes.c_int, [_P, _P, _P, _D]),
    "vx_solve": (ctypes.c_int, [_P, _P, _P, _P, _D, _I32, PROGRESS_CB, _P, _P, _I32,
                                ctypes.POINTER(vx_info)]),
    "vx_solve_batch": (ctypes.c_int, [_P, _I32, _P, _P, _D, _I32, PROGRESS_CB, _P, _P, _I32,
                                      ctypes.POINTER(vx_info)]),
    "vx_solve_jacobi": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _D, _I32, PROGRESS_CB, _P, _P,
                                       _I32, ctypes.POINTER(vx_info)]),
    "vx_parts_create": (ctypes.c_int, [_I32, _P, _D, ctypes.POINTER(vx_options),
                                       ctypes.POINTER(_P)]),
    "vx_parts_destroy": (None, [_P]),
    "vx_parts_last_error": (ctypes.c_char_p, [_P]),
    "vx_parts_set_stream": (ctypes.c_int, [_P, _P]),
    "vx_parts_num_levels": (ctypes.c_int, [_P, _I32]),
    "vx_parts_set_fixed": (ctypes.c_int, [_P, _P, _P]),
    "vx_parts_set_density": (ctypes.c_int, [_P, _P, _P, _D]),
    "vx_parts_autotune": (ctypes.c_int, [_P, _I32]),
    "vx_parts_zchunk": (ctypes.c_int, [_P, _I32, _I32]),
    "vx_solve_parts": (ctypes.c_int, [_P, _P, _P, _D, _I32, PROGRESS_CB, _P, _P, _I32,
                                      ctypes.POINTER(vx_info)]),
    "vx_matvec": (ctypes.c_int, [_P, _I32, _I32, _P, _P]),
    "vx_smooth": (ctypes.c_int, [_P, _I32, _I32, _P, _P, _I32, _P]),
    "vx_restrict": (ctypes.c_int, [_P, _I32, _P, _P]),
    "vx_prolong": (ctypes.c_int, [_P, _I32, _P, _P]),
    "vx_vcycle": (ctypes.c_int, [_P, _I32, _P, _P]),
    "vx_inv_diag": (ctypes.c_int, [_P, _I32, _P]),
    "vx_von_mises": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "vx_strain_energy": (ctypes.c_int, [_P, _P, _P, _P]),
    "vx_post_fields": (ctypes.c_int, [_P, _P, _P, _P, _D, _D, _P, _P, _P, _P]),
    "vx_drop_islands": (ctypes.c_int, [_P, _P, _P, _I64, _P, ctypes.POINTER(_I64)]),
    "vx_drop_islands_noalloc": (ctypes.c_int, [_P, _P, _P, _I64, _P, ctypes.POINTER(_I64)]),
    "vx_time_kernel": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, ctypes.POINTER(_D)]),
    "vx_element_apply_host": (None, [_D, _P, _P]),
    "vx_autotune": (ctypes.c_int, [_P, _I32, _I32]),
    "vx_tuned_zchunk": (ctypes.c_int, [_P, _I32, _I32, _I32]),
    "vx_block_plan": (ctypes.c_int, [_P, _I32, _I32, _P, _P]),
    "vx_set_corrector": (ctypes.c_int, [_P, _P, _I32, ctypes.c_float, _I32, _I32]),
    "vx_corrector_eval": (ctypes.c_int, [_P, _I32, _P, _P, _P]),
    "vx_capture": (ctypes.c_int, [_P, _I32, _P, _I32, _I32]),
    "vx_capture_count": (ctypes.c_int, [_P]),
    "vx_corrector_prepare": (ctypes.c_int, [_P, _I32, _P, _P, _P, _P]),
    "vx_corrector_train": (ctypes.c_int, [_P, _I32, _P, _P, _P, _P, _I32, _I32, ctypes.c_float,
                                          ctypes.c_uint64, _P]),
    "vx_inside_mask": (ctypes.c_int, [_P, _P, _I64, _P, _I64, _P, _I64, _P, _P]),
    "vx_inside_grid": (ctypes.c_int, [_P, _P, _I64, _P, _I64, _I32, _I32, _I32, _P, _D, _I32,
                                      _P, _P]),
    "vx_inside_last_error": (ctypes.c_char_p, []),
    "vxn_create": (ctypes.c_int, [_I32, _I32, _I32, ctypes.POINTER(_P)]),
    "vxn_destroy": (None, [_P]),
    "vxn_last_error": (ctypes.c_char_p, [_P]),
    "vxn_set_stream": (ctypes.c_int, [_P, _P]),
    "vxn_num_params": (_I64, [_P]),
    "vxn_set_params": (ctypes.c_int, [_P, _P]),
    "vxn_get_params": (ctypes.c_int, [_P, _P]),
    "vxn_forward": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _P, _P]),
    "vxn_backward": (ctypes.c_int, [_P, _P, _P]),
    "vxn_adam_step": (ctypes.c_int, [_P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, _I32]),
}

_lib = None


def load():
    """Load the library once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    # torch first, so the HIP runtime it ships is the one this library binds
    # to (both carry the soname libamdhip64.so.7): one runtime per process.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libvxfea.so not found at {LIB_PATH}: build it with "
            f"`python voxelsim-fea_amd/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def default_options(**kw):
    o = vx_options()
    load().vx_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if v is None:
            continue
        if not hasattr(o, k):
            raise TypeError(f"unknown solver option {k!r}")
        setattr(o, k, v)
    return o


class NativeError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"vxfea: {_STATUS.get(rc, rc)}: {msg}")
        self.rc = rc


class StallError(NativeError):
    """The MG-PCG stall/divergence guard fired (core/multigrid.py:524-531)."""


def check(rc, handle=None, allow=()):
    if rc == VX_OK or rc in allow:
        return rc
    msg = ""
    if handle is not None:
        m = load().vx_last_error(handle)
        msg = m.decode() if m else ""
    if rc == VX_ERR_STALL:
        raise StallError(rc, msg)
    raise NativeError(rc, msg)


def check_net(rc, handle):
    """Status check for the network handle (vxn_*)."""
    if rc == VX_OK:
        return rc
    m = load().vxn_last_error(handle) if handle is not None else b""
    raise NativeError(rc, m.decode() if m else "")


def ptr(t):
    """Raw device pointer of a torch tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())
