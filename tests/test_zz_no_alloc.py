"""vx_drop_islands_noalloc (the allocation-free form of vx_drop_islands)
against vx_drop_islands, through the C ABI: two calls in a row on the same
GPU, the second on a smaller grid (its scratch memory still holds the first
call's flags), bit-exact kept masks and equal removal counts; and the
argument checks that need no kernel (grid above the limit, fixed DOF out of
range)."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _handle(dims):
    import torch
    from core import _native
    lib = _native.load()
    h = ctypes.c_void_p()
    opt = _native.default_options(max_levels=1)
    _native.check(lib.vx_create(*dims, 0.3, ctypes.byref(opt), ctypes.byref(h)))
    lib.vx_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return lib, h


def _both(active3d, fixed):
    """(kept, n_removed) from vx_drop_islands and from the no-alloc form."""
    import torch
    from core import _native
    dims = active3d.shape
    lib, h = _handle(dims)
    try:
        a = torch.tensor(active3d.ravel().astype(np.uint8), device="cuda")
        fh = np.ascontiguousarray(fixed, dtype=np.int64)
        fd = torch.tensor(fh, device="cuda")
        k1 = torch.empty_like(a)
        k2 = torch.full_like(a, 7)          # stale output must not leak through
        n1, n2 = ctypes.c_int64(-1), ctypes.c_int64(-1)
        _native.check(lib.vx_drop_islands(h, _native.ptr(a), fh.ctypes.data_as(ctypes.c_void_p),
                                          fh.size, _native.ptr(k1), ctypes.byref(n1)), h)
        _native.check(lib.vx_drop_islands_noalloc(h, _native.ptr(a), _native.ptr(fd), fh.size,
                                                  _native.ptr(k2), ctypes.byref(n2)), h)
        return (k1.cpu().numpy(), n1.value), (k2.cpu().numpy(), n2.value)
    finally:
        lib.vx_destroy(h)


def _case(seed, dims, frac):
    r = np.random.default_rng(seed)
    act = r.random(dims) < frac
    nx, ny, _ = dims
    fixed = np.arange(3 * (nx + 1) * (ny + 1))          # the z = 0 node plane
    return act, fixed


def test_noalloc_equals_allocating_form_two_calls_shrinking():
    print("case noalloc islands: 48x40x36 then 20x13x9", flush=True)
    from oracle.fea_ref import drop_unsupported_islands as ref_drop
    for seed, dims in ((1, (48, 40, 36)), (2, (20, 13, 9))):
        act, fixed = _case(seed, dims, 0.42)
        (k1, n1), (k2, n2) = _both(act, fixed)
        assert n1 == n2 and np.array_equal(k1, k2), dims
        kr, nr = ref_drop(act, fixed)
        assert n2 == nr and np.array_equal(k2.astype(bool).reshape(dims), kr)


def test_noalloc_nothing_supported_leaves_part():
    print("case noalloc islands: unsupported part", flush=True)
    act = np.zeros((6, 6, 6), dtype=bool)
    act[2:4, 2:4, 3:5] = True
    (k1, n1), (k2, n2) = _both(act, np.arange(3 * 49))
    assert n1 == n2 == 0 and np.array_equal(k1, k2)
    assert np.array_equal(k2.astype(bool), act.ravel())


def test_noalloc_argument_errors():
    print("case noalloc islands: argument errors", flush=True)
    import torch
    from core import _native
    # a fixed DOF id past the grid
    lib, h = _handle((4, 4, 4))
    try:
        a = torch.ones(64, dtype=torch.uint8, device="cuda")
        k = torch.empty_like(a)
        bad = torch.tensor([0, 3 * 125], dtype=torch.int64, device="cuda")
        n = ctypes.c_int64()
        rc = lib.vx_drop_islands_noalloc(h, _native.ptr(a), _native.ptr(bad), 2, _native.ptr(k),
                                         ctypes.byref(n))
        assert rc == _native.VX_ERR_ARG and b"out of range" in lib.vx_last_error(h)
    finally:
        lib.vx_destroy(h)
    # a grid above the scratch limit (331^3 = 36,264,691 nodes > 2^25) is
    # refused before any kernel runs
    lib, h = _handle((330, 330, 330))
    try:
        a = torch.ones(330 ** 3, dtype=torch.uint8, device="cuda")
        k = torch.empty_like(a)
        rc = lib.vx_drop_islands_noalloc(h, _native.ptr(a), None, 0, _native.ptr(k), None)
        assert rc == _native.VX_ERR_ARG and b"no-allocation limit" in lib.vx_last_error(h)
    finally:
        lib.vx_destroy(h)
