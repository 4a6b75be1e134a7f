"""Voxelisation inside test on the GPU (core/inside_worker.py).

Same API and results as the reference's ``inside_mask`` numpy path
(core/inside_worker.py:118-162, 188-205): one tilted ray per point,
crossing parity, the reference's tolerances and operation order.  The
reference hands meshes of 600 triangles and more to trimesh + rtree
(core/inside_worker.py:196-205); this build runs the same single-ray test at
every size (a 2-D broad phase along the ray keeps it fast), so for large
meshes it equals the reference's numpy path, not trimesh's.
"""

import ctypes

import numpy as np

from . import _native

RAY_DIR = np.array([1.0, 0.0073301, 0.0031337])
RAY_DIR = RAY_DIR / np.linalg.norm(RAY_DIR)


def _torch():
    import torch
    return torch


def _check(rc):
    if rc != _native.VX_OK:
        m = _native.load().vx_inside_last_error()
        raise _native.NativeError(rc, m.decode() if m else "")


def _mesh_dev(verts, faces):
    torch = _torch()
    v = torch.as_tensor(np.ascontiguousarray(verts, dtype=np.float64)).reshape(-1, 3).cuda()
    f = torch.as_tensor(np.ascontiguousarray(faces, dtype=np.int64)).reshape(-1, 3).cuda()
    if f.numel() and (int(f.min()) < 0 or int(f.max()) >= v.shape[0]):
        raise ValueError("face index out of range")
    return v, f


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def inside_mask(verts, faces, points, direction=None, prefer_trimesh=True, progress_cb=None):
    """Bool mask of the points inside the closed mesh (core/inside_worker.py:188-205).
    ``prefer_trimesh`` is accepted for the reference's signature; the device
    path serves every mesh size."""
    torch = _torch()
    lib = _native.load()
    d = np.ascontiguousarray(RAY_DIR if direction is None else direction, dtype=np.float64)
    v, f = _mesh_dev(verts, faces)
    p = torch.as_tensor(np.ascontiguousarray(points, dtype=np.float64)).reshape(-1, 3).cuda()
    out = torch.empty(p.shape[0], dtype=torch.uint8, device="cuda")
    _check(lib.vx_inside_mask(_stream(), _native.ptr(v), v.shape[0], _native.ptr(f), f.shape[0],
                              _native.ptr(p), p.shape[0], d.ctypes.data_as(ctypes.c_void_p),
                              _native.ptr(out)))
    if progress_cb is not None:
        progress_cb(p.shape[0])
    return out.cpu().numpy().astype(bool)


def inside_grid_device(dims, origin, vsize, verts, faces, target="centers", direction=None):
    """Inside mask of a grid's voxel centres (element order) or nodes
    (FEA node order), generated on the device; returns a uint8 device tensor."""
    torch = _torch()
    lib = _native.load()
    nx, ny, nz = (int(x) for x in dims)
    t = {"centers": 0, "nodes": 1}[target]
    n = nx * ny * nz if t == 0 else (nx + 1) * (ny + 1) * (nz + 1)
    d = np.ascontiguousarray(RAY_DIR if direction is None else direction, dtype=np.float64)
    org = np.ascontiguousarray(origin, dtype=np.float64)
    v, f = _mesh_dev(verts, faces)
    out = torch.empty(n, dtype=torch.uint8, device="cuda")
    _check(lib.vx_inside_grid(_stream(), _native.ptr(v), v.shape[0], _native.ptr(f), f.shape[0],
                              nx, ny, nz, org.ctypes.data_as(ctypes.c_void_p), float(vsize), t,
                              d.ctypes.data_as(ctypes.c_void_p), _native.ptr(out)))
    return out


def inside_grid(dims, origin, vsize, verts, faces, target="centers", direction=None):
    """Host bool mask version of ``inside_grid_device``."""
    return inside_grid_device(dims, origin, vsize, verts, faces, target,
                              direction).cpu().numpy().astype(bool)
