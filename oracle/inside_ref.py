"""CPU restatement of the voxelisation inside test (numpy) -- test
infrastructure only: the checker of the HIP inside test, never called by the
product.

Follows core/inside_worker.py:

* ``RAY_DIR``: the tilted, normalised ray direction (core/inside_worker.py:57-58);
* ``voxel_centers`` / ``node_coords``: grid point generation
  (core/inside_worker.py:87-111, the formulas of core/voxelize.py:51-70);
* ``inside_mask``: single-direction crossing-parity Moeller-Trumbore test of
  ``_inside_mask_numpy`` (core/inside_worker.py:118-162) with its tolerances
  (det 1e-9, barycentric -1e-7 / 1 + 1e-7, t > 1e-7) and arithmetic order.
  The reference switches to trimesh + rtree from 600 triangles up
  (core/inside_worker.py:118, 196-205); trimesh is not vendored in the
  reference and not installed here, so that branch is not restated: results
  for large meshes are pinned to this single-ray algorithm, and to the
  analytic inside tests of closed meshes in tests/test_oracle_inside.py.
"""

import numpy as np

RAY_DIR = np.array([1.0, 0.0073301, 0.0031337])
RAY_DIR = RAY_DIR / np.linalg.norm(RAY_DIR)


def voxel_centers(dims, origin, vsize):
    nx, ny, nz = dims
    origin = np.asarray(origin, dtype=np.float64)
    vsize = float(vsize)
    xs = (np.arange(nx) + 0.5) * vsize + origin[0]
    ys = (np.arange(ny) + 0.5) * vsize + origin[1]
    zs = (np.arange(nz) + 0.5) * vsize + origin[2]
    gx, gy, gz = np.meshgrid(xs, ys, zs, indexing="ij")
    return np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)


def node_coords(dims, origin, vsize):
    nx, ny, nz = dims
    origin = np.asarray(origin, dtype=np.float64)
    n = np.arange((nx + 1) * (ny + 1) * (nz + 1))
    ix = n % (nx + 1)
    iy = (n // (nx + 1)) % (ny + 1)
    iz = n // ((nx + 1) * (ny + 1))
    coords = np.stack([ix, iy, iz], axis=1).astype(np.float64) * float(vsize)
    return coords + origin


def inside_mask(verts, faces, points, direction=None, chunk=4096):
    """Odd number of forward crossings of the ray p + t d (t > 1e-7)."""
    d = np.asarray(RAY_DIR if direction is None else direction, dtype=np.float64)
    verts = np.asarray(verts, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    points = np.asarray(points, dtype=np.float64)
    inside = np.zeros(len(points), dtype=bool)
    if len(points) == 0 or len(faces) == 0:
        return inside
    v0, v1, v2 = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    e1 = v1 - v0
    e2 = v2 - v0
    pvec = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, pvec)
    parallel = np.abs(det) < 1e-9
    inv_det = np.zeros_like(det)
    inv_det[~parallel] = 1.0 / det[~parallel]
    for s in range(0, len(points), chunk):
        p = points[s:s + chunk]
        tvec = p[:, None, :] - v0[None, :, :]
        u = np.einsum("ptj,tj->pt", tvec, pvec) * inv_det[None, :]
        qvec = np.cross(tvec, e1[None, :, :])
        v = np.einsum("j,ptj->pt", d, qvec) * inv_det[None, :]
        t = np.einsum("tj,ptj->pt", e2, qvec) * inv_det[None, :]
        hit = (~parallel[None, :] & (u >= -1e-7) & (v >= -1e-7)
               & (u + v <= 1.0 + 1e-7) & (t > 1e-7))
        inside[s:s + chunk] = (hit.sum(axis=1) % 2) == 1
    return inside
