"""Pins of the inside-test oracle (oracle/inside_ref.py): closed meshes whose
inside is known analytically, the grid point formulas, and the ray set-up."""

import numpy as np

from core import workloads as W
from oracle import inside_ref as R


def test_ray_direction_is_the_reference_constant():
    d = np.array([1.0, 0.0073301, 0.0031337])
    assert np.allclose(R.RAY_DIR, d / np.linalg.norm(d), rtol=0, atol=0)


def test_box_and_sphere_analytic():
    rng = np.random.default_rng(0)
    pts = rng.uniform(-1.5, 1.5, (4000, 3))
    v, f = W.box_mesh((-1.0, -0.7, -0.4), (0.9, 0.6, 1.1))
    lo, hi = np.array([-1.0, -0.7, -0.4]), np.array([0.9, 0.6, 1.1])
    inside = np.all((pts > lo) & (pts < hi), axis=1)
    dist = np.min(np.abs(np.concatenate([pts - lo, hi - pts], axis=1)), axis=1)
    keep = dist > 1e-6
    assert np.array_equal(R.inside_mask(v, f, pts)[keep], inside[keep])
    # sphere mesh: points clearly inside the inscribed / outside the circumscribed sphere
    v, f = W.icosphere(3, (0.1, -0.2, 0.05), 1.2)
    r = np.linalg.norm(pts - np.array([0.1, -0.2, 0.05]), axis=1)
    got = R.inside_mask(v, f, pts)
    assert np.all(got[r < 1.15]) and not np.any(got[r > 1.2])


def test_grid_point_formulas():
    c = R.voxel_centers((3, 4, 5), (1.0, 2.0, 3.0), 0.5)
    # element order e = iz + nz*(iy + ny*ix)
    e = 4 + 5 * (2 + 4 * 1)
    assert np.allclose(c[e], [1.0 + 1.5 * 0.5, 2.0 + 2.5 * 0.5, 3.0 + 4.5 * 0.5])
    n = R.node_coords((3, 4, 5), (1.0, 2.0, 3.0), 0.5)
    k = 2 + 4 * (3 + 5 * 5)     # node id ix + (nx+1)(iy + (ny+1) iz)
    assert np.allclose(n[k], [1.0 + 1.0, 2.0 + 1.5, 3.0 + 2.5])


def test_open_mesh_parity_counts_crossings():
    # two parallel squares crossing the ray: points between them are "inside"
    v = np.array([[0, -1, -1], [0, 1, -1], [0, 1, 1], [0, -1, 1],
                  [1, -1, -1], [1, 1, -1], [1, 1, 1], [1, -1, 1]], float)
    f = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]])
    pts = np.array([[-0.5, 0.1, 0.2], [0.5, 0.1, 0.2], [1.5, 0.1, 0.2]])
    assert R.inside_mask(v, f, pts).tolist() == [False, True, False]
