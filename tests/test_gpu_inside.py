"""GPU voxelisation inside test against the oracle restatement of the
reference's numpy path: masks must be identical."""

import numpy as np
import pytest

from core import workloads as W
from oracle import inside_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from core import inside
    return inside


def _meshes():
    return {
        "box": W.box_mesh((-0.8, -0.6, -0.5), (0.7, 0.65, 0.9)),
        "sphere80": W.icosphere(1, (0.05, 0.02, -0.03), 0.9),
        "sphere1280": W.icosphere(3, (0.0, 0.1, 0.0), 0.95),
    }


@pytest.mark.parametrize("name", ["box", "sphere80", "sphere1280"])
def test_points_match_oracle(gpu, name):
    print(f"inside test points: {name}", flush=True)
    v, f = _meshes()[name]
    rng = np.random.default_rng(1)
    pts = rng.uniform(-1.2, 1.2, (20000, 3))
    # points on vertices and near faces exercise the tolerances
    near = v[f[:, 0]] * (1 + 1e-9 * rng.standard_normal((len(f), 1)))
    pts = np.concatenate([pts, v, near])
    assert np.array_equal(gpu.inside_mask(v, f, pts), R.inside_mask(v, f, pts))


def test_open_triangle_soup_parity(gpu):
    rng = np.random.default_rng(2)
    v = rng.uniform(-1, 1, (300, 3))
    f = rng.integers(0, 300, (400, 3))
    pts = rng.uniform(-1.2, 1.2, (8000, 3))
    assert np.array_equal(gpu.inside_mask(v, f, pts), R.inside_mask(v, f, pts))


@pytest.mark.parametrize("target", ["centers", "nodes"])
def test_grid_matches_oracle(gpu, target):
    v, f = W.icosphere(3, (0.0, 0.1, 0.0), 0.95)
    dims, origin, vs = (24, 20, 18), (-1.2, -1.05, -0.95), 0.1
    pts = R.voxel_centers(dims, origin, vs) if target == "centers" else \
        R.node_coords(dims, origin, vs)
    got = gpu.inside_grid(dims, origin, vs, v, f, target)
    assert np.array_equal(got, R.inside_mask(v, f, pts))
    assert got.sum() > 0


def test_empty_mesh_and_parallel_only(gpu):
    pts = np.random.default_rng(3).uniform(-1, 1, (100, 3))
    v = np.zeros((3, 3))
    f = np.array([[0, 1, 2]])                 # degenerate: parallel, never hit
    assert not gpu.inside_mask(v, f, pts).any()
    assert not gpu.inside_mask(v, np.zeros((0, 3), np.int64), pts).any()
