"""Closed-form pins of the post-solve oracle (core/solve_worker.py:358-431)."""

import numpy as np

from oracle.postsolve_ref import post_fields


def _uniaxial(nx, ny, nz, nu, E, sig):
    """u of a uniform uniaxial stress sig_xx: eps = (s/E, -nu s/E, -nu s/E)."""
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    u = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    e = sig / E
    u[..., 0] = (e * ix).transpose(2, 1, 0)
    u[..., 1] = (-nu * e * iy).transpose(2, 1, 0)
    u[..., 2] = (-nu * e * iz).transpose(2, 1, 0)
    return u.reshape(-1)


def test_uniform_uniaxial_fields():
    nx, ny, nz, nu, E, vsize = 4, 3, 2, 0.3, 200e9, 1e-3
    sig = 80e6
    # unit-cube elements: u in "voxel units" x vsize gives strain sig/E
    u = _uniaxial(nx, ny, nz, nu, E, sig) * vsize
    active = np.ones(nx * ny * nz, bool)
    active[0] = False
    f = post_fields(nx, ny, nz, nu, vsize, u, np.full(nx * ny * nz, E), active, 250e6, True)
    vm = f["stress3d"].ravel()
    assert np.allclose(vm[1:], sig, rtol=1e-6) and vm[0] == 0
    assert np.isclose(f["max_stress_pa"], sig, rtol=1e-12)
    assert np.isclose(f["stress_vmax"], sig, rtol=1e-12)
    assert np.isclose(f["min_safety_factor"], 250e6 / sig, rtol=1e-12)
    assert np.allclose(f["safety3d"].ravel()[1:], 250e6 / sig, rtol=1e-6)
    assert f["safety3d"].ravel()[0] == 0
    # displacement: mean of the corner |u|; max over all nodes
    mag = np.sqrt((u.reshape(-1, 3) ** 2).sum(1))
    assert np.isclose(f["max_displacement_m"], mag.max())
    assert f["reliable"]


def test_safety_clip_and_zero_yield():
    nx = ny = nz = 2
    u = np.zeros(3 * 27)
    active = np.ones(8, bool)
    f = post_fields(nx, ny, nz, 0.3, 1e-3, u, np.full(8, 1e9), active, 100e6, True)
    assert np.all(f["safety3d"] == 10.0)          # zero stress -> capped
    assert f["stress_vmax"] == 1e-6               # floor of the colour ceiling
    g = post_fields(nx, ny, nz, 0.3, 1e-3, u, np.full(8, 1e9), active, 0.0, False)
    assert g["min_safety_factor"] is None and np.all(g["safety3d"] == 0) and not g["reliable"]
