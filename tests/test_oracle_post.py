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


def test_displacement_fields_of_a_linear_field():
    """u = (a x, 0, 0) in metres: node |u| = a ix, so the element mean of its
    8 corners is a (ix + 1/2); max_displacement_m runs over all nodes (void-
    only ones included, core/solve_worker.py:378-380) while
    displacement_vmax_m runs over active elements only (:399-400)."""
    nx, ny, nz, a = 5, 3, 2, 1e-4
    ix = np.arange(nx + 1)
    u = np.zeros(((nz + 1), (ny + 1), (nx + 1), 3))
    u[..., 0] = a * ix[None, None, :]
    u = u.reshape(-1)
    active = np.ones((nx, ny, nz), bool)
    active[nx - 1] = False                     # the last x-layer is void
    f = post_fields(nx, ny, nz, 0.3, 1e-3, u, np.full(nx * ny * nz, 1e9), active.ravel(),
                    250e6, True)
    want = (a * (np.arange(nx) + 0.5))[:, None, None] * np.ones((nx, ny, nz))
    assert np.allclose(f["displacement3d"], want, rtol=1e-6)
    assert np.isclose(f["max_displacement_m"], a * nx, rtol=1e-15)
    assert np.isclose(f["displacement_vmax_m"], a * (nx - 2 + 0.5), rtol=1e-7)
    # a zero field still reports the floor of the colour ceiling
    g = post_fields(nx, ny, nz, 0.3, 1e-3, 0 * u, np.full(nx * ny * nz, 1e9), active.ravel(),
                    250e6, True)
    assert g["displacement_vmax_m"] == 1e-9 and g["max_displacement_m"] == 0


def test_stress_percentile_of_a_graded_chain():
    """A 101-element chain along x with u_x(ix) = vsize ix (ix + 1) / 2: element
    i has the uniform strain (i + 1) in x only, so sigma = E c ((1 - nu), nu,
    nu)(i + 1) and von Mises |s_xx - s_yy| = E (i + 1) / (1 + nu).  With the
    last element void the 100 active values are v_k = E (k + 1) / (1 + nu),
    whose 99th percentile (linear interpolation at rank 0.99 * 99 = 98.01)
    is E 99.01 / (1 + nu) and whose maximum is 100 E / (1 + nu)."""
    nx, ny, nz, nu, E, vsize = 101, 1, 1, 0.3, 2.0e3, 1e-3
    ix = np.arange(nx + 1)
    u = np.zeros(((nz + 1), (ny + 1), (nx + 1), 3))
    u[..., 0] = vsize * (ix * (ix + 1) / 2.0)[None, None, :]
    active = np.ones(nx, bool)
    active[-1] = False
    f = post_fields(nx, ny, nz, nu, vsize, u.reshape(-1), np.full(nx, E), active, 1e6, True)
    unit = E / (1 + nu)
    vm = f["stress3d"].ravel()
    assert np.allclose(vm[:-1], unit * np.arange(1, nx), rtol=1e-6) and vm[-1] == 0
    assert np.isclose(f["stress_vmax"], 99.01 * unit, rtol=1e-12)
    assert np.isclose(f["max_stress_pa"], 100 * unit, rtol=1e-12)
    assert np.isclose(f["min_safety_factor"], 1e6 / (100 * unit), rtol=1e-12)
    sf = f["safety3d"].ravel()
    assert np.allclose(sf[:-1], np.clip(1e6 / (unit * np.arange(1, nx)), 0, 10), rtol=1e-6)
