"""Pins of the oracle's element kernels (core/fea.py:31-148) against closed
forms and invariants, and of the HIP library's host element operator."""

import ctypes
import os
import re

import numpy as np
import pytest

from oracle.fea_ref import hex8_KE, hex8_B_C, build_edof, node_id, von_mises

NU = [0.0, 0.3, 0.33, 0.35, 0.42]


def _mode_KE(nu):
    """Independent derivation: energy of the trilinear field in the monomial
    basis under 2x2x2 Gauss (gradient monomials are orthogonal), assembled
    with a general isotropic C(G) = lam tr(G) I + mu (G + G^T)."""
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu = 1 / (2 * (1 + nu))
    signs = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    sx, sy, sz = signs.T
    chi = {"x": sx, "y": sy, "z": sz, "xy": sx * sy, "yz": sy * sz, "zx": sz * sx,
           "xyz": sx * sy * sz}

    def C(G):
        return lam * np.trace(G) * np.eye(3) + mu * (G + G.T)

    KE = np.zeros((24, 24))
    for j in range(24):
        u = np.zeros((8, 3))
        u[j // 3, j % 3] = 1.0
        h = {m: chi[m] @ u for m in chi}
        # gradient coefficient matrices (d/dx = 2 d/dxi, u = sum h_m m / 8)
        G0 = np.stack([h["x"], h["y"], h["z"]], 1) / 4
        Gxi = np.zeros((3, 3)); Gxi[:, 1] = h["xy"] / 4; Gxi[:, 2] = h["zx"] / 4
        Get = np.zeros((3, 3)); Get[:, 0] = h["xy"] / 4; Get[:, 2] = h["yz"] / 4
        Gze = np.zeros((3, 3)); Gze[:, 0] = h["zx"] / 4; Gze[:, 1] = h["yz"] / 4
        Gyz = np.zeros((3, 3)); Gyz[:, 0] = h["xyz"] / 4
        Gxz = np.zeros((3, 3)); Gxz[:, 1] = h["xyz"] / 4
        Gxy = np.zeros((3, 3)); Gxy[:, 2] = h["xyz"] / 4
        S0, Sxi, Set, Sze = C(G0), C(Gxi), C(Get), C(Gze)
        yh = {"x": S0[:, 0] / 4, "y": S0[:, 1] / 4, "z": S0[:, 2] / 4,
              "xy": (Sxi[:, 1] + Set[:, 0]) / 12, "yz": (Set[:, 2] + Sze[:, 1]) / 12,
              "zx": (Sxi[:, 2] + Sze[:, 0]) / 12,
              "xyz": (C(Gyz)[:, 0] + C(Gxz)[:, 1] + C(Gxy)[:, 2]) / 36}
        y = sum(np.outer(chi[m], yh[m]) for m in yh)
        KE[:, j] = y.ravel()
    return KE


def _rigid_modes():
    signs = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                      [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    modes = []
    for c in range(3):
        m = np.zeros((8, 3)); m[:, c] = 1; modes.append(m.ravel())
    for a, b in ((0, 1), (1, 2), (2, 0)):
        m = np.zeros((8, 3)); m[:, a] = -signs[:, b]; m[:, b] = signs[:, a]; modes.append(m.ravel())
    return np.array(modes).T


@pytest.mark.parametrize("nu", NU)
def test_KE_matches_independent_mode_derivation(nu):
    KE = hex8_KE(1.0, nu)
    assert np.abs(KE - _mode_KE(nu)).max() < 1e-14 * np.abs(KE).max()


@pytest.mark.parametrize("nu", NU)
def test_KE_invariants(nu):
    KE = hex8_KE(1.0, nu)
    assert np.allclose(KE, KE.T, atol=1e-16)
    w = np.linalg.eigvalsh(KE)
    assert w.min() > -1e-13
    assert np.sum(np.abs(w) < 1e-12) == 6
    R = _rigid_modes()
    assert np.abs(KE @ R).max() < 1e-14
    assert np.allclose(hex8_KE(3.5, nu), 3.5 * KE, rtol=1e-15)
    assert np.ptp(np.diag(KE)) < 1e-15


@pytest.mark.parametrize("nu", [0.3, 0.42])
def test_patch_test_uniform_strain(nu):
    """u = G x on the corners: KE u_e = B0^T C sym(G) (unit volume)."""
    rng = np.random.default_rng(3)
    G = rng.standard_normal((3, 3))
    corners = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    u = (corners @ G.T).ravel()
    B0, C0 = hex8_B_C(nu)
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1],
                    G[2, 0] + G[0, 2]])
    assert np.allclose(B0 @ u, eps, atol=1e-15)
    assert np.allclose(hex8_KE(1.0, nu) @ u, B0.T @ (C0 @ eps), atol=1e-13)


def test_B_C_closed_form():
    B0, C0 = hex8_B_C(0.3)
    assert set(np.unique(np.round(np.abs(B0[B0 != 0]), 15))) == {0.25}
    c = 1 / (1.3 * 0.4)
    assert np.isclose(C0[0, 0], c * 0.7) and np.isclose(C0[0, 1], c * 0.3)
    assert np.isclose(C0[3, 3], c * 0.2)


def test_von_mises_closed_forms():
    nu, E = 0.3, 200e9
    corners = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                        [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], float)  # node order x-fastest
    sig = 123e6
    # uniaxial stress sigma_xx: eps = (s/E, -nu s/E, -nu s/E)
    G = np.diag([sig / E, -nu * sig / E, -nu * sig / E])
    u = (corners @ G.T).ravel()
    vm = von_mises(1, 1, 1, nu, u, [E], [True])
    assert np.isclose(vm[0], sig, rtol=1e-12)
    # pure shear tau_xy: gamma = tau / mu
    tau = 50e6
    mu = E / (2 * (1 + nu))
    G = np.zeros((3, 3)); G[0, 1] = tau / mu
    u = (corners @ G.T).ravel()
    vm = von_mises(1, 1, 1, nu, u, [E], [True])
    assert np.isclose(vm[0], np.sqrt(3) * tau, rtol=1e-12)
    assert von_mises(1, 1, 1, nu, u, [E], [False])[0] == 0.0


def test_build_edof_brute_force():
    nx, ny, nz = 3, 2, 4
    edof = build_edof(nx, ny, nz)
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    e = 0
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                want = []
                for ox, oy, oz in offs:
                    n = (ix + ox) + (nx + 1) * ((iy + oy) + (ny + 1) * (iz + oz))
                    want += [3 * n, 3 * n + 1, 3 * n + 2]
                assert list(edof[e]) == want
                e += 1
    assert node_id(1, 1, 1, nx, ny) == 1 + 4 * (1 + 3)


# ---------------------------------------------------------------------------
# the HIP library loads on the CPU and exports its ABI; its host element
# operator (the kernels' own code) equals KE
# ---------------------------------------------------------------------------
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "voxelsim-fea_amd", "lib", "libvxfea.so")
HDR = os.path.join(ROOT, "include", "vxfea.h")


def _lib():
    if not os.path.exists(LIB):
        pytest.fail("libvxfea.so not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def test_library_exports_every_declared_symbol():
    lib = _lib()
    text = open(HDR).read()
    names = set(re.findall(r"\b(vxn?_[a-z_0-9]+)\s*\(", text))
    assert len(names) >= 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    from core import _native
    assert names <= set(_native.SIGNATURES), names - set(_native.SIGNATURES)


@pytest.mark.parametrize("nu", NU + [0.49])
def test_library_host_element_operator(nu):
    lib = _lib()
    lib.vx_element_apply_host.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    KE = hex8_KE(1.0, nu)
    Y = np.zeros((24, 24))
    y = np.zeros(24)
    for j in range(24):
        u = np.zeros(24)
        u[j] = 1.0
        lib.vx_element_apply_host(nu, u.ctypes.data, y.ctypes.data)
        Y[:, j] = y
    assert np.abs(Y - KE).max() < 1e-15 * np.abs(KE).max() * 8


@pytest.mark.parametrize("struct", ["vx_options", "vx_info"])
def test_ctypes_structs_match_the_header(struct, tmp_path):
    """The ctypes mirrors have the C header's size and field offsets (gcc on
    include/vxfea.h)."""
    import os
    import shutil
    import subprocess
    from core import _native
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    cls = getattr(_native, struct)
    names = [f[0] for f in cls._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"vxfea.h\"\nint main(void) {\n"
                   + f'  printf("%zu\\n", sizeof({struct}));\n'
                   + "".join(f'  printf("%zu\\n", offsetof({struct}, {n}));\n' for n in names)
                   + "  return 0;\n}\n")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got[0] == ctypes.sizeof(cls)
    assert got[1:] == [getattr(cls, n).offset for n in names]


def test_default_options_are_reference_defaults():
    from core import _native
    o = _native.default_options()
    assert (o.max_levels, o.min_elems, o.n_smooth, o.coarse_sweeps) == (6, 4, 2, 30)
    assert o.omega == 0.6 and o.precision == _native.VX_FP64
    assert o.struct_size == ctypes.sizeof(_native.vx_options)


def test_level_smooth_option_packing():
    """Per-level smoothing counts reach the C options as 16 bytes (0: n_smooth);
    more than 16 levels or a count outside 0-255 is an error before any GPU
    call."""
    from core import _native
    from core.multigrid import _level_smooth
    o = _native.default_options(level_smooth=_level_smooth([2, 2, 6, 12, 24, 24]))
    assert list(o.level_smooth) == [2, 2, 6, 12, 24, 24] + [0] * 10
    assert list(_native.default_options().level_smooth) == [0] * 16
    assert _native.default_options().eager == 0
    assert _level_smooth(None) is None
    for bad in ([1] * 17, [256], [-1]):
        with pytest.raises(ValueError):
            _level_smooth(bad)


def _create_status(dims=(16, 16, 16), **kw):
    """vx_create's status for options kw; a handle it returns is destroyed."""
    from core import _native
    lib = _native.load()
    h = ctypes.c_void_p()
    rc = lib.vx_create(*dims, 0.3, ctypes.byref(_native.default_options(**kw)), ctypes.byref(h))
    if rc == _native.VX_OK:
        lib.vx_destroy(h)
    return rc


def test_create_rejects_degenerate_chebyshev_intervals():
    """Every Chebyshev pass smooths over [r lambda, lambda] with r = cheb_ratio
    (n_smooth / k)^p for its k steps; r must lie in (0, 1) (r = 1 gives a
    zero-width interval and division by zero in the step coefficients).
    vx_create checks each level's step count and the coarsest sweeps before
    it looks for a device, so a bad set is VX_ERR_ARG on any machine."""
    from core import _native
    from core.multigrid import _level_smooth
    cheb = _native.VX_SMOOTH_CHEBYSHEV
    ok = (_native.VX_OK, _native.VX_ERR_NO_DEVICE)
    bad = _native.VX_ERR_ARG
    # level 1 with k = 2 < n_smooth = 4: r = 0.25 * (4/2)^2 = 1
    assert _create_status(smoother=cheb, cheb_ratio=0.25, n_smooth=4, cheb_ratio_power=2,
                          max_levels=3, level_smooth=_level_smooth([4, 2])) == bad
    # the coarsest sweeps: r = 0.3 * (4/2)^2 = 1.2
    assert _create_status(smoother=cheb, cheb_ratio=0.3, n_smooth=4, cheb_ratio_power=2,
                          max_levels=3, coarse_sweeps=2) == bad
    for r in (0.0, 1.0, -0.1, float("nan")):
        assert _create_status(smoother=cheb, cheb_ratio=r) == bad, r
    # the same ratios with the Jacobi smoother are never used
    assert _create_status(cheb_ratio=1.0) in ok
    # the benchmark's configuration: 0.05 (2/k)^2 for k = 2, 6, 12 and 48
    assert _create_status(smoother=cheb, cheb_ratio=0.05, cheb_ratio_power=2, max_levels=5,
                          level_smooth=_level_smooth([2, 2, 6, 12]), coarse_sweeps=48) in ok
    # a one-level hierarchy (odd edge) uses only the coarsest sweeps
    assert _create_status((15, 16, 16), smoother=cheb, cheb_ratio=0.3, n_smooth=4,
                          cheb_ratio_power=2, coarse_sweeps=2) == bad


def test_create_rejects_bad_block_options():
    """block_steps (steps per launch of the blocked FP32 smoother) is 0-8 and
    block_edge (its owned box edge) 0 or 2-16; vx_create checks both before it
    looks for a device."""
    from core import _native
    ok = (_native.VX_OK, _native.VX_ERR_NO_DEVICE)
    bad = _native.VX_ERR_ARG
    for steps, edge in ((-1, 0), (9, 0), (0, 1), (0, 17), (0, -2)):
        assert _create_status(block_steps=steps, block_edge=edge) == bad, (steps, edge)
    for steps, edge in ((0, 0), (1, 0), (8, 16), (3, 2)):
        assert _create_status(block_steps=steps, block_edge=edge) in ok, (steps, edge)


def test_create_rejects_bad_keep_direction_options():
    """keep_direction is a 0/1 switch and keep_drift a limit >= 0 (0: none);
    a NaN limit is rejected too."""
    from core import _native
    ok = (_native.VX_OK, _native.VX_ERR_NO_DEVICE)
    bad = _native.VX_ERR_ARG
    for keep, drift in ((2, 0.0), (-1, 0.0), (1, -0.1), (1, float("nan"))):
        assert _create_status(keep_direction=keep, keep_drift=drift) == bad, (keep, drift)
    for keep, drift in ((0, 0.0), (1, 0.0), (1, 0.3), (0, 5.0)):
        assert _create_status(keep_direction=keep, keep_drift=drift) in ok, (keep, drift)


# ---------------------------------------------------------------------------
# element_strain_energy (core/fea.py:595-602)
# ---------------------------------------------------------------------------

def _sparse_K(nx, ny, nz, E, nu):
    """Global K = sum_e E_e P_e^T KE P_e by scipy's sparse assembly."""
    import scipy.sparse as sp
    edof = build_edof(nx, ny, nz)
    KE = hex8_KE(1.0, nu)
    n = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    rows = np.repeat(edof, 24, axis=1).ravel()
    cols = np.tile(edof, (1, 24)).ravel()
    vals = (np.asarray(E, float)[:, None, None] * KE[None]).ravel()
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


@pytest.mark.parametrize("dims", [(3, 2, 2), (4, 3, 5)])
def test_strain_energy_sums_to_global_quadratic_form(dims):
    """sum_e E_e ce_e = u^T K u with K assembled by scipy, E_e = 0 on the
    inactive elements (their ce is 0); and a field living on one element's
    nodes only has energy there and in the elements sharing those nodes."""
    from oracle.fea_ref import VoxelFEARef
    nx, ny, nz = dims
    rng = np.random.default_rng(sum(dims))
    active = rng.random(nx * ny * nz) < 0.7
    E = rng.uniform(0.5, 3.0, nx * ny * nz)
    ref = VoxelFEARef(nx, ny, nz, nu=0.3, active_elems=active)
    u = rng.standard_normal(3 * (nx + 1) * (ny + 1) * (nz + 1))
    ce = ref.element_strain_energy(u)
    assert np.all(ce[~active] == 0) and np.all(ce[active] > 0)
    K = _sparse_K(nx, ny, nz, np.where(active, E, 0.0), 0.3)
    assert np.isclose((E * ce).sum(), u @ (K @ u), rtol=1e-12)
    # one element's corner DOFs only: the energy sits in the elements that
    # share at least one of those nodes
    e = int(np.nonzero(active)[0][0])
    v = np.zeros_like(u)
    edof = build_edof(nx, ny, nz)
    v[edof[e]] = rng.standard_normal(24)
    ce1 = ref.element_strain_energy(v)
    nodes = set(edof[e][0::3] // 3)
    share = np.array([bool(nodes & set(edof[k][0::3] // 3)) for k in range(nx * ny * nz)])
    assert np.all(ce1[~share] == 0) and ce1[e] > 0


@pytest.mark.parametrize("nu", [0.0, 0.3, 0.42])
def test_strain_energy_closed_forms(nu):
    """Linear fields in unit cubes with E = 1: uniaxial stress s_xx = e has
    u = (e x, -nu e y, -nu e z) and u_e^T KE u_e = eps.sigma = e^2; simple
    shear u = (g y, 0, 0) has g^2 G with G = 1 / (2 (1 + nu))."""
    from oracle.fea_ref import VoxelFEARef
    nx, ny, nz = 3, 2, 4
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij")

    def field(ux, uy, uz):
        u = np.stack([ux, uy, uz], -1).transpose(2, 1, 0, 3)   # node order x fastest
        return u.reshape(-1)

    ref = VoxelFEARef(nx, ny, nz, nu=nu)
    e, g = 1e-3, 2e-3
    ce = ref.element_strain_energy(field(e * ix, -nu * e * iy, -nu * e * iz))
    assert np.allclose(ce, e * e, rtol=1e-12)
    ce = ref.element_strain_energy(field(g * iy, 0 * iy, 0 * iz))
    assert np.allclose(ce, g * g / (2 * (1 + nu)), rtol=1e-12)
    # a rigid translation plus rotation carries no energy
    ce = ref.element_strain_energy(field(1.0 + 0.2 * iy, -0.2 * ix, 0.5 + 0 * iz))
    assert np.abs(ce).max() < 1e-14


def test_blocked_smoother_small_division():
    """The blocked smoother splits flat indices with q = (n + 0.5) * fl(1/d)
    in FP32 instead of an integer division.  Exhaustively for n < 2^16 and
    d <= 2048 (its indices stay below 2^13 and its divisors below 2^10),
    the truncated product equals n // d."""
    n = np.arange(1 << 16, dtype=np.int64)
    nf = n.astype(np.float32) + np.float32(0.5)
    for d in range(1, 2049):
        rd = np.float32(1.0) / np.float32(d)
        q = (nf * rd).astype(np.int64)
        assert np.array_equal(q, n // d), d


@pytest.mark.parametrize("ratio", [0.05, 0.3])
def test_two_root_weights_give_the_degree2_chebyshev_polynomial(ratio):
    """The FP32 smoother runs a two-step Chebyshev pass from a nonzero start
    as two damped-Jacobi steps weighted 1/(theta + delta/sqrt 2), then
    1/(theta - delta/sqrt 2).  The error polynomial of that pair equals the
    pass's T_2((theta - x)/delta) / T_2(theta/delta) and the smoother's
    three-term recurrence (x1 = x0 + D^-1 r0 / theta, then rho_1 rho_0 and
    2 rho_1 / delta) on the eigenvalues x of D^-1 A; weights from any other
    pair of points (here delta/2 in place of delta/sqrt 2) do not."""
    x = np.linspace(0.0, 1.2, 241)
    a, b = ratio, 1.0
    theta, delta = 0.5 * (b + a), 0.5 * (b - a)
    sigma = theta / delta
    cheb = (2 * ((theta - x) / delta) ** 2 - 1) / (2 * sigma ** 2 - 1)
    rho0 = 1.0 / sigma
    rho1 = 1.0 / (2 * sigma - rho0)
    e1 = 1 - x / theta
    three_term = e1 + rho1 * rho0 * (e1 - 1) - (2 * rho1 / delta) * x * e1
    r = delta / np.sqrt(2.0)
    two_root = (1 - x / (theta + r)) * (1 - x / (theta - r))
    assert np.abs(three_term - cheb).max() < 1e-12
    assert np.abs(two_root - cheb).max() < 1e-12
    wrong = (1 - x / (theta + delta / 2)) * (1 - x / (theta - delta / 2))
    assert np.abs(wrong - cheb).max() > 1e-2
