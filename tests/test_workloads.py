"""The synthetic inputs of the five measured configurations hold what the
configurations state (sizes, materials, supports, load totals), and the
multi-process CPU matvec restatement equals the single-process apply."""

import numpy as np
import pytest

from core import workloads as W


def _fixed_nodes(p, comps=(0, 1, 2)):
    ids = np.asarray(p.fixed_dofs)
    return {c: set((ids[ids % 3 == c] // 3).tolist()) for c in comps}


def test_cantilever():
    p = W.cantilever()
    assert (p.nx, p.ny, p.nz) == (64, 16, 16) and p.tol == 1e-8 and p.max_levels == 3
    assert p.active.all() and np.all(p.E_pa == 200e9) and p.nu == 0.3
    f = _fixed_nodes(p)
    assert all(len(f[c]) == 17 * 17 for c in range(3))          # all DOFs of the x=0 nodes
    assert all(n % 65 == 0 for n in f[0])
    fz = p.force[2::3]
    assert np.isclose(fz.sum(), -1000.0) and np.count_nonzero(fz) == 289
    assert np.allclose(fz[fz != 0], -1000.0 / 289)
    assert not p.force[0::3].any() and not p.force[1::3].any()


def test_voids_block_scaled():
    p = W.voids_block(n=32)
    assert np.all(p.E_pa == 70e9) and p.nu == 0.33
    assert 0.3 < p.active.mean() < 1.0
    f = _fixed_nodes(p)
    assert all(len(f[c]) == 33 * 33 for c in range(3))           # z = 0 plane
    fx = p.force[0::3]
    assert np.isclose(fx.sum(), 5000.0)
    loaded = np.flatnonzero(fx)
    assert np.all(loaded // (33 * 33) == 32)                        # on the z = max plane
    assert not p.force[1::3].any() and not p.force[2::3].any()


def test_multi_rhs_part_scaled():
    p = W.multi_rhs_part(n=24, nrhs=4)
    assert p.force.shape == (4, 3 * 25 ** 3)
    assert 0.40 <= p.active.mean() <= 0.70
    for k in range(4):
        F = p.force[k].reshape(-1, 3)
        tot = F.sum(axis=0)
        assert 100.0 <= np.linalg.norm(tot) <= 5000.0
        nodes = np.flatnonzero(np.abs(F).sum(axis=1))
        ix, iy, iz = nodes % 25, (nodes // 25) % 25, nodes // 625
        assert np.all(iz == 24) and np.ptp(ix) <= 8 and np.ptp(iy) <= 8


def test_uneven_parts_scaled():
    parts = W.uneven_parts(n_parts=3, lo=20, hi=30, seed=7)
    assert len(parts) == 3
    for q in parts:
        assert all(20 <= d <= 30 for d in (q.nx, q.ny, q.nz))
        assert 0.30 <= q.active.mean() <= 0.70
        assert np.all(q.E_pa == 2.2e9) and q.nu == 0.35
        assert 100.0 <= np.linalg.norm(q.force.reshape(-1, 3).sum(axis=0)) <= 5000.0


def test_two_material_scaled():
    n = 32
    p = W.two_material(n=n)
    assert set(np.unique(p.E_pa)) == {200e9, 1.5e9} and p.nu == 0.3
    assert 0.40 < p.active.mean() < 0.60
    f = _fixed_nodes(p)
    x0 = {i for i in range((n + 1) ** 3) if i % (n + 1) == 0}
    xmax = {i for i in range((n + 1) ** 3) if i % (n + 1) == n}
    assert f[0] == x0 and f[1] == x0 and f[2] == x0 | xmax        # roller: z locked at x=max
    assert np.isclose(p.force[2::3].sum(), -20000.0)
    assert not p.force[0::3].any() and not p.force[1::3].any()


def test_cpu_pool_equals_single_process():
    from oracle.fea_ref import build_edof, hex8_KE
    from oracle.parallel_cpu_ref import CPUMatVecPoolRef
    nx, ny, nz = 6, 5, 4
    edof = build_edof(nx, ny, nz)
    KE = hex8_KE(1.0, 0.3)
    ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    rng = np.random.default_rng(0)
    ev = rng.uniform(0.5, 2.0, nx * ny * nz)
    p = rng.standard_normal(ndof)
    want = np.bincount(edof.ravel(), weights=((p[edof] @ KE.T) * ev[:, None]).ravel(),
                       minlength=ndof)
    pool = CPUMatVecPoolRef(edof, KE, ndof, 2)
    try:
        pool.set_density(ev)
        got = pool.apply(p)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        d = pool.diagonal()
        dwant = np.bincount(edof.ravel(), weights=(ev[:, None] * np.diag(KE)[None, :]).ravel(),
                            minlength=ndof)
        assert np.allclose(d, dwant)
    finally:
        pool.close()
