"""GPU parity of the single operators against the oracle (core/multigrid.py
_apply / _smooth / _restrict / _prolongate / _vcycle, set_density Minv).

FP64 kernels: relative max error <= 1e-12 (only the summation order
differs from the oracle's bincount).  FP32 kernels: <= 2e-5 relative
(FP32 storage and arithmetic of the same operator)."""

import numpy as np
import pytest

from oracle.multigrid_ref import MGSolverRef

pytestmark = pytest.mark.gpu

GRIDS = [
    (8, 8, 8), (16, 8, 4), (6, 10, 4), (5, 7, 3), (33, 17, 9), (2, 2, 2), (1, 1, 1),
    (40, 34, 20),
]


def _problem(nx, ny, nz, seed, nfix_frac=0.1):
    rng = np.random.default_rng(seed)
    ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    E = rng.uniform(0.1, 3.0, nx * ny * nz) * 10.0 ** rng.integers(-3, 1, nx * ny * nz)
    fixed = np.unique(rng.choice(ndof, max(1, int(nfix_frac * ndof)), replace=False))
    return E, fixed, ndof, rng


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _pair(nx, ny, nz, fixed, E, nu=0.3, **kw):
    from core.multigrid import MGSolver
    g = MGSolver(nx, ny, nz, fixed, nu=nu, compute_mode="GPU", **kw)
    g.set_density(E)
    o = MGSolverRef(nx, ny, nz, fixed, nu=nu, **{k: v for k, v in kw.items()
                                                   if k in ("max_levels", "n_smooth", "omega",
                                                            "level_smooth")})
    o.set_density(E)
    return g, o


@pytest.mark.parametrize("dims", GRIDS)
def test_matvec_fp64_all_levels(dims):
    print("case matvec", dims, flush=True)
    E, fixed, ndof, rng = _problem(*dims, seed=sum(dims))
    g, o = _pair(*dims, fixed, E)
    assert len(g.levels) == len(o.levels)
    for li, lv in enumerate(o.levels):
        u = rng.standard_normal(lv["ndof"])
        assert _rel(g._apply(li, u), o._apply(li, u)) < 1e-12
    g.close()


@pytest.mark.parametrize("dims", [(8, 8, 8), (33, 17, 9), (40, 34, 20)])
def test_matvec_fp32(dims):
    print("case matvec fp32", dims, flush=True)
    E, fixed, ndof, rng = _problem(*dims, seed=7)
    g, o = _pair(*dims, fixed, E)
    u = rng.standard_normal(ndof)
    assert _rel(g._apply(0, u, precision="mixed"), o._apply(0, u)) < 2e-5
    g.close()


@pytest.mark.parametrize("dims", [(8, 8, 8), (16, 8, 4), (5, 7, 3)])
def test_smooth_restrict_prolong(dims):
    print("case smooth/transfer", dims, flush=True)
    E, fixed, ndof, rng = _problem(*dims, seed=11)
    g, o = _pair(*dims, fixed, E)
    for li, lv in enumerate(o.levels):
        u = rng.standard_normal(lv["ndof"])
        b = rng.standard_normal(lv["ndof"])
        for it in (1, 2, 3):
            assert _rel(g._smooth(li, u, b, it), o._smooth(li, u, b, it)) < 1e-12
        z0 = o._smooth(li, np.zeros(lv["ndof"]), b, 2)
        assert _rel(g._smooth(li, None, b, 2), z0) < 1e-12
        assert _rel(g.inv_diag(li), lv["Minv"]) < 1e-13
        if li + 1 < len(o.levels):
            assert _rel(g._restrict(li, u), o._restrict(li, u)) < 1e-13
            ec = rng.standard_normal(o.levels[li + 1]["ndof"])
            assert _rel(g._prolongate(li, ec), o._prolongate(li, ec)) < 1e-13
    g.close()


@pytest.mark.parametrize("dims,kw", [((16, 8, 8), {}), ((8, 8, 8), {"max_levels": 3}),
                                     ((6, 10, 4), {"n_smooth": 3}), ((5, 7, 3), {}),
                                     ((160, 12, 4), {"max_levels": 2}),
                                     ((16, 8, 8), {"level_smooth": [2, 5, 3]}),
                                     ((8, 8, 8), {"max_levels": 3, "level_smooth": [1, 4]})])
def test_vcycle(dims, kw):
    print("case vcycle", dims, kw, flush=True)
    E, fixed, ndof, rng = _problem(*dims, seed=5)
    g, o = _pair(*dims, fixed, E, **kw)
    b = np.where(o.levels[0]["free"], rng.standard_normal(ndof), 0.0)
    z = o._vcycle(0, b)
    assert _rel(g._vcycle(0, b), z) < 1e-11
    assert _rel(g._vcycle(0, b, precision="mixed"), z) < 1e-4
    g.close()


@pytest.mark.parametrize("smoother,steps,power", [("chebyshev", [2, 3, 9], 0),
                                                  ("jacobi", [2, 2, 5], 0),
                                                  ("chebyshev", [1, 1, 1], 0),
                                                  ("chebyshev", [2, 3, 9], 2)])
def test_vcycle_fused_bottom_levels(smoother, steps, power):
    """FP32 runs the two bottom levels (here 8x4x4 and 4x2x2 elements) in one
    single-work-group launch; FP64 runs them launch by launch.  The cycles
    agree to FP32 rounding, and the Jacobi form also matches the oracle."""
    print("case fused bottom", smoother, steps, power, flush=True)
    from core.multigrid import MGSolver
    dims = (32, 16, 16)
    E, fixed, ndof, rng = _problem(*dims, seed=12)
    kw = dict(max_levels=4, smoother=smoother, level_smooth=steps, cheb_ratio=0.1,
              cheb_ratio_power=power)
    if smoother == "jacobi":
        kw["omega"] = 0.35
    g = MGSolver(*dims, fixed, nu=0.3, compute_mode="GPU", **kw)
    g.set_density(E)
    assert [lv["dims"] for lv in g.levels] == [(32, 16, 16), (16, 8, 8), (8, 4, 4), (4, 2, 2)]
    o = MGSolverRef(*dims, fixed, nu=0.3, max_levels=4, omega=0.35, level_smooth=steps)
    o.set_density(E)
    b = np.where(o.levels[0]["free"], rng.standard_normal(ndof), 0.0)
    z64 = g._vcycle(0, b)
    z32 = g._vcycle(0, b, precision="mixed")
    assert _rel(z32, z64) < 1e-4, _rel(z32, z64)
    if smoother == "jacobi":
        assert _rel(z64, o._vcycle(0, b)) < 1e-11
    g.close()


@pytest.mark.parametrize("dims,kw,direct", [
    ((16, 8, 8), {}, True), ((8, 8, 8), {"max_levels": 3}, True),
    ((6, 10, 4), {"n_smooth": 3}, True), ((24, 16, 16), {"max_levels": 3}, True),
    ((40, 34, 20), {"max_levels": 2}, False),    # 12474 coarse DOFs: sweeps kept
])
def test_vcycle_direct_coarse(dims, kw, direct):
    """coarse="direct": the coarsest level solved exactly (FP64 inverse)."""
    print("case vcycle direct", dims, kw, flush=True)
    from core.multigrid import MGSolver
    nx, ny, nz = dims
    E, _, ndof, rng = _problem(*dims, seed=9)
    face = np.arange(3 * (nx + 1) * (ny + 1))
    fixed = np.unique(np.concatenate([face, rng.choice(ndof, 25, replace=False)]))
    g = MGSolver(nx, ny, nz, fixed, nu=0.3, compute_mode="GPU", coarse="direct", **kw)
    g.set_density(E)
    o = MGSolverRef(nx, ny, nz, fixed, nu=0.3, coarse="direct", **kw)
    o.set_density(E)
    assert (o.coarse_chol is not None) == direct
    b = np.where(o.levels[0]["free"], rng.standard_normal(ndof), 0.0)
    z = o._vcycle(0, b)
    assert _rel(g._vcycle(0, b), z) < 1e-10
    assert _rel(g._vcycle(0, b, precision="mixed"), z) < 1e-4
    g.close()


def test_vcycle_direct_coarse_singular_keeps_sweeps():
    """Supports only on nodes that are not coarse nodes: the coarse free block
    is singular, the library keeps the 30 sweeps."""
    print("case vcycle direct singular", flush=True)
    from core.multigrid import MGSolver
    dims = (6, 6, 6)
    E, _, ndof, rng = _problem(*dims, seed=3)
    node = 1 + 7 * (1 + 7 * 1)
    fixed = np.array([3 * node, 3 * node + 1, 3 * node + 2])
    g = MGSolver(*dims, fixed, nu=0.3, compute_mode="GPU", coarse="direct", max_levels=2)
    g.set_density(E)
    o = MGSolverRef(*dims, fixed, nu=0.3, max_levels=2)
    o.set_density(E)
    b = np.where(o.levels[0]["free"], rng.standard_normal(ndof), 0.0)
    assert _rel(g._vcycle(0, b), o._vcycle(0, b)) < 1e-11
    g.close()


@pytest.mark.parametrize("k,power", [(1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0), (3, 2),
                                     (6, 2), (9, 1)])
@pytest.mark.parametrize("start", ["zero", "u"])
def test_chebyshev_smoother_closed_form(k, power, start):
    """k Chebyshev steps on D^-1 A over [ratio*lam, lam] leave the error
    P_k(D^-1 A) e0 with P_k(x) = T_k((theta - x)/delta) / T_k(theta/delta):
    checked against that matrix polynomial built densely on a tiny grid, in
    FP64 and FP32.  With cheb_ratio_power p the interval of a k-step pass
    starts at ratio * (n_smooth / k)^p * lam."""
    from core.multigrid import MGSolver
    from oracle.fea_ref import hex8_KE
    from oracle.multigrid_ref import MGSolverRef
    print(f"chebyshev closed form k={k} power={power} start={start}", flush=True)
    rng = np.random.default_rng(10 + k)
    nx, ny, nz = 3, 2, 3
    fixed = np.arange(3 * (nx + 1) * (ny + 1))          # z = 0 plane
    E = rng.uniform(0.5, 2.0, nx * ny * nz)
    ratio = 0.3
    ref = MGSolverRef(nx, ny, nz, fixed, max_levels=1)
    ref.set_density(E)
    n = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    A = np.stack([ref._Ku(0, np.eye(n)[j]) for j in range(n)], axis=1)
    free = ref.levels[0]["free"]
    f = np.flatnonzero(free)
    Af = A[np.ix_(f, f)]
    Dinv = 1.0 / np.diag(Af)
    KE = hex8_KE(1.0, 0.3)
    lam = np.linalg.eigvalsh(KE).max() / KE[0, 0]
    a_, b_ = ratio * (2.0 / k) ** power * lam, lam
    theta, delta = 0.5 * (b_ + a_), 0.5 * (b_ - a_)
    M = Dinv[:, None] * Af
    Y = (theta * np.eye(len(f)) - M) / delta
    Tm, T = np.eye(len(f)), Y.copy()
    tm, t = 1.0, theta / delta
    for _ in range(k - 1):
        Tm, T = T, 2 * Y @ T - Tm
        tm, t = t, 2 * (theta / delta) * t - tm
    P = T / t
    b = rng.standard_normal(n)
    b[np.setdiff1d(np.arange(n), f)] = 0.0
    u0 = np.zeros(n) if start == "zero" else rng.standard_normal(n) * (free > 0)
    xs = np.linalg.solve(Af, b[f])
    want = np.zeros(n)
    want[f] = xs - P @ (xs - u0[f])
    mg = MGSolver(nx, ny, nz, fixed, compute_mode="GPU", smoother="chebyshev", cheb_ratio=ratio,
                  max_levels=1, cheb_ratio_power=power)
    mg.set_density(E)
    got = mg._smooth(0, None if start == "zero" else u0, b, k, "fp64")
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    # FP32 (two steps from u run as two damped-Jacobi steps weighted by the
    # reciprocal roots of the same polynomial): FP32 rounding only
    got32 = mg._smooth(0, None if start == "zero" else u0, b, k, "mixed")
    assert np.abs(got32 - want).max() <= 1e-5 * np.abs(want).max()
    mg.close()


def test_tile_results_independent_of_z_chunk():
    """The z-chunk only changes how planes are split over work-groups: every
    FP32/FP64 smoothing step, the operator and the V-cycle are bitwise the
    same for chunks that start on even and on odd planes (the layer loop is
    unrolled by two, and both copies must round alike)."""
    print("case z-chunk independence 40x24x24", flush=True)
    from core.multigrid import MGSolver
    dims = (40, 24, 24)
    E, fixed, ndof, rng = _problem(*dims, seed=21)
    ms = {zc: MGSolver(*dims, fixed, nu=0.3, compute_mode="GPU", smoother="chebyshev",
                       cheb_ratio=0.05, z_chunk=zc) for zc in (10, 5, 7, 25)}
    for m in ms.values():
        m.set_density(E)
    u = rng.standard_normal(ndof)
    b = rng.standard_normal(ndof)
    ops = [lambda m: m._smooth(0, None, b, 2, precision="mixed"),
           lambda m: m._smooth(0, u, b, 2, precision="mixed"),
           lambda m: m._smooth(0, u, b, 3, precision="mixed"),
           lambda m: m._smooth(0, u, b, 2),
           lambda m: m._apply(0, u, precision="mixed"),
           lambda m: m._vcycle(0, b, precision="mixed"),
           lambda m: m._vcycle(0, b)]
    for op in ops:
        ref = op(ms[10])
        for zc in (5, 7, 25):
            assert np.array_equal(op(ms[zc]), ref), zc
    for m in ms.values():
        m.close()


@pytest.mark.parametrize("smoother,start", [("chebyshev", "zero"), ("chebyshev", "u"),
                                            ("jacobi", "zero"), ("jacobi", "u")])
def test_blocked_smoother_matches_step_launches(smoother, start):
    """FP32 smoothing passes on the levels below level 0 with at most 40000
    nodes run several steps per launch (k_smooth_block; each work-group
    repeats the halo work of the steps it fuses).  They agree with one tile
    launch per step (block_steps=1) and with FP64 (the oracle's steps) to FP32
    rounding: the two kernels sum a node's element forces in other orders.
    Between box edges and steps per launch the result is bitwise the same,
    since every node sums the same terms in the same order."""
    from core.multigrid import MGSolver
    print("case blocked smoother", smoother, start, flush=True)
    dims = (40, 24, 32)
    E, fixed, ndof, rng = _problem(*dims, seed=31)
    kw = dict(smoother=smoother, cheb_ratio=0.05, max_levels=3)
    if smoother == "jacobi":
        kw["omega"] = 0.35
    variants = [(1, 0), (0, 0), (2, 0), (5, 0), (8, 0), (3, 2), (3, 5), (4, 7)]
    ms = {v: MGSolver(*dims, fixed, nu=0.3, compute_mode="GPU", block_steps=v[0],
                      block_edge=v[1], **kw) for v in variants}
    for m in ms.values():
        m.set_density(E)
    m0 = ms[(1, 0)]
    for li in (1, 2):
        n = m0.levels[li]["ndof"]
        b = rng.standard_normal(n)
        u = None if start == "zero" else rng.standard_normal(n)
        for k in (2, 3, 7, 12):
            ref64 = m0._smooth(li, u, b, k, "fp64")
            tile = m0._smooth(li, u, b, k, "mixed")
            got = {v: ms[v]._smooth(li, u, b, k, "mixed") for v in variants[1:]}
            tol = 5e-6 * k
            assert _rel(tile, ref64) < tol, (li, k)
            blocked = got[(0, 0)]
            assert _rel(blocked, ref64) < tol, (li, k, _rel(blocked, ref64))
            assert _rel(blocked, tile) < tol, (li, k)
            for v, g in got.items():
                assert np.array_equal(g, blocked), (li, k, v)
    for m in ms.values():
        m.close()


@pytest.mark.parametrize("kw", [
    dict(max_levels=5, level_smooth=[2, 2, 6, 12], coarse_sweeps=48, cheb_ratio_power=2),
    dict(max_levels=4, level_smooth=[2, 3, 9], coarse_sweeps=24),
])
def test_vcycle_blocked_coarse_levels(kw):
    """A mixed V-cycle whose coarse passes run blocked equals the one launched
    step by step to FP32 rounding, and the FP64 cycle to the usual mixed bar."""
    from core.multigrid import MGSolver
    print("case vcycle blocked", kw, flush=True)
    dims = (64, 32, 48)
    E, fixed, ndof, rng = _problem(*dims, seed=32)
    common = dict(smoother="chebyshev", cheb_ratio=0.05, **kw)
    mb = MGSolver(*dims, fixed, nu=0.3, compute_mode="GPU", **common)
    mt = MGSolver(*dims, fixed, nu=0.3, compute_mode="GPU", block_steps=1, **common)
    for m in (mb, mt):
        m.set_density(E)
    b = np.where(rng.random(ndof) < 0.9, rng.standard_normal(ndof), 0.0)
    zb = mb._vcycle(0, b, precision="mixed")
    zt = mt._vcycle(0, b, precision="mixed")
    z64 = mt._vcycle(0, b)
    assert _rel(zb, zt) < 1e-5, _rel(zb, zt)
    assert _rel(zb, z64) < 1e-4
    mb.close()
    mt.close()


def test_autotuned_block_plans_solve_like_tile_launches():
    """vx_autotune times each small level's FP32 pass as tile launches and as
    blocked launches and keeps the faster form; the solve with the tuned
    forms converges like the one with tile launches only (the two differ by
    the rounding of a node's force sum)."""
    import torch
    from core.multigrid import MGSolver
    print("case autotuned block plans", flush=True)
    dims = (64, 64, 32)
    E, fixed, ndof, rng = _problem(*dims, seed=33)
    kw = dict(precision="mixed", smoother="chebyshev", cheb_ratio=0.05, max_levels=5,
              level_smooth=[2, 2, 6, 12], coarse_sweeps=48, cheb_ratio_power=2)
    f = torch.tensor(np.where(rng.random(ndof) < 0.5, rng.standard_normal(ndof), 0.0),
                     device="cuda")
    res = {}
    for name, extra in (("tuned", dict(autotune=True)), ("tile", dict(block_steps=1))):
        m = MGSolver(*dims, fixed, nu=0.3, compute_mode="GPU", **kw, **extra)
        m.set_density(E)
        u = torch.empty_like(f)
        m.solve_device(f, u, tol=1e-8)
        plans = m.block_plans()
        for li, (edge, per) in plans.items():
            assert li >= 1 and 2 <= edge <= 8 and 2 <= per <= 4
        res[name] = (u.cpu().numpy(), m.last_iters, m.last_true_resid_ratio, plans)
        m.close()
    assert res["tile"][3] == {}
    (ut, it_t, tr_t, _), (u1, it_1, tr_1, _) = res["tuned"], res["tile"]
    print("  tuned plans", res["tuned"][3], "iters", it_t, it_1, flush=True)
    assert tr_t < 1e-8 and tr_1 < 1e-8
    assert abs(it_t - it_1) <= 2
    assert _rel(ut, u1) < 1e-6
