"""Parts of uneven size in one solve (vx_solve_parts, core.multigrid.PartsSolver).

* FP64: every part's u and iteration count equal, bit for bit, its own
  MGSolver solve with the same options (the reference solves each part with
  its own VoxelFEA, core/fea.py:244-305) -- for parts of different depths,
  one that does not coarsen at all (odd edges: its level 0 is its coarsest),
  a zero load, the plain reference smoother and the Chebyshev one, padded and
  not.  The tile z-chunks are fixed so that both sides group their dot
  products alike.
* Mixed: the c4 configuration (PLAIN_CYCLE, padded parts) solved in one call
  meets the 1e-5 bar against the reference's solutions of two uneven parts
  (tests/golden/configs).
"""

import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _part(shape, seed, zero_load=False):
    from core import workloads as W
    rng = np.random.default_rng(seed)
    kept, fixed = W._random_part(rng, shape, (0.3, 0.8))
    nx, ny, nz = shape
    ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    f = np.zeros(ndof)
    if not zero_load:
        d = rng.standard_normal(3)
        f = W.spread_load(ndof, W._top_patch_nodes(rng, kept, patch=4), 800.0 * d / np.linalg.norm(d))
    return dict(shape=shape, active=kept.ravel(), E=np.full(nx * ny * nz, 2.0e6), fixed=fixed,
                f=f)


PARTS = [((16, 12, 8), 1, False), ((12, 12, 12), 2, False), ((9, 10, 7), 3, False),
         ((8, 8, 4), 4, True), ((20, 8, 12), 5, False)]


def _solve_both(parts, tol, **kw):
    import torch
    from core.multigrid import MGSolver, PartsSolver
    ps = PartsSolver([p["shape"] for p in parts], [p["fixed"] for p in parts], nu=0.3,
                     compute_mode="GPU", **kw)
    ps.set_density([p["E"] for p in parts], [p["active"] for p in parts])
    U, _ = ps.solve_device([torch.tensor(p["f"], device="cuda") for p in parts], tol=tol)
    singles = []
    for p in parts:
        mg = MGSolver(*p["shape"], p["fixed"], nu=0.3, compute_mode="GPU", **kw)
        E = torch.tensor(p["E"], device="cuda")
        A = torch.tensor(p["active"].astype(np.uint8), device="cuda")
        f = torch.tensor(p["f"], device="cuda")
        u = torch.empty_like(f)
        mg.set_density(E, active=A)
        mg.solve_device(f, u, tol=tol)
        singles.append((u.cpu().numpy(), mg.last_iters, mg.last_converged, len(mg.levels)))
        mg.close()
    res = [(u.cpu().numpy(), info) for u, info in zip(U, ps.last_parts)]
    ps.close()
    return res, singles


@pytest.mark.parametrize("kw,one_level", [
    # the reference's V(2,2) damped Jacobi (omega 0.6): a part that does not
    # coarsen would take 30 of its sweeps, which amplify the highest modes
    # (omega > 2 / lambda_max), as its whole preconditioner, and stall: here
    # only parts that coarsen
    (dict(), False),
    (dict(smoother="chebyshev", cheb_ratio=0.1, system="active", max_levels=4), True),
    (dict(smoother="chebyshev", cheb_ratio=0.05, system="active", max_levels=5, pad=True,
          level_smooth=[2, 2, 6], coarse_sweeps=24, cheb_ratio_power=2), True),
], ids=["jacobi-full", "chebyshev-active", "chebyshev-padded-variable"])
def test_fp64_parts_equal_single_solves(kw, one_level):
    print("case parts fp64", kw, flush=True)
    spec = [p for p in PARTS if one_level or p[0] != (9, 10, 7)]
    parts = [_part(s, seed, z) for s, seed, z in spec]
    res, singles = _solve_both(parts, 1e-8, precision="fp64", z_chunk=4, **kw)
    depths = set()
    for (u, info), (us, it, conv, nl), (_, _, zero) in zip(res, singles, spec):
        print(f"  levels {info['levels']} iters {info['iters']} (single {it})", flush=True)
        depths.add(info["levels"])
        assert info["levels"] == nl
        assert info["converged"] == conv
        assert info["iters"] == it
        assert np.array_equal(u, us)
        if zero:
            assert np.all(u == 0)
    assert len(depths) >= 2          # the parts really differ in depth


def test_mixed_parts_equal_single_mixed_solves():
    """Mixed precision: the parts solve and the single solves reach the same
    u to the tolerance's rounding, each part to its own true residual."""
    print("case parts mixed vs single", flush=True)
    parts = [_part(s, seed, z) for s, seed, z in PARTS]
    kw = dict(precision="mixed", smoother="chebyshev", cheb_ratio=0.05, system="active",
              max_levels=5, pad=True, inner_tol=1e-5)
    res, singles = _solve_both(parts, 1e-9, **kw)
    for (u, info), (us, it, conv, nl) in zip(res, singles):
        print(f"  levels {info['levels']} iters {info['iters']} (single {it}) true "
              f"{info['true_resid_ratio']:.2e}", flush=True)
        assert info["converged"] and info["true_resid_ratio"] < 1e-9
        if np.abs(us).max() > 0:
            assert np.abs(u - us).max() <= 1e-6 * np.abs(us).max()
        else:
            assert np.all(u == 0)


def test_c4_parts_in_one_call_match_reference_solutions():
    """Both uneven-part fixtures (the reference's Jacobi-PCG solutions on the
    active-DOF system) solved in one vx_solve_parts call with the benchmark's
    c4 options (PLAIN_CYCLE on void-padded grids)."""
    import torch
    from core import workloads as W
    from core.fea import VoxelFEA
    from core.multigrid import PLAIN_CYCLE, PartsSolver
    print("case parts c4 fixtures", flush=True)
    gs = []
    for name in ("c4_uneven_part_a", "c4_uneven_part_b"):
        g = np.load(os.path.join(GOLDEN, "configs", name + ".npz"))
        gs.append({k: g[k] for k in g.files})
    dims = [(int(g["nx"]), int(g["ny"]), int(g["nz"])) for g in gs]
    ps = PartsSolver(dims, [g["fixed_dofs"] for g in gs], nu=float(gs[0]["nu"]), compute_mode="GPU",
                     pad=True, **PLAIN_CYCLE)
    ps.set_density([g["E_pa"] * float(g["vsize"]) for g in gs], [g["active"] for g in gs])
    U, _ = ps.solve_device([torch.tensor(g["force"], device="cuda") for g in gs], tol=1e-8)
    for g, d, u, info in zip(gs, dims, U, ps.last_parts):
        u = u.cpu().numpy()
        used = np.repeat(W.element_nodes_mask(g["active"].reshape(d)).ravel(), 3)
        eu = np.abs(u[used] - g["u"][used]).max() / np.abs(g["u"][used]).max()
        vf = VoxelFEA(*d, nu=float(g["nu"]), active_elems=g["active"], compute_mode="GPU")
        vm = vf.element_von_mises_stress(u, g["E_pa"]) / float(g["vsize"])
        vf.close()
        ev = np.abs(vm - g["vm"]).max() / np.abs(g["vm"]).max()
        print(f"  {d}: levels {info['levels']} iters {info['iters']} true "
              f"{info['true_resid_ratio']:.2e} rel u err {eu:.2e} rel vm err {ev:.2e}", flush=True)
        assert info["converged"] and info["true_resid_ratio"] < 1e-8
        assert eu <= 1e-5 and ev <= 1e-5
        assert np.all(u[~used] == 0)
    ps.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_parts_graph_replay_equals_eager_launches(precision):
    """Each PCG step's parts V-cycle replayed from a captured HIP graph gives
    the bitwise result of launching it kernel by kernel, over two solves in a
    row and a new density (the graph is captured once and replayed)."""
    import torch
    from core.multigrid import PartsSolver
    print("case parts graph vs eager", precision, flush=True)
    parts = [_part(s, seed, z) for s, seed, z in PARTS]
    kw = dict(precision=precision, smoother="chebyshev", cheb_ratio=0.05, system="active",
              max_levels=5, pad=True, level_smooth=[2, 2, 6], coarse_sweeps=24,
              cheb_ratio_power=2, inner_tol=1e-5)
    out = {}
    for graphs in (True, False):
        ps = PartsSolver([p["shape"] for p in parts], [p["fixed"] for p in parts], nu=0.3,
                         compute_mode="GPU", graphs=graphs, **kw)
        res = []
        for scale in (1.0, 1.0, 0.5):
            ps.set_density([p["E"] * scale for p in parts], [p["active"] for p in parts])
            U, _ = ps.solve_device([torch.tensor(p["f"], device="cuda") for p in parts], tol=1e-8)
            res.append(([u.cpu().numpy() for u in U], [i["iters"] for i in ps.last_parts]))
        out[graphs] = res
        ps.close()
    for (ug, ig), (ue, ie) in zip(out[True], out[False]):
        assert ig == ie
        assert all(np.array_equal(a, b) for a, b in zip(ug, ue))


def test_parts_autotune_keeps_the_solution():
    """vx_parts_autotune picks each level's tile z-chunk by timing.  The
    z-chunk only regroups the FP64 dot-product partials, so tuned solves agree
    with fixed-z-chunk ones to rounding, in FP64 and mixed precision."""
    import torch
    from core.multigrid import PartsSolver
    print("case parts autotune", flush=True)
    parts = [_part(s, seed, z) for s, seed, z in PARTS]
    kw = dict(smoother="chebyshev", cheb_ratio=0.05, system="active", max_levels=5, pad=True,
              compute_mode="GPU", nu=0.3)
    F = [torch.tensor(p["f"], device="cuda") for p in parts]
    for precision, bar in (("fp64", 1e-9), ("mixed", 1e-6)):
        out = {}
        for name, extra in (("tuned", dict(autotune=True)), ("fixed", dict(z_chunk=4))):
            ps = PartsSolver([p["shape"] for p in parts], [p["fixed"] for p in parts],
                             precision=precision, **kw, **extra)
            ps.set_density([p["E"] for p in parts], [p["active"] for p in parts])
            assert ps.tuned == (name == "tuned")
            zc = ps.zchunks()
            assert all(len(zc[i]) == n and min(zc[i]) >= 1 for i, n in enumerate(ps.levels))
            print(f"  {precision} {name} z-chunks {zc}", flush=True)
            U, _ = ps.solve_device(F, tol=1e-8)
            out[name] = ([u.cpu().numpy() for u in U], [i["iters"] for i in ps.last_parts],
                         [i["converged"] for i in ps.last_parts])
            ps.close()
        (ut, it_t, ct), (uf, it_f, cf) = out["tuned"], out["fixed"]
        print(f"  {precision} iterations tuned {it_t} fixed {it_f}", flush=True)
        assert all(ct) and all(cf)
        assert all(abs(a - b) <= 1 for a, b in zip(it_t, it_f))
        for a, b in zip(ut, uf):
            scale = np.abs(b).max()
            assert scale == 0 or np.abs(a - b).max() <= bar * scale
