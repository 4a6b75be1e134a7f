"""Benchmark: time to a 1e-6 relative residual on the 256^3 two-material hex8
part, MI355X mixed-precision MG-PCG vs the plain FP64 GPU MG-PCG.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
                    [--ref-iters N]

A "step" is one complete solve of the workload -- density set-up of the
hierarchy, PCG with V-cycle preconditioning to tol, FP64 true-residual check --
with f, E and the void mask already resident in HBM.  One JSON line goes to
stdout; progress lines go to stderr.

Comparators (both FP64, same kernels):
  * the reference algorithm (damped Jacobi omega 0.6 V(2,2), 30 coarsest
    sweeps, voids at E*e_min): on this part its stall guard fires and the
    reference falls back to Jacobi-PCG on the active-DOF system
    (core/fea.py:476-489, 497-593); that fallback is run for --ref-iters
    iterations and its time and reached ratio are reported;
  * FP64 MG-PCG with this build's preconditioner (Chebyshev smoother, active
    system), i.e. the same solve without the precision change.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "voxelsim-fea_amd"))

HBM_PEAK_GBS = 8000.0

# the MI355X solver configurations: OPT (variable V-cycle over 5 levels) is
# measured as `value`; OPT_PLAIN (plain 7-level cycle) runs the single-case
# workloads that converge in 19-100 iterations
from core.multigrid import PLAIN_CYCLE as OPT_PLAIN, VARIABLE_CYCLE as OPT  # noqa: E402
# the uneven parts in one vx_solve_parts call: the plain cycle with 16
# coarsest sweeps (the one-work-group coarsest smoothing of their 1000-4000-DOF
# coarsest levels is serial work on a few CUs; 16 sweeps cost 5 % more
# iterations and save 9 % of the time against 30)
OPT_PARTS = dict(OPT_PLAIN, coarse_sweeps=16)

# the roofline kernel: the longest level-0 launch of a mixed PCG step at 256^3
# in the kernel trace (profiles/bench_two_material_256_kernel_trace.md), the
# pre-smoothing pass: two Chebyshev steps from zero in one launch (b 3 x FP32,
# D^-1 FP32 and the mask byte per node and E per element in, x 3 x FP32 out),
# and its counter summary
ROOF_KERNEL = "chebyshev2_zero_fp32"
ROOF_KERNEL_NAME = ("k_tile<float, M_CHEB_Z2> (the pre-smoothing pass: two Chebyshev steps from "
                    "zero in one launch), level 0")
TRAFFIC_PROFILE = "profiles/level0_fp32_256_traffic.md"


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


class _StopSolve(Exception):
    """Raised by the timing callback to end an oracle solve early."""


def cpu_baseline(p, ref_iters, single_iters=2, pool_iters=1):
    """The oracle (NumPy restatement of the reference algorithm) on this
    host: the plain FP64 MG-PCG -- damped Jacobi omega 0.4, V(2,2), 30
    coarsest sweeps, 6 levels, voids at E*e_min on the full grid
    (core/multigrid.py:85-583, core/fea.py:467-469) -- on the same workload,
    in both of the reference's CPU modes.  Each run is a real solve whose
    progress callback stamps the end of every PCG iteration and stops it
    after a few; an iteration is one outer operator, its vector updates and
    one V-cycle.  Estimate = hierarchy build + set_density + ref_iters x the
    measured time per iteration (ref_iters: the same algorithm's iteration
    count on the GPU).  Modes: one process with nproc-1 BLAS threads
    (core/compute_plan.py:80-90, 186-188), and the element-partitioned pool
    of nproc-1 spawned one-thread workers for the outer operator with one
    BLAS thread in the main process (core/parallel_cpu.py:56-205,
    core/compute_plan.py:168-172, 181-188)."""
    import numpy as np
    from oracle.multigrid_ref import MGSolverRef
    # a one-GPU host may grant this process only 16 CPUs, whatever os.cpu_count() says
    threads = max(1, min((os.cpu_count() or 2) - 1, 15))
    try:
        import threadpoolctl
    except Exception:
        threadpoolctl = None

    def blas(n):
        if threadpoolctl is not None:
            threadpoolctl.threadpool_limits(limits=n)

    blas(threads if threadpoolctl else 1)
    log("cpu baseline: building the oracle hierarchy (omega 0.4, 6 levels)")
    t0 = time.perf_counter()
    mg = MGSolverRef(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, omega=0.4, max_levels=6)
    t_build = time.perf_counter() - t0
    E = np.where(p.active, p.E_stiffness, p.E_stiffness * 1e-9)

    def iterations(n):
        t0 = time.perf_counter()
        mg.set_density(E)
        t_density = time.perf_counter() - t0
        stamps = []

        def cb(it, ratio, tol):
            stamps.append(time.perf_counter())
            log(f"cpu baseline: iteration {it} ratio {ratio:.3e} at "
                f"{stamps[-1] - t_start:.1f} s")
            if it >= n:
                raise _StopSolve()
        t_start = time.perf_counter()
        try:
            mg.solve(E, p.force, tol=1e-6, max_cg=n + 1, progress_cb=cb)
        except _StopSolve:
            pass
        # the solve repeats set_density before its first iteration
        per_iter = (stamps[-1] - t_start - t_density) / len(stamps)
        return {"iterations": len(stamps), "set_density_s": t_density,
                "per_iteration_s": per_iter,
                "value": t_build + t_density + ref_iters * per_iter}

    single = iterations(single_iters)
    single["threads"] = threads
    log(f"cpu baseline, one process: {single}")
    pool = None
    try:
        from oracle.parallel_cpu_ref import CPUMatVecPoolRef
        log(f"cpu baseline: starting a {threads}-worker operator pool")
        lv0 = mg.levels[0]
        mg._pool = CPUMatVecPoolRef(lv0["edof"], mg.KE, lv0["ndof"], threads)
        blas(1)
        try:
            pool = iterations(pool_iters)
            pool["workers"] = mg._pool.n_workers
        finally:
            mg.close()
        log(f"cpu baseline, pool: {pool}")
    except Exception as e:   # the timing leg must not stop the bench
        log(f"cpu baseline: pool mode skipped ({e})")
    best = min(single["value"], pool["value"]) if pool else single["value"]
    mode = "pool" if pool and pool["value"] < single["value"] else "single process"
    return {"value": best, "unit": "s", "cores": threads, "kind": "port",
            "sample": (f"the reference algorithm (oracle: Jacobi omega 0.4, V(2,2), 6 levels, "
                       f"e_min voids) on the same {p.nx}^3 workload, stopped after "
                       f"{single['iterations']} complete PCG iterations in one process "
                       f"({single['per_iteration_s']:.1f} s each)"
                       + (f" and {pool['iterations']} with the {pool['workers']}-worker pool "
                          f"({pool['per_iteration_s']:.1f} s each)" if pool else "")
                       + f"; solve = hierarchy {t_build:.1f} s + set_density + {ref_iters} "
                       f"iterations (its GPU count); faster mode: {mode}; the shipped omega "
                       f"0.6 stalls (reference_fp64_mgpcg)"),
            "hierarchy_s": t_build, "single_process": single, "pool": pool,
            "cpu_model": cpu_model(), "ram_gb": ram_gb()}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def ram_gb():
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal"):
                return round(int(line.split()[1]) / 2**20, 1)
    except OSError:
        pass
    return None


def measured_traffic():
    """HBM bytes per launch of the dominant kernel, from the committed
    counter summary (two rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE,
    scaled by a 1 GiB copy); None when the summary is absent."""
    import re
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    try:
        m = re.search(r"= \*\*([0-9.]+) MB\*\*", open(path).read())
        return float(m.group(1)) * 1e6 if m else None
    except OSError:
        return None


def copy_bandwidth(nbytes=2 << 30, reps=10):
    """Achievable HBM bandwidth: torch's device copy of a buffer of known size
    (read + write bytes / time)."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return 2 * nbytes / dt / 1e9


def kernel_only(p, E_dev, act_dev, MGSolver, z_chunk):
    """The level-0 FP32 tile kernels that take longest in a PCG step, 2 + 20
    launches each at one fixed z-chunk: the last post-smoothing step (a
    damped-Jacobi step with its r.z partials), the pre-smoothing pass (two
    Chebyshev steps from zero in one launch) and the operator (masked K u);
    then a plain 1 GiB FP32 copy 5 times (the reference dispatch for counter
    scaling).  The z-chunk is fixed so that every counter pass sees the same
    launches."""
    import torch
    opt = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU",
                   **dict(OPT, autotune=False, z_chunk=z_chunk))
    opt.set_density(E_dev, active=act_dev)
    nn = (p.nx + 1) * (p.ny + 1) * (p.nz + 1)
    rec = {}
    for kind, nbytes in (("jacobi_dot", 36 * nn + nn + 4 * p.nelem),
                         ("chebyshev2_zero", 24 * nn + 4 * nn + nn + 4 * p.nelem),
                         ("matvec", 24 * nn + nn + 4 * p.nelem)):
        rec[kind] = {"ms": opt.time_kernel(kind, 0, "mixed", 20), "algorithmic_bytes": nbytes}
    a = torch.ones((1 << 30) // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"level0_fp32": rec, "z_chunk": z_chunk,
                      "copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30}), flush=True)


def _oracle_solve(p, max_levels=6, active_system=False):
    """Complete CPU oracle solve of a problem (reference algorithm)."""
    import numpy as np
    from oracle.fea_ref import VoxelFEARef
    vf = VoxelFEARef(p.nx, p.ny, p.nz, nu=p.nu, active_elems=p.active,
                     use_multigrid=not active_system, mg_kwargs={"max_levels": max_levels})
    vf.set_fixed(p.fixed_dofs)
    t0 = time.perf_counter()
    u = vf.solve(p.E_stiffness, p.force if p.force.ndim == 1 else p.force[0], tol=p.tol)
    return u, time.perf_counter() - t0, vf.last_iters, vf.last_converged


def vm_rel_err(q, u, ref):
    """Largest element von Mises difference over the largest von Mises of the
    reference displacement (both with the part's E; the GPU kernel of
    core/fea.py:604-626)."""
    import numpy as np
    from core.fea import VoxelFEA
    vf = VoxelFEA(q.nx, q.ny, q.nz, nu=q.nu, active_elems=q.active, compute_mode="GPU")
    a = vf.element_von_mises_stress(u, q.E_stiffness)
    b = vf.element_von_mises_stress(ref, q.E_stiffness)
    vf.close()
    return float(np.abs(a - b).max() / np.abs(b).max())


def uneven_parts_record():
    """c4: 16 uneven ABS parts on void-padded grids (plain cycle) in one
    vx_solve_parts call, beside the same parts solved one after another;
    error against FP64 Jacobi-PCG of each part to 1e-10."""
    import numpy as np
    import torch
    from core import workloads as W
    from core.multigrid import MGSolver, PartsSolver

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def used_of(q):
        return np.repeat(W.element_nodes_mask(q.active.reshape(q.nx, q.ny, q.nz)).ravel(), 3)

    def rel(u, ref, used):
        return float(np.abs(u[..., used] - ref[..., used]).max() / np.abs(ref[..., used]).max())

    def dev(q):
        return (torch.tensor(q.E_stiffness, device="cuda"),
                torch.tensor(q.active.astype(np.uint8), device="cuda"),
                torch.tensor(q.force, device="cuda"))

    # c4: 16 uneven ABS parts on void-padded grids (plain cycle), all in one
    # vx_solve_parts call; beside it the same parts solved one after another
    log("configs: c4 uneven parts")
    parts = W.uneven_parts()
    devs = [dev(q) for q in parts]
    exact, errs, single_total = [], [], 0.0
    for q, (E, A, f) in zip(parts, devs):
        # FP64 Jacobi-PCG on the same active-DOF system (the reference's
        # fallback, core/fea.py:497-593): the FP64 MG form's stall guard can
        # fire on these parts' oscillating residual
        u = torch.empty_like(f)
        ref = MGSolver(q.nx, q.ny, q.nz, q.fixed_dofs, nu=q.nu, compute_mode="GPU", max_levels=1)
        ref.solve_jacobi_device(E, A, f, u, tol=1e-10)
        exact.append(u.cpu().numpy())
        ref.close()
        mg = MGSolver(q.nx, q.ny, q.nz, q.fixed_dofs, nu=q.nu, compute_mode="GPU", pad=True,
                      **OPT_PLAIN)
        single_total += timed(lambda: (mg.set_density(E, active=A),
                                       mg.solve_device(f, u, tol=q.tol, max_cg=5000)))
        mg.close()
    ps = PartsSolver([(q.nx, q.ny, q.nz) for q in parts], [q.fixed_dofs for q in parts],
                     nu=parts[0].nu, compute_mode="GPU", pad=True, **OPT_PARTS)
    got = {}

    def run_parts():
        ps.set_density([d[0] for d in devs], [d[1] for d in devs])
        got["U"], _ = ps.solve_device([d[2] for d in devs], tol=parts[0].tol, max_cg=5000)
    t = timed(run_parts)
    vm_errs = []
    for q, u, ue in zip(parts, got["U"], exact):
        errs.append(rel(u.cpu().numpy(), ue, used_of(q)))
        vm_errs.append(vm_rel_err(q, u.cpu().numpy(), ue))
    rec = {
        "time_s": t, "form": "one vx_solve_parts call (per-part recurrences)",
        "solver": OPT_PARTS,
        "one_solve_per_part_time_s": single_total, "speedup_vs_one_solve_per_part": single_total / t,
        "iters": [i["iters"] for i in ps.last_parts],
        "levels": [i["levels"] for i in ps.last_parts],
        "true_resid_max": max(i["true_resid_ratio"] for i in ps.last_parts),
        "all_converged": all(i["converged"] for i in ps.last_parts),
        "rel_disp_err_vs_fp64_jacobi_1e-10": max(errs),
        "rel_vm_err_vs_fp64_jacobi_1e-10": max(vm_errs)}
    ps.close()
    return rec


def config_records():
    """Compact records of the other four configurations for the headline line:
    time to tolerance with inputs resident in HBM (second of two runs; the
    first builds the autotuned launches and graphs), iterations, true
    residual and the relative displacement error on the nodes of active
    elements and the relative element von Mises error -- against the reference's own solution where this exact
    problem has one (c1: tests/golden/cantilever.npz, written by the
    reference), else against an FP64 solve to 1e-10 (c2, c3: with the same
    preconditioner; c4: Jacobi-PCG, the reference's fallback) at full size;
    their scaled-down instances are compared with the reference's solutions
    in tests/test_gpu_parity_configs.py."""
    import numpy as np
    import torch
    from core import workloads as W
    from core.multigrid import MGSolver

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def used_of(q):
        return np.repeat(W.element_nodes_mask(q.active.reshape(q.nx, q.ny, q.nz)).ravel(), 3)

    def rel(u, ref, used):
        return float(np.abs(u[..., used] - ref[..., used]).max() / np.abs(ref[..., used]).max())

    def dev(q):
        return (torch.tensor(q.E_stiffness, device="cuda"),
                torch.tensor(q.active.astype(np.uint8), device="cuda"),
                torch.tensor(q.force, device="cuda"))

    out = {}
    # c1: the reference algorithm in FP64 (damped Jacobi 0.6, 3 levels, tol 1e-8)
    log("configs: c1 cantilever")
    q = W.cantilever()
    E, A, f = dev(q)
    u = torch.empty_like(f)
    ref_u = np.load(os.path.join(ROOT, "tests", "golden", "cantilever.npz"))["u"]
    rec = {}
    for name, kw in (("reference_fp64", dict(max_levels=q.max_levels)),
                     ("mixed", dict(OPT_PLAIN, max_levels=q.max_levels))):
        mg = MGSolver(q.nx, q.ny, q.nz, q.fixed_dofs, nu=q.nu, compute_mode="GPU", **kw)
        t = timed(lambda: (mg.set_density(E, active=A, e_min=1e-9),
                           mg.solve_device(f, u, tol=q.tol)))
        rec[name] = {"time_s": t, "iters": mg.last_iters,
                     "true_resid": mg.last_true_resid_ratio,
                     "rel_disp_err_vs_reference": rel(u.cpu().numpy(), ref_u, used_of(q)),
                     "rel_vm_err_vs_reference": vm_rel_err(q, u.cpu().numpy(), ref_u)}
        mg.close()
    out["c1_cantilever_64x16x16"] = rec
    # c2: 128^3 aluminium block with 40 spherical voids, plain 7-level cycle
    log("configs: c2 voids block")
    q = W.voids_block()
    E, A, f = dev(q)
    u = torch.empty_like(f)
    exact = MGSolver(q.nx, q.ny, q.nz, q.fixed_dofs, nu=q.nu, compute_mode="GPU",
                     **dict(OPT_PLAIN, precision="fp64"))
    exact.set_density(E, active=A)
    exact.solve_device(f, u, tol=1e-10)
    u_ex = u.cpu().numpy()
    exact.close()
    mg = MGSolver(q.nx, q.ny, q.nz, q.fixed_dofs, nu=q.nu, compute_mode="GPU", **OPT_PLAIN)
    t = timed(lambda: (mg.set_density(E, active=A), mg.solve_device(f, u, tol=q.tol)))
    out["c2_voids_block_128"] = {"time_s": t, "iters": mg.last_iters,
                                 "true_resid": mg.last_true_resid_ratio,
                                 "rel_disp_err_vs_fp64_1e-10": rel(u.cpu().numpy(), u_ex,
                                                                   used_of(q)),
                                 "rel_vm_err_vs_fp64_1e-10": vm_rel_err(q, u.cpu().numpy(),
                                                                        u_ex)}
    mg.close()
    # c3: 96^3 part, 32 load cases in one batched solve (variable cycle)
    log("configs: c3 multi-RHS part")
    q = W.multi_rhs_part()
    E, A, F = dev(q)
    U = torch.empty_like(F)
    nrhs = F.shape[0]
    exact = MGSolver(q.nx, q.ny, q.nz, q.fixed_dofs, nu=q.nu, compute_mode="GPU",
                     max_batch=nrhs, **dict(OPT, precision="fp64"))
    exact.set_density(E, active=A)
    exact.solve_batch_device(F, U, tol=1e-10)
    U_ex = U.cpu().numpy()
    exact.close()
    mg = MGSolver(q.nx, q.ny, q.nz, q.fixed_dofs, nu=q.nu, compute_mode="GPU", max_batch=nrhs,
                  **OPT)
    t = timed(lambda: (mg.set_density(E, active=A), mg.solve_batch_device(F, U, tol=q.tol)))
    out["c3_multi_rhs_96x32"] = {
        "time_s": t, "load_cases_per_s": nrhs / t,
        "iters_max": max(b["iters"] for b in mg.last_batch),
        "true_resid_max": max(b["true_resid_ratio"] for b in mg.last_batch),
        "rel_disp_err_vs_fp64_1e-10": rel(U.cpu().numpy(), U_ex, used_of(q)),
        "rel_vm_err_vs_fp64_1e-10": max(vm_rel_err(q, Ur, Ue)
                                        for Ur, Ue in zip(U.cpu().numpy(), U_ex))}
    mg.close()
    out["c4_uneven_parts_16"] = uneven_parts_record()
    return out


def other_workload(args):
    """Measurement of one of the other configurations (not the headline)."""
    import numpy as np
    import torch
    from core import _native
    from core import workloads as W
    from core.multigrid import MGSolver

    def dev(a, dt=None):
        return torch.tensor(a, device="cuda") if dt is None else torch.tensor(a, device="cuda",
                                                                              dtype=dt)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    out = {"workload": args.workload}
    if args.workload != "inside":
        out["solver"] = OPT_PLAIN
    if args.workload in ("cantilever", "voids_block"):
        p = W.cantilever() if args.workload == "cantilever" else W.voids_block()
        E, A, f = dev(p.E_stiffness), dev(p.active.astype(np.uint8)), dev(p.force)
        u = torch.empty_like(f)
        used = np.repeat(W.element_nodes_mask(p.active.reshape(p.nx, p.ny, p.nz)).ravel(), 3) \
            if not p.active.all() else np.ones(p.ndof, bool)
        res, sol = {}, {}
        variants = (("reference_fp64", dict(max_levels=p.max_levels)),
                    ("mixed_full", dict(OPT_PLAIN, system="full", max_levels=p.max_levels)),
                    ("mixed_active", dict(OPT_PLAIN, max_levels=p.max_levels)))
        for name, kw in variants:
            mg = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU", **kw)
            run = lambda: (mg.set_density(E, active=A, e_min=1e-9),
                           mg.solve_device(f, u, tol=p.tol))
            try:
                timed(run)
                t, _ = timed(run)
                res[name] = {"time_s": t, "iters": mg.last_iters,
                             "converged": mg.last_converged,
                             "true_resid": mg.last_true_resid_ratio}
                sol[name] = u.cpu().numpy()
            except _native.StallError as exc:
                res[name] = {"stalled": str(exc)}
            mg.close()
        out.update(res)

        def rel(x, y):
            return float(np.abs(x[used] - y[used]).max() / np.abs(y[used]).max())
        if "reference_fp64" in sol:
            for k in ("mixed_full", "mixed_active"):
                if k in sol:
                    out[k + "_rel_err_vs_fp64"] = rel(sol[k], sol["reference_fp64"])
        if args.oracle_full or args.workload == "cantilever":
            log("oracle: complete CPU solve")
            uo, t_o, it_o, conv_o = _oracle_solve(p, p.max_levels)
            out["cpu_oracle"] = {"time_s": t_o, "iters": it_o, "converged": conv_o,
                                 "cores": 1}
            for k in sol:
                out[k + "_rel_err_vs_oracle"] = rel(sol[k], uo)
    elif args.workload == "multi_rhs_part":
        p = W.multi_rhs_part()
        nrhs = p.force.shape[0]
        E, A, F = dev(p.E_stiffness), dev(p.active.astype(np.uint8)), dev(p.force)
        U = torch.empty_like(F)
        # one launch serves all 32 cases, so the coarse levels are no longer
        # latency-bound per case: the variable cycle pays here
        kw = dict(OPT)
        out["solver"] = OPT
        bat = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU",
                       max_batch=nrhs, **kw)
        run = lambda: (bat.set_density(E, active=A), bat.solve_batch_device(F, U, tol=p.tol))
        timed(run)
        t_b, _ = timed(run)
        out["batched_mixed"] = {"time_s": t_b, "nrhs": nrhs,
                                "iters_max": max(i["iters"] for i in bat.last_batch),
                                "all_converged": all(i["converged"] for i in bat.last_batch),
                                "true_resid_max": max(i["true_resid_ratio"] for i in bat.last_batch)}
        Ub = U.clone()
        one = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU", **kw)
        one.set_density(E, active=A)
        u = torch.empty_like(F[0])

        def seq():
            one.set_density(E, active=A)
            for q in range(nrhs):
                one.solve_device(F[q], u, tol=p.tol)
                U[q].copy_(u)
        timed(seq)
        t_s, _ = timed(seq)
        out["sequential_mixed"] = {"time_s": t_s}
        out["batched_speedup"] = t_s / t_b
        out["batched_vs_sequential_max_rel_diff"] = float(
            ((Ub - U).abs().amax(dim=1) / U.abs().amax(dim=1)).max().item())
        out["load_cases_per_s"] = nrhs / t_b
    elif args.workload == "uneven_parts":
        out["solver"] = OPT_PARTS
        out.update(uneven_parts_record())
    elif args.workload == "inside":
        # voxelisation inside test: a 81,920-triangle sphere in a 256^3 grid
        from core import inside
        from oracle import inside_ref
        v, fc = W.icosphere(6, (0.5, 0.5, 0.5), 0.45)
        dims, origin, vs = (256, 256, 256), (0.0, 0.0, 0.0), 1.0 / 256
        run = lambda: inside.inside_grid_device(dims, origin, vs, v, fc, "centers")
        timed(run)
        t, m = timed(run)
        npts = dims[0] * dims[1] * dims[2]
        out.update({"triangles": int(len(fc)), "points": npts, "gpu_time_s": t,
                    "points_per_s": npts / t, "inside_fraction": float(m.float().mean())})
        # CPU: the oracle restatement of the reference's numpy path on a sample
        rng = np.random.default_rng(0)
        idx = rng.choice(npts, 2000, replace=False)
        pts = inside_ref.voxel_centers(dims, origin, vs)[idx]
        t0 = time.perf_counter()
        ref = inside_ref.inside_mask(v, fc, pts, chunk=64)
        tc = time.perf_counter() - t0
        out["cpu_numpy_sample"] = {"points": 2000, "time_s": tc,
                                   "extrapolated_full_s": tc * npts / 2000}
        out["sample_matches"] = bool(np.array_equal(m.cpu().numpy()[idx].astype(bool), ref))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ref-iters", type=int, default=3000,
                    help="Jacobi-PCG fallback iterations of the reference comparator")
    ap.add_argument("--workload", default="two_material",
                    choices=["two_material", "cantilever", "voids_block", "multi_rhs_part",
                             "uneven_parts", "inside"],
                    help="two_material is the headline line; the others print their own "
                         "per-configuration measurement")
    ap.add_argument("--z-chunk", type=int, default=10,
                    help="--kernel-only: node planes per work-group of the timed kernel")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the compact records of the other four configurations")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip training / measuring the learned corrector")
    ap.add_argument("--oracle-full", action="store_true",
                    help="per-configuration modes: run the complete CPU oracle solve")
    ap.add_argument("--kernel-only", action="store_true",
                    help="run only the dominant kernel (and a plain copy of known size) for "
                         "hardware-counter passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solve runs whole-domain on one MI355X "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from core import _native
    from core import workloads as W
    from core.multigrid import MGSolver

    if args.workload != "two_material":
        print(json.dumps(other_workload(args)), flush=True)
        return

    log(f"generating the two-material workload, n={args.n}")
    p = W.two_material(n=args.n)
    nn = (p.nx + 1) * (p.ny + 1) * (p.nz + 1)
    ne = p.nelem
    E_dev = torch.tensor(p.E_stiffness, device="cuda")
    act_dev = torch.tensor(p.active.astype(np.uint8), device="cuda")
    f_dev = torch.tensor(p.force, device="cuda")
    u_dev = torch.empty_like(f_dev)
    tol = 1e-6

    if args.kernel_only:
        kernel_only(p, E_dev, act_dev, MGSolver, args.z_chunk)
        return

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    # --- comparator 1: the reference algorithm in FP64 on the GPU ----------
    log("reference FP64 MG-PCG (Jacobi omega 0.6, voids at e_min)")
    ref = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU")

    def run_ref():
        ref.set_density(E_dev, active=act_dev, e_min=1e-9)
        try:
            ref.solve_device(f_dev, u_dev, tol=tol)
            return {"stalled": False}
        except _native.StallError as exc:
            return {"stalled": True, "msg": str(exc).split(": ", 2)[-1]}

    t_mg, st = timed(run_ref)
    refinfo = {"mg_time_s": t_mg, "mg_iters": ref.last_iters, **st}
    if st["stalled"]:
        log(f"reference MG stalled ({st['msg']}); Jacobi-PCG fallback, "
            f"{args.ref_iters} iterations")
        t_jac, rc = timed(lambda: ref.solve_jacobi_device(
            E_dev, act_dev, f_dev, u_dev, tol=tol, max_cg=args.ref_iters))
        refinfo.update({"fallback_time_s": t_jac, "fallback_iters": ref.last_iters,
                        "fallback_ratio": ref.last_resid_ratio,
                        "fallback_converged": ref.last_converged,
                        "time_s": t_mg + t_jac})
        log(f"fallback: {ref.last_iters} its, ratio {ref.last_resid_ratio:.2e}, {t_jac:.2f} s")
    else:
        refinfo.update({"time_s": t_mg, "true_resid": ref.last_true_resid_ratio})
    ref.close()

    # --- comparator 1b: the same algorithm with a stable Jacobi weight -----
    # omega = 0.6 exceeds 2 / lambda_max(D^-1 A) (lambda_max ~ 5.25 for this
    # part), which is why comparator 1 stalls; 0.4 is the largest of the
    # tried weights that converges.  The reference's own system (voids at
    # E*e_min, full grid) and hierarchy (max_levels 6).
    log("plain FP64 MG-PCG (the reference algorithm with Jacobi omega 0.4)")
    plain = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU", omega=0.4)

    def run_plain():
        plain.set_density(E_dev, active=act_dev, e_min=1e-9)
        return plain.solve_device(f_dev, u_dev, tol=tol)

    try:
        timed(run_plain)
        t_plain, _ = timed(run_plain)
        plaininfo = {"time_s": t_plain, "iters": plain.last_iters,
                     "converged": plain.last_converged,
                     "true_resid": plain.last_true_resid_ratio, "omega": 0.4,
                     "system": "full grid, voids at E*e_min", "levels": len(plain.levels)}
    except _native.StallError as exc:
        plaininfo = {"stalled": True, "msg": str(exc).split(": ", 2)[-1], "omega": 0.4}
    log(f"plain FP64: {plaininfo}")
    plain.close()

    # --- comparator 2: FP64 MG-PCG with this build's preconditioner --------
    log("FP64 MG-PCG, Chebyshev / active system")
    kw64 = dict(OPT)
    kw64["precision"] = "fp64"
    mg64 = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU", **kw64)

    def run64():
        mg64.set_density(E_dev, active=act_dev)
        return mg64.solve_device(f_dev, u_dev, tol=tol)

    timed(run64)
    t64, _ = timed(run64)
    fp64 = {"time_s": t64, "iters": mg64.last_iters, "converged": mg64.last_converged,
            "true_resid": mg64.last_true_resid_ratio}
    log(f"FP64 Chebyshev: {fp64}")
    u64 = u_dev.clone()
    ms_mv64 = mg64.time_kernel("matvec", 0, "fp64", 50)
    mg64.close()

    # --- the measured solver ------------------------------------------------
    log(f"MI355X mixed-precision solver {OPT}")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    opt = MGSolver(p.nx, p.ny, p.nz, p.fixed_dofs, nu=p.nu, compute_mode="GPU", **OPT)

    def run_opt():
        opt.set_density(E_dev, active=act_dev)
        return opt.solve_device(f_dev, u_dev, tol=tol)

    for w in range(args.warmup):
        run_opt()
        log(f"warmup {w}: iters={opt.last_iters} true={opt.last_true_resid_ratio:.2e}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        run_opt()
        log(f"step {k}: iters={opt.last_iters} outer={opt.last_outer_iters} "
            f"true={opt.last_true_resid_ratio:.2e} device_ms={opt.last_solve_ms:.1f}")
    torch.cuda.synchronize()
    t_opt = (time.perf_counter() - t0) / args.steps
    dev_mem_gb = (free0 - torch.cuda.mem_get_info()[0]) / 1e9
    rel_err = float(((u_dev - u64).abs().max() / u64.abs().max()).item())
    mixed_iters, mixed_outer = opt.last_iters, opt.last_outer_iters
    tuned, tune_s = opt.tuned, opt.autotune_s
    plans = {str(k): list(v) for k, v in opt.block_plans().items()}
    mixed_true, mixed_conv = opt.last_true_resid_ratio, opt.last_converged

    # the same solve from host numpy arrays (the reference's calling form):
    # host->device copies of E, active, f and the copy of u back included
    def run_host():
        E_h = torch.from_numpy(p.E_stiffness).cuda()
        a_h = torch.from_numpy(p.active.astype(np.uint8)).cuda()
        f_h = torch.from_numpy(p.force).cuda()
        uh = torch.empty_like(f_h)
        opt.set_density(E_h, active=a_h)
        opt.solve_device(f_h, uh, tol=tol)
        return uh.cpu().numpy()
    t_pcie, _ = timed(run_host)

    # --- learned coarse-grid corrector, measured beside the plain solver ---
    learned = None
    if not args.no_learned:
        from core.learned import train_corrector
        lc = min(3, len(opt.levels) - 1)
        log(f"learned corrector: training at level {lc} {opt.levels[lc]['dims']}")
        opt.set_density(E_dev, active=act_dev)
        t_tr0 = time.perf_counter()
        net, rep = train_corrector(opt, f_dev, lc, n_samples=96, steps=300, batch=4, lr=3e-4)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t_tr0
        t_l, _ = timed(run_opt)
        learned = {"time_s": t_l, "iters": opt.last_iters, "corrector": opt.last_corrector,
                   "true_resid": opt.last_true_resid_ratio,
                   "rel_disp_diff_vs_fp64": float(((u_dev - u64).abs().max()
                                                   / u64.abs().max()).item()),
                   "training_s": t_train, "training": rep}
        log(f"learned: {learned['iters']} its ({learned['corrector']}), {t_l:.3f} s; "
            f"training {t_train:.1f} s")
        opt.detach_corrector()
        net.close()

    # --- the level-0 FP32 tile kernels of a PCG step, each at its tuned
    # z-chunk.  The roofline kernel is ROOF_KERNEL, the longest of them in the
    # solve (profiles/bench_two_material_256_kernel_trace.md)
    reps = 50
    nc = (p.nx // 2 + 1) * (p.ny // 2 + 1) * (p.nz // 2 + 1)
    level0 = {
        # kind: (time_kernel kind, algorithmic bytes per launch)
        "chebyshev2_zero_fp32": ("chebyshev2_zero", 24 * nn + 4 * nn + nn + 4 * ne),
        "jacobi_dot_fp32": ("jacobi_dot", 36 * nn + nn + 4 * ne),
        "jacobi_fp32": ("jacobi", 36 * nn + nn + 4 * ne),
        "residual_restrict_fp32": ("residual_restrict", 24 * nn + nn + 4 * ne + 12 * nc),
        "matvec_fp32": ("matvec", 24 * nn + nn + 4 * ne),
    }
    kernels = {}
    for name, (kind, nbytes) in level0.items():
        ms_k = opt.time_kernel(kind, 0, "mixed", reps)
        kernels[name] = {"ms": ms_k, "GBs": nbytes / (ms_k * 1e-3) / 1e9,
                         "algorithmic_bytes": nbytes}
    ms_jac32 = kernels[ROOF_KERNEL]["ms"]
    bytes_sweep = kernels[ROOF_KERNEL]["algorithmic_bytes"]
    bytes_mv64 = 48 * nn + nn + 8 * ne
    gbs = bytes_sweep / (ms_jac32 * 1e-3) / 1e9
    opt.close()

    copy_gbs = copy_bandwidth()
    log(f"plain device copy: {copy_gbs:.0f} GB/s")

    configs = None
    if not args.no_configs:
        configs = config_records()
        log(f"configs: {configs}")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(p, plaininfo.get("iters") or fp64["iters"])
        except Exception as exc:
            log(f"cpu baseline failed: {exc}")

    out = {
        "metric": "solve time @1e-6 rel. residual, 256^3 hex8: mixed+learned vs FP64 MG-PCG",
        "value": t_opt,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_opt * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 V-cycle + f64 refinement",
        "data": "synthetic",
        "config": {"workload": (f"two_material {p.nx}^3: steel/PP split by a smooth random "
                                f"level set, occupancy {p.meta['occupancy']:.3f}, x=0 fixed, "
                                f"roller z at x=max, 20 kN -z on the top nodes, nu 0.3"),
                   "solver": OPT, "tol": tol},
        "speedup_vs_fp64_same_preconditioner": t64 / t_opt,
        "reference_fp64_mgpcg": refinfo,
        "plain_fp64_mgpcg": plaininfo,
        "speedup_vs_plain_fp64": (plaininfo["time_s"] / t_opt
                                  if plaininfo.get("converged") else None),
        "fp64_mgpcg_chebyshev": fp64,
        "value_solver": "mixed FP32/FP64 Chebyshev MG-PCG (variable V-cycle) on the active-DOF "
                        "system; the learned corrector is measured beside it ('learned') and "
                        "is slower on this part",
        "mixed": {"iters": mixed_iters, "outer": mixed_outer,
                  "true_resid": mixed_true, "converged": mixed_conv,
                  "rel_disp_diff_vs_fp64": rel_err,
                  "autotuned_zchunk": {str(k): v for k, v in (tuned or {}).items()},
                  "autotune_s": tune_s,
                  "blocked_smoothing_plans": plans},
        "learned": learned,
        "pcie_inclusive_s": t_pcie,
        "device_mem_gb": dev_mem_gb,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic() if args.n == 256 else None,
                     "traffic_source": TRAFFIC_PROFILE,
                     "kernel": ROOF_KERNEL_NAME,
                     "kernel_ms": ms_jac32, "algorithmic_bytes": bytes_sweep,
                     "copy_GBs": copy_gbs, "frac_of_copy": gbs / copy_gbs},
        "kernels": dict(kernels, matvec_fp64={"ms": ms_mv64,
                                              "GBs": bytes_mv64 / (ms_mv64 * 1e-3) / 1e9}),
        "configs": configs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
