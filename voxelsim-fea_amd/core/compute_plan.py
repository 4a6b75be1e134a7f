"""Compute-plan selection (API of core/compute_plan.py:51-194).

Same names, thresholds and mode strings as the reference.  What changes is
where the work goes: every plan that reaches a usable GPU runs the HIP solver
(``gpu``; ``MULTI_GPU`` degrades to the single GPU exactly as the reference
does with one device, core/compute_plan.py:162).  AUTO picks the GPU for any
size when one is usable -- the reference's 50,000-DOF crossover was measured
for host-staged CuPy calls on consumer GPUs and does not apply to a solver
whose whole iteration stays on the device.  ``CPU`` / ``MULTI_CPU`` plans
(or AUTO without a GPU) have no solver in this build: ``VoxelFEA`` /
``MGSolver`` raise for them rather than compute on the host.
"""

import os

from . import backend

GPU_DOF = 50_000
MULTI_GPU_DOF = 50_000_000
MULTI_CPU_DOF = 500_000


class ComputePlan:
    __slots__ = ("kind", "xp", "n_workers", "label")

    def __init__(self, kind, xp, n_workers, label):
        self.kind = kind
        self.xp = xp
        self.n_workers = n_workers
        self.label = label

    @property
    def parallel(self):
        return self.kind in ("multi_gpu", "multi_cpu")

    @property
    def on_gpu(self):
        return self.kind in ("gpu", "multi_gpu")


def cpu_worker_count(requested=0):
    """core/compute_plan.py:80-90."""
    total = os.cpu_count() or 1
    if requested and requested > 0:
        return max(1, min(int(requested), total))
    return max(1, total - 1)


def configure_cpu_threads(n_threads, verbose=False):
    """Cap the BLAS pool (core/compute_plan.py:93-125); a no-op without
    threadpoolctl."""
    try:
        import threadpoolctl
        return threadpoolctl.threadpool_limits(limits=int(n_threads))
    except Exception as exc:
        if verbose:
            print(f"[VoxelSim FEA] BLAS thread cap unavailable ({exc})")
        return None


def choose(ndof, mode="AUTO", cpu_threads=0, verbose=False):
    """Resolve a ComputePlan for ``ndof`` DOFs (core/compute_plan.py:128-194)."""
    mode = (mode or "AUTO").upper()
    n_cpu = cpu_worker_count(cpu_threads)
    gpu_ok = backend.is_gpu_build() and backend.gpu_usable()
    n_gpus = backend.gpu_device_count() if gpu_ok else 0

    def _gpu():
        return ComputePlan("gpu", backend.get_xp(True), 1, "single GPU (MI355X HIP)")

    def _cpu():
        return ComputePlan("cpu", backend.get_xp(False), 1,
                           f"CPU ({n_cpu} threads; no solver in this build)")

    if mode in ("GPU", "MULTI_GPU", "AUTO"):
        plan = _gpu() if gpu_ok else _cpu()
    elif mode == "CPU":
        plan = _cpu()
    elif mode == "MULTI_CPU":
        plan = ComputePlan("multi_cpu", backend.get_xp(False), n_cpu,
                           f"multi-process CPU ({n_cpu} workers; no solver in this build)")
    else:
        raise ValueError(f"unknown compute_mode {mode!r}")
    if verbose:
        print(f"[VoxelSim FEA] compute plan: ndof={ndof} mode={mode} -> {plan.label} "
              f"(gpu_ok={gpu_ok}, n_gpus={n_gpus})")
    return plan
