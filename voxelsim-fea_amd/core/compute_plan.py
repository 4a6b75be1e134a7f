"""Compute-plan selection (API of core/compute_plan.py:51-194).

Same names, thresholds, mode strings and branches as the reference.  What
changes is where the work goes: every plan that reaches a usable GPU runs the
HIP solver (``gpu``; ``MULTI_GPU`` resolves to one GPU, as the reference does
with one device, core/compute_plan.py:162).  AUTO picks the GPU for any size
when one is usable -- the reference's 50,000-DOF crossover was measured for
host-staged CuPy calls on consumer GPUs and chooses between two solvers, and
this build has only the device one.  ``CPU`` / ``MULTI_CPU`` plans (or AUTO
without a GPU) have no solver in this build: ``VoxelFEA`` / ``MGSolver``
raise for them rather than compute on the host.
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


_THREADPOOL_LIMITER = None   # the active cap, kept for the process lifetime


def configure_cpu_threads(n_threads, verbose=False):
    """Cap the BLAS pool (core/compute_plan.py:93-125), replacing an earlier
    cap; a no-op without threadpoolctl."""
    global _THREADPOOL_LIMITER
    try:
        import threadpoolctl
        if _THREADPOOL_LIMITER is not None:
            _THREADPOOL_LIMITER.unregister()
        _THREADPOOL_LIMITER = threadpoolctl.threadpool_limits(limits=int(n_threads))
        return _THREADPOOL_LIMITER
    except Exception as exc:
        if verbose:
            print(f"[VoxelSim FEA] BLAS thread cap unavailable ({exc})")
        return None


def choose(ndof, mode="AUTO", cpu_threads=0, verbose=False):
    """Resolve a ComputePlan for ``ndof`` DOFs (core/compute_plan.py:128-194).

    The branches are the reference's: ``CPU``; ``GPU`` (CPU when no GPU is
    usable); ``MULTI_GPU`` (the single GPU when there is one, since this
    build solves one whole domain per GPU); ``MULTI_CPU`` (CPU when only one
    worker is available); anything else is AUTO.  AUTO differs in one point:
    with a usable GPU it picks the GPU at every size, below ``GPU_DOF`` too,
    because this build has no host solver to route small problems to (the
    reference's 50,000-DOF crossover is a host-vs-device choice between two
    solvers).  Without a GPU, AUTO picks ``multi_cpu`` at ``MULTI_CPU_DOF``
    and more, else ``cpu``, as the reference does; those plans carry no
    solver here.  The main process's BLAS threads are capped as in the
    reference: one for the pooled plans, the worker count otherwise."""
    mode = (mode or "AUTO").upper()
    n_cpu = cpu_worker_count(cpu_threads)
    gpu_ok = backend.is_gpu_build() and backend.gpu_usable()
    n_gpus = backend.gpu_device_count() if gpu_ok else 0

    def _gpu(note=""):
        return ComputePlan("gpu", backend.get_xp(True), 1, "single GPU (MI355X HIP)" + note)

    def _cpu():
        return ComputePlan("cpu", backend.get_xp(False), 1,
                           f"CPU ({n_cpu} threads; no solver in this build)")

    def _multi_cpu():
        return ComputePlan("multi_cpu", backend.get_xp(False), n_cpu,
                           f"multi-process CPU ({n_cpu} workers; no solver in this build)")

    one_gpu = f"; {n_gpus} devices, whole domain on one" if n_gpus > 1 else ""
    if mode == "CPU":
        plan = _cpu()
    elif mode == "GPU":
        plan = _gpu() if gpu_ok else _cpu()
    elif mode == "MULTI_GPU":
        plan = _gpu(one_gpu) if gpu_ok else _cpu()
    elif mode == "MULTI_CPU":
        plan = _multi_cpu() if n_cpu > 1 else _cpu()
    else:   # AUTO (and any unrecognised string, as in the reference)
        if gpu_ok:
            plan = _gpu(one_gpu if ndof >= MULTI_GPU_DOF else "")
        elif ndof >= MULTI_CPU_DOF and n_cpu > 1:
            plan = _multi_cpu()
        else:
            plan = _cpu()
    configure_cpu_threads(1 if plan.kind in ("multi_cpu", "multi_gpu") else n_cpu, verbose)
    if verbose:
        print(f"[VoxelSim FEA] compute plan: ndof={ndof} mode={mode} -> {plan.label} "
              f"(gpu_ok={gpu_ok}, n_gpus={n_gpus}, cpu_workers={n_cpu})")
    return plan
