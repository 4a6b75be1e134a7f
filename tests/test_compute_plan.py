"""Compute-plan routing (core/compute_plan.py:128-194) on the CPU: every mode
string, the thresholds, the MULTI_GPU / MULTI_CPU step-downs and the BLAS
thread cap, with the device probe replaced by fixed answers.

One documented difference from the reference: AUTO with a usable GPU picks
the GPU below GPU_DOF too (the reference would pick its host solver there;
this build has none), which ``test_auto_small_problem_stays_on_gpu`` states.
"""

import numpy as np
import pytest

from core import backend, compute_plan


@pytest.fixture
def machine(monkeypatch):
    """Set (gpu_ok, n_gpus, n_cpu) for choose(); returns the recorded BLAS caps."""
    caps = []

    def setup(gpu_ok, n_gpus, n_cpu):
        monkeypatch.setattr(backend, "is_gpu_build", lambda: True)
        monkeypatch.setattr(backend, "gpu_usable", lambda: gpu_ok)
        monkeypatch.setattr(backend, "gpu_device_count", lambda: n_gpus if gpu_ok else 0)
        monkeypatch.setattr(compute_plan.os, "cpu_count", lambda: n_cpu)
        monkeypatch.setattr(compute_plan, "configure_cpu_threads",
                            lambda n, verbose=False: caps.append(int(n)))
        return caps
    return setup


SMALL, MID, LARGE, HUGE = 1_000, 100_000, 1_000_000, 60_000_000


@pytest.mark.parametrize("ndof", [SMALL, MID, LARGE, HUGE])
def test_forced_cpu(machine, ndof):
    caps = machine(True, 1, 16)
    p = compute_plan.choose(ndof, "CPU")
    assert p.kind == "cpu" and p.n_workers == 1 and p.xp is np and not p.on_gpu
    assert caps == [15]          # all cores but one (core/compute_plan.py:80-90)


def test_forced_gpu(machine):
    machine(True, 1, 8)
    p = compute_plan.choose(SMALL, "gpu")          # case-insensitive
    assert p.kind == "gpu" and p.on_gpu and p.xp is backend.torch_xp
    machine(False, 0, 8)
    assert compute_plan.choose(MID, "GPU").kind == "cpu"   # no GPU: CPU


def test_multi_gpu_resolves_to_one_gpu(machine):
    caps = machine(True, 4, 8)
    p = compute_plan.choose(HUGE, "MULTI_GPU")
    # the reference builds a 4-device pool here; this build solves the whole
    # domain on one GPU, as the reference does when one device is visible
    assert p.kind == "gpu" and p.n_workers == 1 and "4 devices" in p.label
    assert caps == [7]
    machine(True, 1, 8)
    assert compute_plan.choose(HUGE, "MULTI_GPU").kind == "gpu"
    machine(False, 0, 8)
    assert compute_plan.choose(HUGE, "MULTI_GPU").kind == "cpu"


def test_multi_cpu(machine):
    caps = machine(True, 1, 8)
    p = compute_plan.choose(SMALL, "MULTI_CPU")
    assert p.kind == "multi_cpu" and p.n_workers == 7 and p.parallel and not p.on_gpu
    assert caps == [1]           # the pool's workers do the matvec
    caps.clear()
    machine(True, 1, 1)          # one core: nothing to pool
    p = compute_plan.choose(LARGE, "MULTI_CPU")
    assert p.kind == "cpu" and caps == [1]


def test_cpu_threads_request(machine):
    machine(False, 0, 16)
    assert compute_plan.choose(SMALL, "MULTI_CPU", cpu_threads=4).n_workers == 4
    assert compute_plan.choose(SMALL, "MULTI_CPU", cpu_threads=99).n_workers == 16


def test_auto_without_gpu_follows_thresholds(machine):
    machine(False, 0, 8)
    assert compute_plan.choose(SMALL).kind == "cpu"
    assert compute_plan.choose(compute_plan.MULTI_CPU_DOF - 1).kind == "cpu"
    assert compute_plan.choose(compute_plan.MULTI_CPU_DOF).kind == "multi_cpu"
    assert compute_plan.choose(HUGE).kind == "multi_cpu"
    machine(False, 0, 1)
    assert compute_plan.choose(HUGE).kind == "cpu"


def test_auto_with_gpu(machine):
    machine(True, 1, 8)
    assert compute_plan.choose(compute_plan.GPU_DOF).kind == "gpu"
    assert compute_plan.choose(LARGE).kind == "gpu"   # never multi_cpu with a GPU
    machine(True, 2, 8)
    p = compute_plan.choose(HUGE)
    assert p.kind == "gpu" and "2 devices" in p.label


def test_auto_small_problem_stays_on_gpu(machine):
    """The one deviation: the reference picks its host solver below GPU_DOF;
    this build has only the device solver, so AUTO keeps the GPU."""
    machine(True, 1, 8)
    assert compute_plan.choose(compute_plan.GPU_DOF - 1).kind == "gpu"


@pytest.mark.parametrize("mode", [None, "", "auto", "bogus"])
def test_unknown_or_empty_mode_is_auto(machine, mode):
    machine(False, 0, 8)
    assert compute_plan.choose(LARGE, mode).kind == "multi_cpu"
    machine(True, 1, 8)
    assert compute_plan.choose(SMALL, mode).kind == "gpu"


def test_thresholds_are_the_references():
    assert (compute_plan.GPU_DOF, compute_plan.MULTI_CPU_DOF, compute_plan.MULTI_GPU_DOF) == \
        (50_000, 500_000, 50_000_000)


def test_solver_refuses_a_cpu_plan(machine):
    """No host fallback: a plan without a GPU makes the solver raise."""
    from core.multigrid import _require_gpu
    machine(False, 0, 8)
    with pytest.raises(RuntimeError, match="no usable GPU"):
        _require_gpu(compute_plan.choose(MID, "AUTO"), "MGSolver")
