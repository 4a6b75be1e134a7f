"""CPU checks of bench.py's reporting contract: the one-GPU guard, the HBM
traffic figure read from the committed counter summary, and the CPU
baseline record (both of the reference's CPU modes) on a tiny workload."""

import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "one MI355X" in r.stderr
    assert r.stdout.strip() == ""


def test_traffic_figure_matches_the_counter_summary():
    sys.path.insert(0, ROOT)
    import bench
    # the summary of the kernel the bench line names as its roofline kernel
    text = open(os.path.join(ROOT, bench.TRAFFIC_PROFILE)).read()
    mb = float(re.search(r"\*\*([0-9.]+) MB\*\*", text).group(1))
    assert bench.measured_traffic() == pytest.approx(mb * 1e6, rel=1e-9)


def test_cpu_baseline_record():
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "voxelsim-fea_amd"))
    import bench
    from core import workloads as W
    p = W.two_material(n=16)
    r = bench.cpu_baseline(p, 10, single_iters=2, pool_iters=1)
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in r
    assert r["kind"] == "port" and r["unit"] == "s" and 1 <= r["cores"] <= 15
    assert r["value"] > 0
    single = r["single_process"]
    # measured iterations of the reference algorithm, not an operator count
    assert single["iterations"] == 2 and single["per_iteration_s"] > 0
    assert single["value"] == pytest.approx(r["hierarchy_s"] + single["set_density_s"]
                                            + 10 * single["per_iteration_s"])
    if r["pool"]:
        assert r["pool"]["iterations"] == 1
    assert r["value"] == (min(single["value"], r["pool"]["value"]) if r["pool"]
                          else single["value"])
    assert "omega 0.4" in r["sample"] and "complete PCG iterations" in r["sample"]
