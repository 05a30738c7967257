"""bench.py without a device: it refuses more than one GPU, every workload
stops with a message when no HIP device is present, and the reference-side
CPU baselines of the QFT and observe workloads run on small sizes (where
oracle/_ref was built)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True,
                          timeout=300, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


@pytest.mark.parametrize("workload", ["fused30", "batch", "qft", "observe", "config5"])
def test_workloads_need_a_device(workload):
    import qcs_amd
    hip, _ = qcs_amd.load()
    if hip.qcs_hip_device_available():
        pytest.skip("a device is present here")
    r = run("--workload", workload, "--no-cpu")
    assert r.returncode != 0
    assert "no HIP device" in (r.stderr + r.stdout)


def test_reference_baselines_small():
    sys.path.insert(0, ROOT)
    import bench
    q = bench.ref_qft_baseline(10)
    if q is None:
        pytest.skip("oracle/_ref not built")
    assert q["kind"] == "reference" and q["value"] > 0
    assert "55 gates" in q["sample"]          # 10 H + 45 controlled phases
    o = bench.ref_observe_baseline(10, 64, grover_n=8)
    assert set(o["ops"]) == {"argmax", "sample_64_shots", "measure_qubit_7", "grover_round_8q"}
    assert all(v["ms"] > 0 for v in o["ops"].values())
