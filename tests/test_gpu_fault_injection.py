"""The device-error path that replaces the reference's silent CPU fallback
(src/gpu_complex.c:533-560): a failed backend call prints
"Error: <what> failed: <message>" on stderr and leaves a sticky error code in
the circuit (qc_last_error); nothing is computed on the host instead.

Failures are injected on the host (QCS_HIP_INJECT_FAULT, read once per
process, so each case runs in a child process): the N-th kernel-launch check
reports a failure although the kernel ran, or the N-th state allocation is
refused before hipMalloc.  The device never faults."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
sys.path.insert(0, %r)
import qcs_amd
_, q = qcs_amd.load()
c = q.qc_create(6)
if not c:
    print("CREATE_NULL", flush=True)
    sys.exit(0)
print("ERR_BEFORE", q.qc_last_error(c), flush=True)
for t in range(6):
    q.qc_h(c, t)
print("ERR_AFTER", q.qc_last_error(c), flush=True)
# the state stays usable: later gates and observations run on the device
q.qc_x(c, 0)
print("PROB", q.qc_get_probability(c, 1), flush=True)
print("ERR_END", q.qc_last_error(c), flush=True)
q.qc_destroy(c)
"""


def run_child(spec):
    env = dict(os.environ, QCS_HIP_INJECT_FAULT=spec)
    p = subprocess.run([sys.executable, "-c", CHILD % ROOT], env=env, capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stdout, p.stderr


def test_injected_launch_failure_is_reported_and_sticky(qlib):
    # launch checks: 1 = |0> fill of qc_create, 2.. = the H gates
    out, err = run_child("launch:4")
    assert "ERR_BEFORE 0" in out
    assert "ERR_AFTER -4" in out  # QCS_HIP_EDEVICE, kept after later good calls
    assert "ERR_END -4" in out
    assert "Error: " in err and "injected failure" in err
    assert err.count("injected failure") == 1
    # H on every qubit then X(0): |a_1|^2 = 1/64, computed on the device
    prob = float(out.split("PROB ")[1].split()[0])
    assert abs(prob - 1.0 / 64) < 1e-15


def test_injected_allocation_failure_returns_null(qlib):
    out, err = run_child("alloc:1")
    assert "CREATE_NULL" in out
    assert "injected failure" in err


def test_no_injection_no_error(qlib):
    out, err = run_child("none:0")
    assert "ERR_AFTER 0" in out and "ERR_END 0" in out
    assert "Error" not in err
