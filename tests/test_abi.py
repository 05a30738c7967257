"""The C ABI: both libraries load on a machine without a GPU and export
every function the headers declare; without a device the product fails
loudly instead of falling back to host loops."""
import ctypes
import os
import re
import subprocess

import pytest

import qcs_amd
from qcs_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(%s\w+)\s*\(" % prefix, text)))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], check=True,
                         capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_hip_header_symbols_exported():
    names = declared("qcs_hip.h", "qcs_hip_")
    assert len(names) >= 30
    missing = set(names) - exported(_native.HIP_LIB_PATH)
    assert not missing, missing


def test_qcs_header_symbols_exported():
    names = declared("qcs.h", "qc_")
    # the reference's 30 functions are all declared
    for ref_fn in ["qc_create", "qc_destroy", "qc_h", "qc_x", "qc_y", "qc_z",
                   "qc_cnot", "qc_phase", "qc_rx", "qc_ry", "qc_rz", "qc_barrier",
                   "qc_reset", "qc_measure", "qc_measure_all", "qc_run",
                   "qc_run_shots", "qc_find_most_likely_state",
                   "qc_get_probability", "qc_print_state", "qc_print_circuit",
                   "qc_grover_search", "qc_quantum_fourier_transform",
                   "qc_bernstein_vazirani", "qc_ghz_state", "qc_get_num_qubits",
                   "qc_get_num_gates", "qc_optimize"]:
        assert ref_fn in names, ref_fn
    missing = set(names) - exported(_native.QCS_LIB_PATH)
    assert not missing, missing


def test_bindings_cover_headers():
    hip, qcs = qcs_amd.load()
    for n in declared("qcs_hip.h", "qcs_hip_"):
        assert n in _native._HIP_PROTOS, n
    for n in declared("qcs.h", "qc_"):
        assert n in _native._QCS_PROTOS, n


def test_struct_layouts_match_header():
    assert ctypes.sizeof(_native.HipGate) == 16 + 64
    assert ctypes.sizeof(_native.HipBlock) == 16 + 32


def test_no_device_fails_loudly():
    hip, qcs = qcs_amd.load()
    if hip.qcs_hip_device_available():
        pytest.skip("a device is present here")
    h = ctypes.c_void_p()
    rc = hip.qcs_hip_state_create(ctypes.byref(h), 4, _native.C128, 0)
    assert rc == -2 and not h.value  # QCS_HIP_ENODEV
    assert b"device" in hip.qcs_hip_last_error()
    assert not qcs.qc_create(4)
    with pytest.raises(RuntimeError):
        qcs_amd.Circuit(4)


def test_invalid_arguments_rejected_without_device():
    hip, _ = qcs_amd.load()
    h = ctypes.c_void_p()
    assert hip.qcs_hip_state_create(ctypes.byref(h), 0, 0, 0) == -1
    assert hip.qcs_hip_state_create(ctypes.byref(h), 35, 0, 0) == -1
    assert hip.qcs_hip_state_create(ctypes.byref(h), 4, 7, 0) == -1
    assert hip.qcs_hip_apply_1q(None, 0, None) == -1
