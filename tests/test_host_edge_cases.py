"""Edge cases of the C API that need no device: NULL handles, deferred
circuits (gates recorded, no state), option and size validation, and the
batch API without a device.  The reference returns early (or 0 / 0.0) on
these (src/qcs.c:245-248 for qc_measure, src/qcs.c:391-395 for
qc_get_probability); a NULL circuit, which the reference would dereference,
is a no-op here."""
import ctypes

import pytest

import qcs_amd
from qcs_amd import _native


@pytest.fixture(scope="module")
def q():
    return qcs_amd.load()[1]


def test_null_circuit_is_a_no_op(q):
    buf = (ctypes.c_int * 8)()
    lbuf = (ctypes.c_long * 8)()
    dbuf = (ctypes.c_double * 8)()
    for name in ("qc_h", "qc_x", "qc_y", "qc_z", "qc_reset"):
        getattr(q, name)(None, 0)
    for name in ("qc_phase", "qc_rx", "qc_ry", "qc_rz"):
        getattr(q, name)(None, 0, 0.5)
    q.qc_cnot(None, 0, 1)
    q.qc_cphase(None, 0, 1, 0.5)
    q.qc_swap(None, 0, 1)
    for name in ("qc_barrier", "qc_run", "qc_flush", "qc_optimize", "qc_print_circuit",
                 "qc_quantum_fourier_transform", "qc_ghz_state", "qc_destroy"):
        getattr(q, name)(None)
    q.qc_print_state(None, 0)
    q.qc_grover_search(None, 1)
    q.qc_bernstein_vazirani(None, 1)
    q.qc_measure_all(None, buf)
    q.qc_run_shots(None, 4, buf)
    q.qc_sample(None, 4, lbuf)
    q.qc_set_basis_state(None, 0)
    assert q.qc_measure(None, 0) == 0
    assert q.qc_get_probability(None, 0) == 0.0
    assert q.qc_get_probability_l(None, 0) == 0.0
    assert q.qc_find_most_likely_state(None) == 0
    assert q.qc_norm_sq(None) == 0.0
    assert q.qc_get_device_state(None) is None
    assert q.qc_get_amplitudes(None, 0, 1, dbuf) < 0
    assert q.qc_get_num_qubits(None) == 0
    assert q.qc_get_num_gates(None) == 0
    assert q.qc_set_amplitudes(None, 0, 1, dbuf) < 0


def test_create_validation(q, capfd):
    # n <= 0: as the reference (q_state_init's message, and qc_create keeps
    # the circuit without a state, src/q_state.c:38-41, src/qcs.c:54); gate
    # calls on it are rejected, observations give their defaults
    for n in (0, -3):
        c = q.qc_create_ex(n, _native.DEFERRED)
        assert c
        q.qc_h(c, 0)
        assert q.qc_get_probability(c, 0) == 0.0
        q.qc_destroy(c)
    assert "Error: Invalid arguments for 1-qubit gate application." in capfd.readouterr().err
    q.qc_destroy(q.qc_create_ex(0, _native.DEFERRED))
    assert not q.qc_create_ex(23, _native.DEFERRED)  # deferred circuits: 1..22 qubits
    assert not q.qc_create_ex(4, _native.FUSED_FP32 | _native.REFERENCE_COMPAT)
    assert not q.qc_create_ex(4, _native.FUSE_SPLIT_BF16)
    assert not q.qc_create_ex(4, _native.DEFERRED | _native.FUSED_FP32 | _native.FUSE_SPLIT_BF16)
    err = capfd.readouterr().err
    assert "Number of qubits must be positive" in err
    assert "QCS_REFERENCE_COMPAT applies to the plain forms only" in err
    assert "QCS_FUSE_SPLIT_BF16 applies to the fused form only" in err


def test_deferred_circuit_records_and_rejects(q, capfd):
    c = q.qc_create_ex(3, _native.DEFERRED)
    assert c
    try:
        q.qc_h(c, 0)
        q.qc_h(c, 3)          # out of range: message, logged like the reference
        q.qc_cnot(c, 1, 1)    # control == target
        q.qc_cnot(c, 0, -1)
        q.qc_rz(c, 2, 0.1)
        err = capfd.readouterr().err
        assert err.count("Error: Invalid arguments for 1-qubit gate application.") == 1
        assert err.count("Error: Invalid arguments for 2-qubit gate application.") == 2
        assert q.qc_get_num_qubits(c) == 3
        assert q.qc_get_options(c) == _native.DEFERRED
        # no device state: observations return their defaults
        assert q.qc_get_probability(c, 0) == 0.0
        assert q.qc_measure(c, 0) == 0
        assert q.qc_find_most_likely_state(c) == 0
        assert q.qc_get_device_state(c) is None
        # a deferred circuit starts from |0...0>: other basis states are refused
        q.qc_set_basis_state(c, 1)
        assert "starts from |0...0>" in capfd.readouterr().err
        n = q.qc_get_num_gates(c)
        q.qc_set_basis_state(c, 0)  # drops the recorded gates, keeps the history
        assert q.qc_get_num_gates(c) == n
    finally:
        q.qc_destroy(c)


def test_invalid_measure_draws_no_uniform(q):
    """The reference returns before rand() on bad arguments (src/qcs.c:245-248),
    so the uniform stream is unchanged."""
    libc = qcs_amd._native.libc()
    c = q.qc_create_ex(2, _native.DEFERRED)
    try:
        libc.srand(5)
        first = libc.rand()
        libc.srand(5)
        assert q.qc_measure(c, 7) == 0
        assert q.qc_measure(c, -1) == 0
        assert libc.rand() == first
    finally:
        q.qc_destroy(c)


def test_batch_without_device(q, capfd):
    """Creating a batch needs the device: here it fails with a message, and
    calls given no batch are rejected.  (Argument checks on a real batch run
    in test_gpu_edge_cases.py.)"""
    assert not q.qc_batch_create(0x100)
    assert "unknown batch options" in capfd.readouterr().err
    hip = qcs_amd.load()[0]
    if hip.qcs_hip_device_available():
        pytest.skip("a device is present here")
    assert not q.qc_batch_create(0)
    c = q.qc_create_ex(3, _native.DEFERRED)
    try:
        arr = (ctypes.c_void_p * 1)(c)
        idx = (ctypes.c_long * 4)()
        assert q.qc_batch_sample(None, arr, 1, 4, idx) < 0
        assert q.qc_batch_set_threads(None, 512) < 0
        assert q.qc_batch_autotune(None, arr, 1) < 0
    finally:
        q.qc_destroy(c)
    q.qc_batch_destroy(None)
