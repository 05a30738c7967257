"""Edge cases on the device: one-qubit circuits in every form, indices out
of range, invalid measurements and zero-shot calls (which draw no uniform,
as the reference returns before rand(): src/qcs.c:245-248, 581-582),
invalid gates on a fused circuit, amplitude ranges past the state, the
batch API's argument checks on a real batch, and the device gate entry
points on every bit and bit pair."""
import ctypes
import math

import numpy as np
import pytest

import qcs_amd
from qcs_amd import _native

pytestmark = pytest.mark.gpu

FORMS = [qcs_amd.PLAIN_FP64, qcs_amd.PLAIN_FP32, qcs_amd.FUSED_FP32,
         qcs_amd.REFERENCE_COMPAT]


@pytest.mark.parametrize("options", FORMS)
def test_one_qubit_circuits(qlib, cl, options):
    tol = 1e-12 if options in (qcs_amd.PLAIN_FP64, qcs_amd.REFERENCE_COMPAT) else 1e-6
    with qcs_amd.Circuit(1, options) as c:
        c.h(0)
        assert abs(c.get_probability(0) - 0.5) < tol and abs(c.get_probability(1) - 0.5) < tol
        c.h(0)
        c.x(0)
        assert c.find_most_likely_state() == 1
        c.ry(0, math.pi)          # |1> -> -|0>
        assert abs(c.get_probability(0) - 1.0) < tol
        c.quantum_fourier_transform()   # one qubit: H
        assert abs(c.get_probability(1) - 0.5) < tol
        cl.srand(17)
        hist = c.run_shots_hist(1000)
        assert hist.sum() == 1000 and 400 < hist[0] < 600   # test/test_qc_run_shots.c:11-12
        assert c.measure(0) in (0, 1)
        assert abs(c.norm_sq() - 1.0) < tol


def test_indices_out_of_range(qlib):
    with qcs_amd.Circuit(3) as c:
        c.x(0)
        assert c.get_probability(1) == 1.0
        for i in (-1, 8, 1 << 30):
            assert c.get_probability(i) == 0.0
        assert c.get_probability_l(8) == 0.0 and c.get_probability_l(-5) == 0.0
        out = (ctypes.c_double * 8)()
        assert c.get_amplitudes(6, 3, out) < 0      # past the end of the state
        assert c.get_amplitudes(-1, 1, out) < 0
        assert c.lib.qc_set_amplitudes(c.handle, 7, 2, out) < 0
        assert c.get_amplitudes(0, 0, out) == 0     # nothing to copy
        assert c.get_probability(1) == 1.0          # unchanged


def test_invalid_measure_and_zero_shots_draw_nothing(qlib, cl):
    with qcs_amd.Circuit(4) as c:
        c.h(0)
        cl.srand(3)
        first = cl.rand()
        cl.srand(3)
        assert c.measure(4) == 0 and c.measure(-1) == 0
        hist = np.full(16, 7, np.int32)
        c.run_shots(0, hist.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        assert (hist == 7).all()                   # not zeroed, as the reference
        idx = np.full(4, -9, np.int64)
        c.sample(0, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        assert (idx == -9).all()
        assert cl.rand() == first
        assert abs(c.get_probability(0) - 0.5) < 1e-15   # the state is untouched


def test_fused_circuit_rejects_invalid_gates(qlib, capfd):
    with qcs_amd.Circuit(6, qcs_amd.FUSED_FP32) as f, qcs_amd.Circuit(6) as p:
        for c in (f, p):
            c.h(0)
            c.cnot(0, 5)
            c.h(9)
            c.cnot(2, 2)
            c.swap(1, 1)
            c.rz(-1, 0.3)
        err = capfd.readouterr().err
        assert err.count("Error: Invalid arguments for 1-qubit gate application.") == 4
        assert err.count("Error: Invalid arguments for 2-qubit gate application.") == 2
        assert err.count("Error: Invalid arguments for swap.") == 2
        omf, mx = f.compare_with(p)
        assert omf < 1e-12 and mx < 1e-6
        assert f.get_num_gates() == p.get_num_gates() == 6


def test_bernstein_vazirani_needs_two_qubits(qlib, capfd):
    with qcs_amd.Circuit(1) as c:
        c.bernstein_vazirani(1)
        assert "requires at least 2 qubits" in capfd.readouterr().err
        assert c.get_probability(0) == 1.0


def test_batch_argument_checks(qlib):
    _, q = qlib
    b = q.qc_batch_create(0)
    assert b
    c = q.qc_create_ex(3, _native.DEFERRED)
    plain = q.qc_create(3)
    try:
        arr = (ctypes.c_void_p * 1)(c)
        idx = (ctypes.c_long * 4)()
        assert q.qc_batch_sample(b, arr, 0, 4, idx) == 0       # nothing to run
        assert q.qc_batch_sample(None, arr, 1, 4, idx) < 0
        assert q.qc_batch_sample(b, None, 1, 4, idx) < 0
        assert q.qc_batch_sample(b, arr, 1, -1, idx) < 0
        assert q.qc_batch_sample(b, (ctypes.c_void_p * 1)(plain), 1, 4, idx) < 0  # not deferred
        assert q.qc_batch_set_threads(b, 384) < 0              # 256 or 512 only
        assert q.qc_batch_set_plan(b, 0, 12, 14, 0) < 0        # complex128 tiles <= 13 bits
        dbuf = (ctypes.c_double * 2)()
        assert q.qc_batch_get_amplitudes(b, 0, 0, 1, dbuf) < 0  # no states kept
        q.qc_h(c, 0)
        assert q.qc_batch_sample(b, arr, 1, 4, idx) == 0
        assert all(i in (0, 1) for i in idx)
    finally:
        q.qc_destroy(c)
        q.qc_destroy(plain)
        q.qc_batch_destroy(b)


def test_argmax_ties_go_to_the_first_index(qlib):
    """qc_find_most_likely_state keeps the first maximum in index order
    (strict >, src/qcs.c:464-478), also when the equal maxima lie in
    different work-groups' ranges and the state is uniform."""
    n = 22
    N = 1 << n
    with qcs_amd.Circuit(n) as c:
        for q in range(n):
            c.h(q)
        assert c.find_most_likely_state() == 0
        for i, j in ((5, N - 3), (123457, 3000001), (N // 2 + 7, N - 1)):
            amps = np.zeros(N, np.complex128)
            amps[[i, j]] = 1 / math.sqrt(2)
            c.set_amplitudes(amps)
            assert c.find_most_likely_state() == i
            amps[j] = -1j / math.sqrt(2)          # equal |a|^2, other phase
            amps[N - 2] = 0.5                      # smaller
            c.set_amplitudes(amps)
            assert c.find_most_likely_state() == i


@pytest.mark.parametrize("options,tol", [(qcs_amd.PLAIN_FP64, 1e-14), (qcs_amd.PLAIN_FP32, 2e-6)])
def test_device_gates_on_every_bit_pair(qlib, options, tol):
    """qcs_hip_apply_1q / qcs_hip_apply_cu with a general matrix (and X) on
    every target, and every (control, target) pair, of an 8-qubit state:
    each launch path (a pair within one 128-byte line, one lane per
    amplitude; low control bits, whole-line sweeps; the half-state kernels)
    against the numpy statement of the reference's loops
    (src/q_gates.c:146-148, 277-292)."""
    import helpers as HP
    hip, _ = qlib
    n = 8
    rng = np.random.default_rng(5)
    u, _ = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))
    x = np.array([[0, 1], [1, 0]], complex)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    with qcs_amd.Circuit(n, options) as c:
        st = c.device_state
        for t in range(n):
            for m in (u, x):
                c.set_amplitudes(psi)
                mm = (ctypes.c_double * 8)(*HP.m8(m))
                assert hip.qcs_hip_apply_1q(st, t, mm) == 0
                got, want = c.amplitudes(), HP.apply_1q(psi, n, t, m)
                assert np.max(np.abs(got - want)) < tol, ("1q", t)
            for ctl in range(n):
                if ctl == t:
                    continue
                for m in (u, x):
                    c.set_amplitudes(psi)
                    mm = (ctypes.c_double * 8)(*HP.m8(m))
                    assert hip.qcs_hip_apply_cu(st, ctl, t, mm) == 0
                    got, want = c.amplitudes(), HP.apply_cu(psi, n, ctl, t, m)
                    assert np.max(np.abs(got - want)) < tol, ("cu", ctl, t)


def test_timing_counts_algorithmic_and_moved_bytes(qlib):
    """qcs_hip_timing_read counts the amplitudes a gate changes (read and
    written); qcs_hip_timing_moved the 128-byte lines they lie in.  On a
    complex128 state (8 amplitudes a line) a CNOT whose control is qubit 0
    or 2 changes half of every line and moves the whole state twice; one
    whose control is qubit 10 moves only its algorithmic bytes; a phase gate
    on qubit 1 or 9 likewise; a SWAP of qubits 1 and 8 touches every line."""
    hip, _ = qlib
    n = 16
    N, s = float(1 << n), 16.0
    x = (ctypes.c_double * 8)(0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0)
    ph = (ctypes.c_double * 8)(1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    with qcs_amd.Circuit(n) as c:
        st = c.device_state
        cases = [(qcs_amd.K_CU, lambda: hip.qcs_hip_apply_cu(st, 0, 5, x), N * s, 2 * N * s),
                 (qcs_amd.K_CU, lambda: hip.qcs_hip_apply_cu(st, 2, 5, x), N * s, 2 * N * s),
                 (qcs_amd.K_CU, lambda: hip.qcs_hip_apply_cu(st, 10, 5, x), N * s, N * s),
                 (qcs_amd.K_CU, lambda: hip.qcs_hip_apply_swap(st, 1, 8), N * s, 2 * N * s),
                 (qcs_amd.K_CU, lambda: hip.qcs_hip_apply_swap(st, 4, 8), N * s, N * s),
                 (qcs_amd.K_1Q, lambda: hip.qcs_hip_apply_1q(st, 1, ph), N * s, 2 * N * s),
                 (qcs_amd.K_1Q, lambda: hip.qcs_hip_apply_1q(st, 9, ph), N * s, N * s)]
        for cls, fn, algo, mv in cases:
            assert hip.qcs_hip_timing_enable(st, 1) == 0
            assert fn() == 0, hip.qcs_hip_last_error()
            ms, nl, b, f = (ctypes.c_double(), ctypes.c_long(), ctypes.c_double(),
                            ctypes.c_double())
            got = ctypes.c_double()
            assert hip.qcs_hip_timing_read(st, cls, ctypes.byref(ms), ctypes.byref(nl),
                                           ctypes.byref(b), ctypes.byref(f)) == 0
            assert hip.qcs_hip_timing_moved(st, cls, ctypes.byref(got)) == 0
            assert nl.value == 1 and b.value == algo and got.value == mv, (cls, b.value, got.value)
        assert hip.qcs_hip_timing_moved(st, qcs_amd.K_COUNT, ctypes.byref(got)) == -1
        assert hip.qcs_hip_timing_moved(st, qcs_amd.K_CU, None) == -1
        assert hip.qcs_hip_timing_enable(st, 0) == 0
