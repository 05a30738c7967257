"""Plain-form parity: HIP kernels vs the CPU oracle and the reference's golden
vectors, through the C ABI (qc_* over libqcs_hip.so).

Bar: gate application is bit-exact in complex128 (same operations, no FMA
contraction, the reference's association order); steps with a reduction
(normalise, measure, Grover's mean, diffusion) are within 1e-12 abs per
amplitude, because the device sums in a tree where the reference sums
sequentially.  Outcomes of measurement and sampling from the same srand()
seed are identical.
"""
import math

import numpy as np
import pytest

import helpers as HP
import qcs_amd
from oracle.oracle import OracleState
from qcs_amd import circuits as C

pytestmark = pytest.mark.gpu
RAND_MAX = 2147483647
ALL_KINDS = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE]


def cases(golden, prefix):
    return sorted({k.rsplit("/", 1)[0] for k in golden.files if k.startswith(prefix)})


def ops_of(golden, key):
    ops = C.Ops()
    for a in zip(golden[key + "/op"], golden[key + "/tq"], golden[key + "/cq"],
                 golden[key + "/par"]):
        ops.add(int(a[0]), int(a[1]), int(a[2]), float(a[3]))
    return ops


def gpu_run(n, ops, options=qcs_amd.PLAIN_FP64):
    c = qcs_amd.Circuit(n, options)
    c.apply(ops)
    return c


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 8, 10, 13, 16, 20])
def test_one_qubit_gates_bit_exact(qlib, n):
    print(f"[parity] 1q n={n}", flush=True)
    ops = C.random_circuit(n, 60, 31 * n, kinds=[C.H, C.X, C.Y, C.Z, C.P, C.RX,
                                                  C.RY, C.RZ])
    st = OracleState(n)
    st.run(ops)
    with gpu_run(n, ops) as c:
        assert np.array_equal(c.amplitudes(), st.vector())


@pytest.mark.parametrize("n", [2, 3, 4, 6, 9, 12, 15, 18])
def test_mixed_gates_bit_exact_textbook(qlib, n):
    print(f"[parity] mixed n={n}", flush=True)
    ops = C.random_circuit(n, 80, 97 * n + 1, kinds=ALL_KINDS + [C.SWAP])
    st = OracleState(n)
    st.run(ops)
    with gpu_run(n, ops) as c:
        assert np.array_equal(c.amplitudes(), st.vector())
        np.testing.assert_allclose(c.amplitudes(), HP.simulate(n, ops), atol=1e-12)


def test_every_target_and_control_pair(qlib):
    n = 9
    for t in range(n):
        for k in range(n):
            if k == t:
                continue
            ops = C.Ops().add(C.H, t).add(C.RY, k, -1, 0.7).add(C.RX, t, -1, 1.3)
            ops.add(C.CNOT, t, k).add(C.CPHASE, k, t, 0.4).add(C.RZ, t, -1, 2.2)
            st = OracleState(n)
            st.run(ops)
            with gpu_run(n, ops) as c:
                assert np.array_equal(c.amplitudes(), st.vector()), (t, k)


def test_compat_circuits_equal_reference_golden(qlib, golden):
    for key in cases(golden, "circ/"):
        n = int(golden[key + "/n"])
        with gpu_run(n, ops_of(golden, key), qcs_amd.REFERENCE_COMPAT) as c:
            assert np.array_equal(c.amplitudes(), golden[key + "/vec"]), key


def test_compat_scratch_buffer_matches_reference(qlib, golden):
    import ctypes
    hip, qcs = qlib
    for key in cases(golden, "circ/rand_n4"):
        n = int(golden[key + "/n"])
        c = gpu_run(n, ops_of(golden, key), qcs_amd.REFERENCE_COMPAT)
        state = c.device_state
        out = np.empty(2 << n, np.float64)
        rc = hip.qcs_hip_read_scratch(state, 0, 1 << n,
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert rc == 0
        assert np.array_equal(out.view(np.complex128), golden[key + "/scr"]), key
        c.close()


def test_compat_grover_matches_reference(qlib, golden):
    for key in cases(golden, "grover/"):
        n, m = int(golden[key + "/n"]), int(golden[key + "/m"])
        with qcs_amd.Circuit(n, qcs_amd.REFERENCE_COMPAT) as c:
            c.grover_search(m)
            np.testing.assert_allclose(c.amplitudes(), golden[key + "/vec"],
                                       rtol=0, atol=1e-12)
            assert c.get_num_gates() == int(golden[key + "/num_gates"])


def test_compat_qft_matches_reference(qlib, golden):
    for key in cases(golden, "qft/"):
        n, b = int(golden[key + "/n"]), int(golden[key + "/b"])
        with qcs_amd.Circuit(n, qcs_amd.REFERENCE_COMPAT) as c:
            for q in range(n):
                if (b >> q) & 1:
                    c.x(q)
            c.quantum_fourier_transform()
            assert np.array_equal(c.amplitudes(), golden[key + "/vec"]), key


@pytest.mark.parametrize("compat", [False, True])
def test_measure_sequences_match_reference(qlib, golden, cl, compat):
    opt = qcs_amd.REFERENCE_COMPAT if compat else qcs_amd.PLAIN_FP64
    for key in cases(golden, "measure/"):
        n = int(golden[key + "/n"])
        with gpu_run(n, ops_of(golden, key), opt) as c:
            cl.srand(int(golden[key + "/seed"]))
            outs = [c.measure(int(q)) for q in golden[key + "/order"]]
            assert outs == list(golden[key + "/outcomes"]), key
            np.testing.assert_allclose(c.amplitudes(), golden[key + "/vec"],
                                       rtol=0, atol=1e-12)


@pytest.mark.parametrize("compat", [False, True])
def test_shot_histograms_match_reference(qlib, golden, cl, compat):
    opt = qcs_amd.REFERENCE_COMPAT if compat else qcs_amd.PLAIN_FP64
    for key in cases(golden, "shots/"):
        n, shots = int(golden[key + "/n"]), int(golden[key + "/shots"])
        with gpu_run(n, ops_of(golden, key), opt) as c:
            cl.srand(int(golden[key + "/seed"]))
            hist = c.run_shots_hist(shots)
            assert np.array_equal(hist, golden[key + "/hist"]), key


def test_layered_compat_shots_match_reference(qlib, golden, cl):
    for key in cases(golden, "circ/layered"):
        n, shots = int(golden[key + "/n"]), int(golden[key + "/shots"])
        with gpu_run(n, ops_of(golden, key), qcs_amd.REFERENCE_COMPAT) as c:
            assert np.array_equal(c.amplitudes(), golden[key + "/vec"]), key
            cl.srand(int(golden[key + "/shots_seed"]))
            assert np.array_equal(c.run_shots_hist(shots), golden[key + "/hist"]), key


@pytest.mark.parametrize("n", [10, 14, 17, 20])
def test_sampling_matches_oracle_large(qlib, cl, n):
    print(f"[parity] sampling n={n}", flush=True)
    ops = C.layered(n, 2, 40 + n)
    st = OracleState(n)
    st.run(ops)
    with gpu_run(n, ops) as c:
        cl.srand(99)
        idx = c.sample_indices(3000)
        cl.srand(99)
        u = np.array([cl.rand() / RAND_MAX for _ in range(3000)])
        ref = st.sample(u)
        assert np.array_equal(idx, ref)


def test_sampling_edge_uniforms(qlib):
    import ctypes
    hip, _ = qlib
    n = 3
    with qcs_amd.Circuit(n) as c:
        c.set_amplitudes(np.full(8, 1 / math.sqrt(8), complex))
        state = c.device_state
        u = np.array([0.0, 0.124, 0.125, 0.5, 0.999, 1.0, 1.5])
        out = np.empty(len(u), np.int64)
        rc = hip.qcs_hip_sample(state, len(u), u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        assert rc == 0
        st = OracleState(n)
        st.set_vector(np.full(8, 1 / math.sqrt(8), complex))
        assert list(out) == list(st.sample(u))
        assert out[-1] == -1 and out[-2] == -1


@pytest.mark.parametrize("n", [4, 8, 12, 16, 20])
def test_normalize_measure_argmax_probability(qlib, cl, n):
    ops = C.random_circuit(n, 40, 5 * n, kinds=[C.H, C.RY, C.RZ, C.CNOT])
    st = OracleState(n)
    st.run(ops)
    with gpu_run(n, ops) as c:
        assert c.find_most_likely_state() == st.argmax()
        for i in (0, 1, (1 << n) - 1, 12345 % (1 << n)):
            assert c.get_probability(i) == st.probability(i)
        assert c.get_probability(-1) == 0.0 and c.get_probability(1 << n) == 0.0
        assert abs(c.norm_sq() - st.norm_sq()) < 1e-12
        cl.srand(5)
        outs = [c.measure(q) for q in range(0, n, 3)]
        cl.srand(5)
        ref = [st.measure(q, cl.rand() / RAND_MAX) for q in range(0, n, 3)]
        assert outs == ref
        np.testing.assert_allclose(c.amplitudes(), st.vector(), rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,m", [(2, 1), (6, 42), (12, 3000), (18, 77777)])
def test_grover_textbook(qlib, n, m):
    print(f"[parity] grover n={n}", flush=True)
    k = math.floor(math.pi / 4 * math.sqrt(2 ** n))
    with qcs_amd.Circuit(n) as c:
        c.grover_search(m)
        a = c.amplitudes()
        np.testing.assert_allclose(a, HP.grover_closed_form(n, m, k), rtol=0,
                                   atol=1e-11)
        assert c.get_num_gates() == n + 2 * k
        if n <= 12:
            st = OracleState(n)
            st.grover(m, k)
            np.testing.assert_allclose(a, st.vector(), rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [3, 8, 14])
def test_qft_textbook_closed_form(qlib, n):
    b = (0x9E3779B9 * n) & ((1 << n) - 1)
    with qcs_amd.Circuit(n) as c:
        for q in range(n):
            if (b >> q) & 1:
                c.x(q)
        c.quantum_fourier_transform()
        np.testing.assert_allclose(c.amplitudes(), HP.qft_closed_form(n, b),
                                   rtol=0, atol=1e-12)


def test_ghz_and_invalid_arguments(qlib, capfd):
    with qcs_amd.Circuit(5) as c:
        c.ghz_state()
        a = c.amplitudes()
        assert abs(a[0] - 1 / math.sqrt(2)) < 1e-15 and abs(a[31] - 1 / math.sqrt(2)) < 1e-15
        before = c.amplitudes()
        c.h(7)
        c.cnot(1, 1)
        assert np.array_equal(c.amplitudes(), before)
        assert c.get_num_gates() == 1 + 4 + 2  # history is logged anyway
    err = capfd.readouterr().err
    assert "Error: Invalid arguments for 1-qubit gate application." in err
    assert "Error: Invalid arguments for 2-qubit gate application." in err


def test_plain_fp32_close_to_fp64(qlib):
    n = 16
    ops = C.layered(n, 3, 8)
    with gpu_run(n, ops) as a, gpu_run(n, ops, qcs_amd.PLAIN_FP32) as b:
        omf, mx = a.compare_with(b)
        assert omf < 1e-9 and mx < 1e-5


@pytest.mark.parametrize("n,opt", [(4, qcs_amd.PLAIN_FP64), (9, qcs_amd.PLAIN_FP64),
                                   (14, qcs_amd.PLAIN_FP32), (20, qcs_amd.PLAIN_FP64),
                                   (11, qcs_amd.REFERENCE_COMPAT)])
def test_measure_all_on_device_equals_measure_loop(qlib, cl, n, opt):
    # qc_measure_all decides every outcome on the device (one host round
    # trip); it must give what qc_measure of each qubit in turn gives from
    # the same rand() stream, outcomes and final state
    ops = C.random_circuit(n, 8 * n, 555 + n, kinds=ALL_KINDS)
    with gpu_run(n, ops, opt) as a, gpu_run(n, ops, opt) as b:
        for seed in (1, 2, 3):
            cl.srand(seed)
            got = a.measure_all_list()
            cl.srand(seed)
            want = [b.measure(q) for q in range(n)]
            assert got == want, (seed, got, want)
            np.testing.assert_allclose(a.amplitudes(), b.amplitudes(), rtol=0, atol=1e-12)
            if opt != qcs_amd.REFERENCE_COMPAT:  # compat gates need not keep the norm
                assert a.get_probability(int("".join(map(str, got[::-1])), 2)) > 0.999
            # a fresh superposition for the next seed
            a.apply(ops)
            b.apply(ops)
