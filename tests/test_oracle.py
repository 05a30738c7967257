"""The CPU oracle, pinned before it is trusted (no GPU needed).

* compat mode == the compiled reference's golden vectors, bit for bit;
* textbook mode == two independent numpy evolutions (tensor contraction and
  explicit 2^n x 2^n Kronecker products) and the closed forms of QFT and
  Grover;
* planted mistakes (transposed / conjugated matrix, wrong qubit, dropped
  term) are shown to fail those same checks.
"""
import math

import numpy as np
import pytest

import helpers as HP
from oracle.oracle import OracleState, gate
from qcs_amd import circuits as C

RAND_MAX = 2147483647


def cases(golden, prefix):
    return sorted({k.rsplit("/", 1)[0] for k in golden.files if k.startswith(prefix)})


def ops_of(golden, key):
    ops = C.Ops()
    for a in zip(golden[key + "/op"], golden[key + "/tq"], golden[key + "/cq"],
                 golden[key + "/par"]):
        ops.add(int(a[0]), int(a[1]), int(a[2]), float(a[3]))
    return ops


@pytest.mark.parametrize("kind", [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ])
def test_gate_matrices_are_the_closed_forms(kind):
    for a in (0.0, 0.3, 1.0, math.pi, 2.5, 5.9):
        np.testing.assert_allclose(gate(kind, a), HP.m8(HP.textbook(kind, a)),
                                   rtol=0, atol=2e-16)


def test_golden_file_has_every_family(golden):
    for fam in ("circ/", "grover/", "qft/", "measure/", "shots/"):
        assert cases(golden, fam), fam


def test_compat_oracle_equals_reference_circuits(golden):
    for key in cases(golden, "circ/"):
        n = int(golden[key + "/n"])
        st = OracleState(n, compat=True)
        st.run(ops_of(golden, key))
        assert np.array_equal(st.vector(), golden[key + "/vec"]), key
        assert np.array_equal(st.scratch(), golden[key + "/scr"]), key


def test_compat_oracle_equals_reference_grover(golden):
    for key in cases(golden, "grover/"):
        n, m = int(golden[key + "/n"]), int(golden[key + "/m"])
        st = OracleState(n, compat=True)
        st.grover(m, st.L.oq_grover_iterations(n, 1))
        assert np.array_equal(st.vector(), golden[key + "/vec"]), key
        assert np.array_equal(st.scratch(), golden[key + "/scr"]), key
        # history: n H gates + 2 entries per iteration
        assert n + 2 * st.L.oq_grover_iterations(n, 1) == int(golden[key + "/num_gates"])


def test_compat_oracle_equals_reference_qft(golden):
    for key in cases(golden, "qft/"):
        n, b = int(golden[key + "/n"]), int(golden[key + "/b"])
        st = OracleState(n, compat=True)
        ops = C.Ops()
        for q in range(n):
            if (b >> q) & 1:
                ops.add(C.X, q)
        st.run(ops)
        st.qft()
        assert np.array_equal(st.vector(), golden[key + "/vec"]), key


@pytest.mark.parametrize("compat", [True, False])
def test_oracle_measure_equals_reference(golden, cl, compat):
    # 1-qubit-only circuits: both semantics must agree with the reference
    for key in cases(golden, "measure/"):
        n = int(golden[key + "/n"])
        st = OracleState(n, compat=compat)
        st.run(ops_of(golden, key))
        cl.srand(int(golden[key + "/seed"]))
        outs = [st.measure(int(q), cl.rand() / RAND_MAX) for q in golden[key + "/order"]]
        assert outs == list(golden[key + "/outcomes"]), key
        assert np.array_equal(st.vector(), golden[key + "/vec"]), key


@pytest.mark.parametrize("compat", [True, False])
def test_oracle_shots_equal_reference(golden, cl, compat):
    for key in cases(golden, "shots/"):
        n, shots = int(golden[key + "/n"]), int(golden[key + "/shots"])
        st = OracleState(n, compat=compat)
        st.run(ops_of(golden, key))
        assert np.array_equal(st.vector(), golden[key + "/vec"]), key
        cl.srand(int(golden[key + "/seed"]))
        u = np.array([cl.rand() / RAND_MAX for _ in range(shots)])
        idx = st.sample(u)
        hist = np.bincount(idx[idx >= 0], minlength=1 << n)
        assert np.array_equal(hist, golden[key + "/hist"]), key


def test_layered_shots_equal_reference_compat(golden, cl):
    for key in cases(golden, "circ/layered"):
        n, shots = int(golden[key + "/n"]), int(golden[key + "/shots"])
        st = OracleState(n, compat=True)
        st.run(ops_of(golden, key))
        cl.srand(int(golden[key + "/shots_seed"]))
        u = np.array([cl.rand() / RAND_MAX for _ in range(shots)])
        idx = st.sample(u)
        hist = np.bincount(idx[idx >= 0], minlength=1 << n)
        assert np.array_equal(hist, golden[key + "/hist"]), key


def test_numpy_references_agree():
    # the tensor-contraction evolution and the explicit Kronecker unitary are
    # written independently; they must agree before either pins anything
    for n in range(1, 6):
        ops = C.random_circuit(n, 30, 42 + n, kinds=list(range(11)) if n > 1 else None)
        psi = HP.simulate(n, ops)
        e0 = np.zeros(1 << n, complex)
        e0[0] = 1
        np.testing.assert_allclose(psi, HP.dense_unitary(n, ops) @ e0, atol=1e-13)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_textbook_oracle_vs_dense_unitary(n):
    for seed in range(3):
        kinds = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
        ops = C.random_circuit(n, 40, 77 * n + seed, kinds=kinds if n > 1 else kinds[:8])
        st = OracleState(n)
        st.run(ops)
        e0 = np.zeros(1 << n, complex)
        e0[0] = 1
        np.testing.assert_allclose(st.vector(), HP.dense_unitary(n, ops) @ e0,
                                   rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_textbook_oracle_qft_closed_form(n):
    for b in {0, 1, (1 << n) - 1, (0x5A5A5 & ((1 << n) - 1))}:
        st = OracleState(n)
        ops = C.Ops()
        for q in range(n):
            if (b >> q) & 1:
                ops.add(C.X, q)
        st.run(ops)
        st.qft()
        np.testing.assert_allclose(st.vector(), HP.qft_closed_form(n, b),
                                   rtol=0, atol=1e-13)


def test_textbook_oracle_qft_with_swaps_closed_form():
    n, b = 6, 37
    st = OracleState(n)
    ops = C.Ops()
    for q in range(n):
        if (b >> q) & 1:
            ops.add(C.X, q)
    for a in zip(*C.qft(n, swaps=True).arrays()):
        ops.add(*[int(a[0]), int(a[1]), int(a[2]), float(a[3])])
    st.run(ops)
    np.testing.assert_allclose(st.vector(), HP.qft_closed_form(n, b, swaps=True),
                               rtol=0, atol=1e-13)


@pytest.mark.parametrize("n,m", [(2, 1), (3, 5), (6, 42), (10, 700)])
def test_textbook_oracle_grover_closed_form(n, m):
    k = int(OracleState(n).L.oq_grover_iterations(n, 0))
    st = OracleState(n)
    st.grover(m, k)
    np.testing.assert_allclose(st.vector(), HP.grover_closed_form(n, m, k),
                               rtol=0, atol=1e-12)
    if n == 6:
        assert abs(st.probability(m) - 0.9966) < 1e-4  # the reference test's case


def test_grover_iteration_counts():
    st = OracleState(1)
    for n in range(1, 31):
        assert st.L.oq_grover_iterations(n, 0) == math.floor(math.pi / 4 * math.sqrt(2 ** n))
    assert st.L.oq_grover_iterations(33, 0) == 72792
    assert st.L.oq_grover_iterations(31, 1) == 0
    assert st.L.oq_grover_iterations(33, 1) == 1


def test_normalize_and_measure_basics():
    st = OracleState(3)
    st.set_vector(np.arange(8) * (1 + 1j))
    ns = st.normalize()
    assert ns == sum(2 * i * i for i in range(8))
    assert abs(st.norm_sq() - 1.0) < 1e-15
    st2 = OracleState(2)
    st2.set_vector(np.array([0, 0, 1, 0], complex))  # |10>: qubit 1 = 1
    assert st2.measure(0, 0.3) == 0 and st2.measure(1, 0.3) == 1


def test_sampling_rule_and_dropped_shots():
    st = OracleState(2)
    st.set_vector(np.array([0.5, 0.5, 0.5, 0.5], complex))  # p = 0.25 each
    u = np.array([0.0, 0.2499, 0.25, 0.7499, 0.75, 0.999999, 1.0])
    # first i with u < cumulative; u = 1.0 reaches no index: dropped (-1)
    assert list(st.sample(u)) == [0, 0, 1, 2, 3, 3, -1]


def test_argmax_first_maximum():
    st = OracleState(3)
    st.set_vector(np.array([0, 0.5, 0, 0.5, 0.5, 0, 0.5, 0], complex))
    assert st.argmax() == 1
    st.set_vector(np.zeros(8, complex))
    assert st.argmax() == 0


@pytest.mark.parametrize("mistake", ["transpose", "conjugate", "wrong_qubit",
                                     "drop_term", "cu_swapped_roles"])
def test_planted_mistakes_are_caught(mistake):
    # the numpy checks above must reject a plausibly wrong oracle
    n = 4
    ops = C.random_circuit(n, 30, 5, kinds=[C.H, C.RX, C.RY, C.RZ, C.P, C.CNOT, C.CPHASE])
    good = HP.dense_unitary(n, ops)[:, 0]
    psi = np.zeros(1 << n, complex)
    psi[0] = 1
    for op, t, c, p in zip(ops.op, ops.tq, ops.cq, ops.par):
        m = HP.textbook(op, p)
        t, c = int(t), int(c)
        if mistake == "transpose":
            m = m.T
        elif mistake == "conjugate":
            m = m.conj()
        elif mistake == "drop_term":
            m = np.array([[m[0, 0], 0], [m[1, 0], m[1, 1]]])
        if mistake == "wrong_qubit":
            t = (t + 1) % n
            if c == t:
                c = (c + 1) % n
        if op in (C.CNOT, C.CPHASE):
            if mistake == "cu_swapped_roles":
                c, t = t, c
            psi = HP.apply_cu(psi, n, c, t, m)
        else:
            psi = HP.apply_1q(psi, n, t, m)
    assert np.max(np.abs(psi - good)) > 1e-3


def test_openmp_restatement_matches_sequential():
    n = 12
    ops = C.random_circuit(n, 20, 3, kinds=[C.H, C.RX, C.RY, C.RZ])
    a, b = OracleState(n), OracleState(n)
    a.run(ops)
    for op, t, c, p in zip(ops.op, ops.tq, ops.cq, ops.par):
        m = gate(int(op), float(p))
        b.L.oq_apply_1q_omp(b.s, int(t), m.ctypes.data_as(__import__("ctypes").POINTER(
            __import__("ctypes").c_double)), 0)
    assert np.array_equal(a.vector(), b.vector())


@pytest.mark.parametrize("compat", [False, True])
def test_openmp_race_fixed_circuits_match_sequential(compat):
    # the race-fixed 1q and 2q OpenMP loops on 4 threads give the sequential
    # oracle's states bit for bit, in both semantics (textbook and the
    # reference's as-written controlled gate)
    n = 13
    ops = C.random_circuit(n, 60, 17, kinds=[C.H, C.RX, C.RY, C.RZ, C.P, C.CNOT, C.CPHASE])
    a, b = OracleState(n, compat=compat), OracleState(n, compat=compat)
    a.L.oq_set_threads(4)
    a.run(ops)
    b.run_omp(ops, racy=False)
    assert np.array_equal(a.vector(), b.vector())
    assert np.array_equal(a.scratch(), b.scratch())
    a.close()
    b.close()
