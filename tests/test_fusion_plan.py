"""Host logic of the fused form: the greedy planner (libqcs.so,
qcs_fusion_plan) checked on the CPU -- every block acts on <= k qubits in
ascending order, every gate is placed once, and multiplying the blocks'
unitaries (numpy, FP64) reproduces the circuit."""
import ctypes

import numpy as np
import pytest

import helpers as HP
import qcs_amd
from qcs_amd import circuits as C
from qcs_amd._native import HipBlock, HipGate


def to_gates(ops):
    arr = (HipGate * max(len(ops), 1))()
    for i, (op, t, c, p) in enumerate(zip(ops.op, ops.tq, ops.cq, ops.par)):
        g = arr[i]
        if op == C.SWAP:
            g.kind, g.target, g.control = 2, int(t), int(c)
            continue
        m = HP.m8(HP.textbook(op, p))
        g.kind = 1 if op in (C.CNOT, C.CPHASE) else 0
        g.target, g.control = int(t), int(c) if op in (C.CNOT, C.CPHASE) else -1
        for j in range(8):
            g.m[j] = m[j]
    return arr


def plan(ops, n, k):
    _, qcs = qcs_amd.load()
    gates = to_gates(ops)
    pb = ctypes.POINTER(HipBlock)()
    pg = ctypes.POINTER(HipGate)()
    nb = ctypes.c_int()
    rc = qcs.qcs_fusion_plan(gates, len(ops), n, k, ctypes.byref(pb),
                             ctypes.byref(nb), ctypes.byref(pg))
    assert rc == 0
    blocks = [pb[i] for i in range(nb.value)]
    ordered = [pg[i] for i in range(len(ops))]
    return blocks, ordered


def gate_matrix_full(n, g):
    m = np.array([[g.m[0] + 1j * g.m[1], g.m[2] + 1j * g.m[3]],
                  [g.m[4] + 1j * g.m[5], g.m[6] + 1j * g.m[7]]])
    if g.kind == 2:
        return np.eye(1 << n, dtype=complex)[HP.apply_swap(np.arange(1 << n), n, g.target, g.control)]
    if g.kind == 1:
        return HP.embed(n, g.target, m, g.control)
    return HP.embed(n, g.target, m)


@pytest.mark.parametrize("n,k", [(3, 2), (4, 3), (5, 5), (6, 4), (7, 5), (7, 3)])
def test_plan_reproduces_circuit(n, k):
    kinds = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
    for seed in range(3):
        ops = C.random_circuit(n, 50, 1234 + 17 * n + seed, kinds=kinds)
        blocks, ordered = plan(ops, n, k)
        assert sum(b.num_gates for b in blocks) == len(ops)
        U = np.eye(1 << n, dtype=complex)
        pos = 0
        for b in blocks:
            assert 1 <= b.k <= k
            tg = [b.targets[i] for i in range(b.k)]
            assert tg == sorted(set(tg))
            assert b.first_gate == pos
            for g in ordered[b.first_gate:b.first_gate + b.num_gates]:
                qs = {g.target} if g.kind == 0 else {g.target, g.control}
                assert qs <= set(tg)
                U = gate_matrix_full(n, g) @ U
            pos += b.num_gates
        np.testing.assert_allclose(U, HP.dense_unitary(n, ops), atol=1e-12)


def test_plan_layered_30q_block_counts():
    ops = C.layered(30, 20, 2)
    assert len(ops) == 2410
    counts = {k: len(plan(ops, 30, k)[0]) for k in (3, 4, 5)}
    assert counts == {3: 290, 4: 144, 5: 99}


def test_plan_state_level_12q():
    n = 12
    ops = C.layered(n, 4, 9)
    blocks, ordered = plan(ops, n, 5)
    psi = np.zeros(1 << n, complex)
    psi[0] = 1
    for b in blocks:
        sub = C.Ops()
        for g in ordered[b.first_gate:b.first_gate + b.num_gates]:
            if g.kind == 1:
                op = C.CNOT if g.m[2] == 1.0 else C.CPHASE
            else:
                op = None
            if op is None:
                m = np.array([[g.m[0] + 1j * g.m[1], g.m[2] + 1j * g.m[3]],
                              [g.m[4] + 1j * g.m[5], g.m[6] + 1j * g.m[7]]])
                psi = HP.apply_1q(psi, n, g.target, m)
            else:
                psi = HP.apply_cu(psi, n, g.control, g.target, HP.textbook(C.CNOT))
    np.testing.assert_allclose(psi, HP.simulate(n, ops), atol=1e-12)


def pass_plan(ops, n, tb=13, low=6, k=4, max_blocks=22):
    _, qcs = qcs_amd.load()
    gates = to_gates(ops)
    pp = ctypes.POINTER(qcs_amd.HipPass)()
    pb = ctypes.POINTER(HipBlock)()
    pg = ctypes.POINTER(HipGate)()
    np_, nb = ctypes.c_int(), ctypes.c_int()
    rc = qcs.qcs_pass_plan(gates, len(ops), n, tb, low, k, max_blocks, ctypes.byref(pp),
                           ctypes.byref(np_), ctypes.byref(pb), ctypes.byref(nb),
                           ctypes.byref(pg))
    assert rc == 0
    return ([pp[i] for i in range(np_.value)], [pb[i] for i in range(nb.value)],
            [pg[i] for i in range(len(ops))])


@pytest.mark.parametrize("n,tb,low,k,mb", [(7, 5, 2, 2, 23), (8, 6, 3, 3, 2),
                                          (9, 7, 3, 4, 1), (9, 7, 3, 4, 23)])
def test_pass_plan_reproduces_circuit(n, tb, low, k, mb):
    kinds = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
    for seed in range(3):
        ops = C.random_circuit(n, 60, 99 + 13 * n + seed, kinds=kinds)
        passes, blocks, ordered = pass_plan(ops, n, tb, low, k, mb)
        assert max(p.num_blocks for p in passes) <= mb
        U = np.eye(1 << n, dtype=complex)
        nextb = 0
        for ps in passes:
            bits = [ps.bits[j] for j in range(ps.tile_bits)]
            assert ps.tile_bits == tb and bits == sorted(set(bits))
            assert bits[:low] == list(range(low))
            assert ps.first_block == nextb
            nextb += ps.num_blocks
            for b in blocks[ps.first_block:ps.first_block + ps.num_blocks]:
                tg = [b.targets[i] for i in range(b.k)]
                assert 1 <= b.k <= k and set(tg) <= set(bits)
                for g in ordered[b.first_gate:b.first_gate + b.num_gates]:
                    U = gate_matrix_full(n, g) @ U
        assert nextb == len(blocks)
        assert sum(b.num_gates for b in blocks) == len(ops)
        np.testing.assert_allclose(U, HP.dense_unitary(n, ops), atol=1e-12)


def test_pass_plan_layered_30q_counts():
    # the circuit is fused as a whole (144 blocks, as one pass per block at
    # k = 4) and the blocks dealt into passes, so no block is cut at a pass
    # boundary
    assert len(plan(C.layered(30, 20, 2), 30, 4)[0]) == 144
    passes, blocks, _ = pass_plan(C.layered(30, 20, 2), 30)
    assert (len(passes), len(blocks)) == (53, 144)
    # the tile programs' default: tiles hold bits 0..3 (128-byte runs)
    passes, blocks, _ = pass_plan(C.layered(30, 20, 2), 30, low=4)
    assert (len(passes), len(blocks)) == (31, 144)
    assert sum(b.k == 4 for b in blocks) == 143
    assert max(p.num_blocks for p in passes) <= 22


def test_circuit_gate_counts():
    assert C.gate_count_layered(20, 10) == len(C.layered(20, 10, 0)) == 810
    assert C.gate_count_layered(28, 20) == len(C.layered(28, 20, 0)) == 2248
    assert C.gate_count_layered(30, 20) == 2410
    assert C.gate_count_layered(34, 20) == 2734
    assert len(C.qft(28)) == 28 * 29 // 2
