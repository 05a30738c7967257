"""Host side of the batched form, on the CPU: the in-order pass planner
(qcs_gate_pass_plan) keeps the circuit's gate order and cuts it into passes
whose qubits fit one tile; deferred circuits record without a device; a
batch needs a device and says so."""
import ctypes

import numpy as np
import pytest

import qcs_amd
from qcs_amd import circuits as C
from qcs_amd._native import HipGate, HipPass

from test_fusion_plan import to_gates


def gate_plan(ops, n, tb, low=6):
    _, qcs = qcs_amd.load()
    gates = to_gates(ops)
    pp = ctypes.POINTER(HipPass)()
    np_ = ctypes.c_int()
    rc = qcs.qcs_gate_pass_plan(gates, len(ops), n, tb, low, ctypes.byref(pp),
                                ctypes.byref(np_))
    assert rc == 0
    return [pp[i] for i in range(np_.value)], gates


def qubits(g):
    return {g.target} if g.kind == 0 else {g.target, g.control}


@pytest.mark.parametrize("n,tb", [(3, 13), (13, 13), (14, 13), (16, 8), (20, 13),
                                  (20, 14), (9, 6), (12, 2)])
def test_gate_passes_keep_order_and_fit(n, tb):
    kinds = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
    for ops in (C.random_circuit(n, 120, 77 + n, kinds=kinds), C.layered(n, 3, 5)):
        passes, gates = gate_plan(ops, n, tb)
        eff = min(tb, n)
        nxt = 0
        for ps in passes:
            bits = [ps.bits[j] for j in range(ps.tile_bits)]
            assert ps.tile_bits == eff
            assert bits == sorted(set(bits)) and all(0 <= b < n for b in bits)
            low = min(6, eff) if n <= tb else min(6, max(eff - 2, 0))
            assert bits[:low] == list(range(low))
            assert ps.first_block == nxt  # consecutive gates, circuit order
            for g in range(ps.first_block, ps.first_block + ps.num_blocks):
                assert qubits(gates[g]) <= set(bits)
            nxt += ps.num_blocks
        assert nxt == len(ops)
        if n <= tb:
            assert len(passes) == 1
        # maximal runs: the first gate of each later pass did not fit
        for a, b in zip(passes, passes[1:]):
            used = set(range(min(6, max(eff - 2, 0))))
            for g in range(a.first_block, a.first_block + a.num_blocks):
                used |= qubits(gates[g])
            assert len(used | qubits(gates[b.first_block])) > eff


def test_gate_pass_plan_empty_circuit():
    passes, _ = gate_plan(C.Ops(), 15, 13)
    assert len(passes) == 1 and passes[0].num_blocks == 0
    assert [passes[0].bits[j] for j in range(13)] == list(range(13))


def test_gate_pass_plan_rejects_bad_qubits():
    _, qcs = qcs_amd.load()
    gates = (HipGate * 1)()
    gates[0].kind, gates[0].target = 0, 7
    pp, n = ctypes.POINTER(HipPass)(), ctypes.c_int()
    assert qcs.qcs_gate_pass_plan(gates, 1, 5, 13, 6, ctypes.byref(pp),
                                  ctypes.byref(n)) == -1


def test_deferred_circuit_records_without_device():
    c = qcs_amd.Circuit(12, qcs_amd.DEFERRED)
    c.apply(C.layered(12, 2, 1))
    assert c.get_num_gates() == len(C.layered(12, 2, 1))
    c.close()
    c = qcs_amd.Circuit(9, qcs_amd.DEFERRED | qcs_amd.FUSED_FP32)
    c.apply(C.qft(9, swaps=True))
    c.close()
    _, qcs = qcs_amd.load()
    # deferred circuits are complex128 plain or fused complex64, <= 22 qubits
    assert not qcs.qc_create_ex(23, qcs_amd.DEFERRED)
    assert not qcs.qc_create_ex(10, qcs_amd.DEFERRED | qcs_amd.PLAIN_FP32)
    assert not qcs.qc_create_ex(10, qcs_amd.DEFERRED | qcs_amd.REFERENCE_COMPAT)


def test_batch_needs_device():
    hip, qcs = qcs_amd.load()
    if hip.qcs_hip_device_available():
        pytest.skip("a device is present here")
    assert not qcs.qc_batch_create(0)
    h = ctypes.c_void_p()
    assert hip.qcs_hip_batch_create(ctypes.byref(h), 0) == -2
    assert not h.value


def hoist(gates, ng, n):
    _, qcs = qcs_amd.load()
    po = ctypes.POINTER(HipGate)()
    assert qcs.qcs_gate_hoist_exact(gates, ng, n, ctypes.byref(po)) == 0
    return [po[i] for i in range(ng)]


def apply_np(state, g, n):
    """One gate with per-gate rounding: out0 = m00*v0 + m01*v1 (numpy)."""
    m = np.array([[g.m[0] + 1j * g.m[1], g.m[2] + 1j * g.m[3]],
                  [g.m[4] + 1j * g.m[5], g.m[6] + 1j * g.m[7]]])
    idx = np.arange(1 << n)
    tb = 1 << g.target
    if g.kind == 2:
        cb = 1 << g.control
        sel = idx[((idx & tb) != 0) & ((idx & cb) == 0)]
        out = state.copy()
        out[sel], out[sel ^ tb ^ cb] = state[sel ^ tb ^ cb], state[sel]
        return out
    sel = idx[(idx & tb) == 0]
    if g.kind == 1:
        sel = sel[(sel & (1 << g.control)) != 0]
    v0, v1 = state[sel], state[sel | tb]
    out = state.copy()
    out[sel] = m[0, 0] * v0 + m[0, 1] * v1
    out[sel | tb] = m[1, 0] * v0 + m[1, 1] * v1
    return out


@pytest.mark.parametrize("n", [3, 6, 9])
def test_hoisting_exact_gates_keeps_amplitudes(n):
    _, qcs = qcs_amd.load()
    kinds = [C.H, C.X, C.Y, C.Z, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
    for seed in range(4):
        ops = C.random_circuit(n, 80, 31 * n + seed, kinds=kinds)
        gates = to_gates(ops)
        out = hoist(gates, len(ops), n)
        key = lambda g: (g.kind, g.target, g.control, tuple(g.m))  # noqa: E731
        assert sorted(map(key, out)) == sorted(key(gates[i]) for i in range(len(ops)))
        nonexact = [key(gates[i]) for i in range(len(ops))
                    if not qcs.qcs_gate_is_exact(ctypes.byref(gates[i]))]
        assert [key(g) for g in out if not qcs.qcs_gate_is_exact(ctypes.byref(g))] == nonexact
        rng = np.random.default_rng(seed)
        s0 = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        a, b = s0.copy(), s0.copy()
        for i in range(len(ops)):
            a = apply_np(a, gates[i], n)
        for g in out:
            b = apply_np(b, g, n)
        assert np.array_equal(a, b)


def test_hoisting_cuts_layered_passes():
    n, d = 20, 10
    ops = C.layered(n, d, 1)
    gates = to_gates(ops)
    out = hoist(gates, len(ops), n)
    arr = (HipGate * len(out))(*out)
    _, qcs = qcs_amd.load()
    pp, np_ = ctypes.POINTER(HipPass)(), ctypes.c_int()
    assert qcs.qcs_gate_pass_plan(arr, len(out), n, 13, 4, ctypes.byref(pp),
                                  ctypes.byref(np_)) == 0
    hoisted = np_.value
    passes, _ = gate_plan(ops, n, 13, 4)
    assert hoisted * 2 < len(passes)
    assert hoisted <= 2.2 * d


def test_exact_gate_classes():
    _, qcs = qcs_amd.load()
    ops = C.Ops()
    for k in (C.X, C.Y, C.Z, C.H, C.RZ):
        ops.add(k, 0, -1, 0.3)
    ops.add(C.CNOT, 1, 0).add(C.SWAP, 1, 0).add(C.CPHASE, 1, 0, 0.5)
    g = to_gates(ops)
    assert [qcs.qcs_gate_is_exact(ctypes.byref(g[i])) for i in range(len(ops))] == \
        [1, 1, 1, 0, 0, 1, 1, 0]
