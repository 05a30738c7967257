"""Host side of the batched form, on the CPU: the in-order pass planner
(qcs_gate_pass_plan) keeps the circuit's gate order and cuts it into passes
whose qubits fit one tile; deferred circuits record without a device; a
batch needs a device and says so."""
import ctypes

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
