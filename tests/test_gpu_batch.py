"""Batched circuits on the GPU (qc_batch_*, one persistent launch).

complex128 circuits must equal a circuit of their own (qc_create + the same
gates + qc_sample) bit for bit: every amplitude and every sampled index from
the same rand() stream.  Fused complex64 circuits must stay within 1 - F <=
1e-6 of the complex128 result and sample the same indices wherever the FP32
error cannot move a CDF boundary across a uniform.  Sizes cover states that
stay in LDS (n <= tile bits), states streamed through LDS tiles, the
boundary between them, one-qubit and gate-free circuits."""
import numpy as np
import pytest

import qcs_amd
from qcs_amd import circuits as C

pytestmark = pytest.mark.gpu

ALL = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]


def make_ops(n, seed, count=None):
    if count is None:
        return C.layered(n, 3, seed) if seed % 2 else C.random_circuit(
            n, 12 * n, seed, kinds=ALL)
    return C.random_circuit(n, count, seed, kinds=ALL)


def batch_run(specs, options, shots, cl, seed, batch_opts=qcs_amd.BATCH_KEEP_STATES,
              plans=(), threads=None):
    circs = []
    for n, ops in specs:
        c = qcs_amd.Circuit(n, qcs_amd.DEFERRED | options)
        c.apply(ops)
        circs.append(c)
    b = qcs_amd.Batch(batch_opts)
    for p in plans:
        b.set_plan(*p)
    if threads:
        b.set_threads(threads)
    cl.srand(seed)
    idx = b.sample(circs, shots)
    amps = [b.amplitudes(i, n) for i, (n, _) in enumerate(specs)]
    for c in circs:
        c.close()
    return b, amps, idx


SIZES = [1, 2, 3, 5, 9, 12, 13, 14, 15, 16]


def test_c128_batch_bit_identical(qlib, cl):
    specs = [(n, make_ops(n, 100 + n)) for n in SIZES]
    b, amps, idx = batch_run(specs, qcs_amd.PLAIN_FP64, 200, cl, 7)
    # the single runs draw the uniforms circuit after circuit, like the batch
    cl.srand(7)
    for i, (n, ops) in enumerate(specs):
        c = qcs_amd.Circuit(n)
        c.apply(ops)
        ref_i = c.sample_indices(200)
        ref_a = c.amplitudes()
        c.close()
        assert np.array_equal(amps[i], ref_a), f"n={n}: amplitudes differ"
        assert np.array_equal(idx[i], ref_i), f"n={n}: samples differ"
    b.close()


@pytest.mark.parametrize("tb", [6, 8, 11])
def test_c128_streamed_tiles_match_resident(qlib, cl, tb):
    # states larger than the tile go through HBM pass by pass
    specs = [(n, make_ops(n, 300 + n)) for n in (tb + 1, tb + 3, 14)]
    plans = [(0, n, tb, 0) for n, _ in specs]
    _, amps, idx = batch_run(specs, qcs_amd.PLAIN_FP64, 64, cl, 3, plans=plans)
    _, amps0, idx0 = batch_run(specs, qcs_amd.PLAIN_FP64, 64, cl, 3)
    for i in range(len(specs)):
        assert np.array_equal(amps[i], amps0[i])
        assert np.array_equal(idx[i], idx0[i])


def fidelity(a, b):
    return 1.0 - abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def test_c64_fused_batch_within_tolerance(qlib, cl):
    specs = [(n, make_ops(n, 500 + n)) for n in SIZES + [17]]
    _, amps, idx = batch_run(specs, qcs_amd.FUSED_FP32, 256, cl, 11)
    _, amps64, idx64 = batch_run(specs, qcs_amd.PLAIN_FP64, 256, cl, 11)
    for i, (n, _) in enumerate(specs):
        omf = fidelity(amps[i], amps64[i])
        assert omf <= 1e-6, (n, omf)
        assert np.max(np.abs(amps[i] - amps64[i])) <= 1e-5, n
        assert abs(np.vdot(amps[i], amps[i]).real - 1.0) <= 1e-6, n
        assert np.mean(idx[i] == idx64[i]) >= 0.99, n


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_c64_fusion_widths_and_streaming(qlib, cl, k):
    specs = [(n, make_ops(n, 700 + n + k)) for n in (6, 11, 15)]
    plans = [(1, n, 0, k) for n, _ in specs] + [(1, 15, 9, k)]
    _, amps, _ = batch_run(specs, qcs_amd.FUSED_FP32, 16, cl, 5, plans=plans)
    _, amps64, _ = batch_run(specs, qcs_amd.PLAIN_FP64, 16, cl, 5)
    for i, (n, _) in enumerate(specs):
        assert fidelity(amps[i], amps64[i]) <= 1e-6, (n, k)


def test_mixed_batch_matches_separate_batches(qlib, cl):
    specs = [(n, make_ops(n, 900 + n)) for n in (4, 10, 14, 16)]
    circs = []
    for j, (n, ops) in enumerate(specs * 2):
        opt = qcs_amd.FUSED_FP32 if j >= len(specs) else qcs_amd.PLAIN_FP64
        c = qcs_amd.Circuit(n, qcs_amd.DEFERRED | opt)
        c.apply(ops)
        circs.append(c)
    b = qcs_amd.Batch(qcs_amd.BATCH_KEEP_STATES)
    cl.srand(21)
    idx = b.sample(circs, 32)
    _, a128, i128 = batch_run(specs, qcs_amd.PLAIN_FP64, 32, cl, 21)
    for i, (n, _) in enumerate(specs):
        assert np.array_equal(b.amplitudes(i, n), a128[i])
        assert np.array_equal(idx[i], i128[i])
        assert fidelity(b.amplitudes(len(specs) + i, n), a128[i]) <= 1e-6
    for c in circs:
        c.close()
    b.close()


def test_threads_and_graph_replay(qlib, cl):
    specs = [(n, make_ops(n, 40 + n)) for n in (3, 8, 13, 15)]
    _, a512, i512 = batch_run(specs, qcs_amd.PLAIN_FP64, 48, cl, 9)
    _, a256, i256 = batch_run(specs, qcs_amd.PLAIN_FP64, 48, cl, 9, threads=256)
    for i in range(len(specs)):
        assert np.array_equal(a512[i], a256[i]) and np.array_equal(i512[i], i256[i])
    # graph: first run captures, later runs with the same shape replay; new
    # angles (same structure) must give the new results
    b = qcs_amd.Batch(qcs_amd.BATCH_GRAPH | qcs_amd.BATCH_KEEP_STATES)
    for rep, seed in enumerate((9, 9, 33)):
        circs = []
        for n, _ in specs:
            c = qcs_amd.Circuit(n, qcs_amd.DEFERRED)
            c.apply(C.layered(n, 2, seed + n))
            circs.append(c)
        cl.srand(seed)
        idx = b.sample(circs, 48)
        _, _, replayed = b.last_timing()
        assert replayed == (rep > 0)
        ref_specs = [(n, C.layered(n, 2, seed + n)) for n, _ in specs]
        _, ra, ri = batch_run(ref_specs, qcs_amd.PLAIN_FP64, 48, cl, seed)
        for i, (n, _) in enumerate(specs):
            assert np.array_equal(b.amplitudes(i, n), ra[i])
            assert np.array_equal(idx[i], ri[i])
        for c in circs:
            c.close()
    b.close()


def test_edges(qlib, cl):
    # gate-free circuits sample |0...0>; zero shots; run_shots histograms
    specs = [(n, C.Ops()) for n in (1, 7, 15)]
    _, amps, idx = batch_run(specs, qcs_amd.PLAIN_FP64, 10, cl, 1)
    for i, (n, _) in enumerate(specs):
        assert amps[i][0] == 1 and np.count_nonzero(amps[i]) == 1
        assert np.all(idx[i] == 0)
    _, _, idx = batch_run(specs, qcs_amd.FUSED_FP32, 0, cl, 1)
    assert idx.shape == (3, 0)
    circs = []
    for n in (2, 9):
        c = qcs_amd.Circuit(n, qcs_amd.DEFERRED)
        c.h(0)
        circs.append(c)
    b = qcs_amd.Batch()
    cl.srand(4)
    hists = b.run_shots(circs, 1000)
    for h, c in zip(hists, circs):
        assert h.sum() == 1000 and 400 <= h[0] <= 600 and h[0] + h[1] == 1000
        c.close()
    b.close()


def test_rejects_non_deferred(qlib, cl):
    c = qcs_amd.Circuit(4)
    b = qcs_amd.Batch()
    with pytest.raises(RuntimeError):
        b.sample([c], 4)
    c.close()
    b.close()


def test_optimize_cancels_recorded_pairs(qlib, cl):
    # X X, Y Y, Z Z, CNOT CNOT pairs vanish from the recorded gates; the
    # batch still equals the eager plain form, which computes every gate
    n = 9
    ops = C.Ops()
    for q in range(n):
        ops.add(C.H, q).add(C.X, q).add(C.X, q).add(C.RY, q, -1, 0.3 + q)
    ops.add(C.CNOT, 1, 0).add(C.CNOT, 1, 0).add(C.Z, 4).add(C.Z, 4).add(C.Y, 2).add(C.Y, 2)
    d = qcs_amd.Circuit(n, qcs_amd.DEFERRED)
    d.apply(ops)
    d.optimize()
    assert d.get_num_gates() == len(ops.op) - 2 * (n + 3)
    b = qcs_amd.Batch(qcs_amd.BATCH_KEEP_STATES)
    cl.srand(8)
    idx = b.sample([d], 32)
    with qcs_amd.Circuit(n) as e:
        e.apply(ops)
        cl.srand(8)
        assert np.array_equal(b.amplitudes(0, n), e.amplitudes())
        assert np.array_equal(idx[0], e.sample_indices(32))
    d.close()
    b.close()
    # the fused form also drops H H pairs
    with qcs_amd.Circuit(n, qcs_amd.FUSED_FP32) as f, qcs_amd.Circuit(n) as e:
        f.apply(ops)
        f.h(3)
        f.h(3)
        f.optimize()
        e.apply(ops)
        omf, _ = f.compare_with(e)
        assert omf <= 1e-6


@pytest.mark.parametrize("graph", [0, qcs_amd.BATCH_GRAPH])
def test_rerun_reuses_the_last_plan(qlib, cl, capfd, monkeypatch, graph):
    # sampling the same, unchanged circuits again reuses the device plan and
    # only draws new uniforms; every run must equal a batch that plans from
    # scratch each time (QCS_BATCH_REPLAN) on the same rand() stream
    monkeypatch.setenv("QCS_BATCH_PROFILE", "1")
    circs = []
    for n in (2, 9, 13, 15, 17):
        for opt in (qcs_amd.PLAIN_FP64, qcs_amd.FUSED_FP32):
            c = qcs_amd.Circuit(n, qcs_amd.DEFERRED | opt)
            c.apply(make_ops(n, 500 + n))
            circs.append(c)
    opts = qcs_amd.BATCH_KEEP_STATES | graph
    b = qcs_amd.Batch(opts)
    r = qcs_amd.Batch(opts | qcs_amd.BATCH_REPLAN)

    def both(seed, shots, reruns):
        cl.srand(seed)
        got = [b.sample(circs, shots) for _ in range(3)]
        assert capfd.readouterr().err.count("rerun of the last plan") == reruns
        cl.srand(seed)
        want = [r.sample(circs, shots) for _ in range(3)]
        assert "rerun" not in capfd.readouterr().err
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
        for i, c in enumerate(circs):
            assert np.array_equal(b.amplitudes(i, c.n), r.amplitudes(i, c.n))
        return got

    got = both(21, 40, 2)
    assert not np.array_equal(got[0], got[1])  # new uniforms each run
    # re-recorded circuit: planned again, then reused
    circs[3].set_basis_state(0)
    circs[3].apply(make_ops(9, 77))
    both(5, 40, 2)
    # another shot count: planned again (the first of the three)
    both(6, 24, 2)
    # a changed work-group size invalidates the plan
    b.set_threads(256)
    r.set_threads(256)
    both(7, 24, 2)
    for c in circs:
        c.close()
    b.close()
    r.close()


def test_shared_large_circuits_equal_standalone(qlib, cl, monkeypatch):
    # streamed complex128 circuits of 18-20 qubits at the default tile bits,
    # beside small ones: the large ones are shared by several work-groups
    # (their tiles dealt out per pass, the parts meeting between passes), and
    # must give what one work-group gives (QCS_BATCH_MAX_PARTS=1) and what a
    # circuit of its own gives, amplitudes and samples bit for bit
    specs = [(20, C.layered(20, 5, 61)), (19, C.layered(19, 7, 62)),
             (18, make_ops(18, 63))] + [(n, make_ops(n, 70 + n)) for n in (3, 9, 12, 13, 14)]
    monkeypatch.setenv("QCS_BATCH_MAX_PARTS", "1")
    _, a1, i1 = batch_run(specs, qcs_amd.PLAIN_FP64, 300, cl, 17)
    monkeypatch.delenv("QCS_BATCH_MAX_PARTS")
    b, a2, i2 = batch_run(specs, qcs_amd.PLAIN_FP64, 300, cl, 17,
                          batch_opts=qcs_amd.BATCH_KEEP_STATES | qcs_amd.BATCH_GRAPH)
    cl.srand(17)
    for i, (n, ops) in enumerate(specs):
        with qcs_amd.Circuit(n) as c:
            c.apply(ops)
            own_idx = c.sample_indices(300)
            own = c.amplitudes()
        assert np.array_equal(a2[i], own) and np.array_equal(i2[i], own_idx), n
        assert np.array_equal(a1[i], own) and np.array_equal(i1[i], own_idx), n
    b.close()


def euler_and_permutation_ops(n, seed):
    """Rotation triples RZ RY RZ on one qubit (one op on the device), CNOT
    ladders and SWAPs that end register groups or sit inside them (moved
    past the ops they share no qubit with, or stored through permuted
    slots), and triples broken by a gate on their qubit."""
    rng = np.random.default_rng(seed)
    ops = C.Ops()
    for layer in range(4):
        for q in range(n):
            ops.add(C.RZ, q, p=float(rng.uniform(-3, 3)))
            ops.add(C.RY, q, p=float(rng.uniform(-3, 3)))
            if (q + layer) % 5 == 0:
                ops.add(C.CNOT, q, (q + 1) % n)  # inside the triple: no fusion
            ops.add(C.RZ, q, p=float(rng.uniform(-3, 3)))
        for q in range(n - 1):
            ops.add(C.CNOT, q + 1, q)
            if (q + layer) % 3 == 0:
                ops.add(C.SWAP, q, (q + 2) % n)
        ops.add(C.H, layer % n)
    return ops


@pytest.mark.parametrize("n", [5, 13, 14, 16])
def test_c128_fused_rotations_and_folded_permutations(qlib, cl, n):
    specs = [(n, euler_and_permutation_ops(n, 40 + n))]
    b, amps, idx = batch_run(specs, qcs_amd.PLAIN_FP64, 256, cl, 11)
    cl.srand(11)
    c = qcs_amd.Circuit(n)
    c.apply(specs[0][1])
    ref_i = c.sample_indices(256)
    ref_a = c.amplitudes()
    c.close()
    assert np.array_equal(amps[0], ref_a)
    assert np.array_equal(idx[0], ref_i)
    b.close()


def test_batch_run_reports_first_invalid_circuit(qlib):
    """qcs_hip_batch_run checks the circuits on host threads in chunks; the
    error it reports is still the first failing circuit in order (checked
    again on the calling thread for the message), whichever chunk finishes
    first.  No device work runs for a rejected batch."""
    import ctypes
    hip, _ = qlib
    nc = 200
    h = 2 ** -0.5
    G = (qcs_amd.HipGate * nc)()
    P = (qcs_amd.HipPass * nc)()
    for i in range(nc):
        G[i].kind, G[i].target, G[i].control = qcs_amd._native.G_1Q, i % 4, -1
        for j, v in enumerate((h, 0, h, 0, h, 0, -h, 0)):
            G[i].m[j] = v
        P[i].tile_bits, P[i].first_block, P[i].num_blocks = 4, i, 1
        for j in range(4):
            P[i].bits[j] = j

    def circuits():
        cs = (qcs_amd.HipBatchCircuit * nc)()
        for i in range(nc):
            cs[i].num_qubits, cs[i].precision, cs[i].tile_bits = 4, qcs_amd._native.C128, 4
            cs[i].first_pass, cs[i].num_passes = i, 1
        return cs

    b = ctypes.c_void_p()
    assert hip.qcs_hip_batch_create(ctypes.byref(b), 0) == 0
    try:
        u = (ctypes.c_double * nc)(*([0.5] * nc))
        out = (ctypes.c_long * nc)()

        def run(cs):
            return hip.qcs_hip_batch_run(b, cs, nc, P, nc, None, 0, G, nc, 1, u, out)

        assert run(circuits()) == 0  # the valid batch runs
        cs = circuits()
        cs[170].precision = 7
        assert run(cs) == -1
        assert b"circuit 170: unknown precision" in hip.qcs_hip_last_error()
        cs[150].num_qubits = 0
        assert run(cs) == -1
        assert b"circuit 150: num_qubits 0" in hip.qcs_hip_last_error()
        cs = circuits()
        G[190].target = 9  # outside circuit 190's 4 qubits and its pass bits
        assert run(cs) == -1
        assert b"pass 190: gate 190 outside the pass bits" in hip.qcs_hip_last_error()
        G[190].target = 1
        assert run(circuits()) == 0
    finally:
        hip.qcs_hip_batch_destroy(b)
