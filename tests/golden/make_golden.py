"""Write the golden vectors of tests/golden/ref_vectors.npz from the reference.

Runs oracle/_ref/libqcsref.so -- SegusFaultise/qcs compiled from its own
sources by oracle/Makefile (`make -C oracle ref`) -- on small seeded circuits
and records its outputs: amplitudes of both buffers (vector and
scratch_vector), measurement outcomes under srand() seeds, shot histograms,
Grover / QFT states.  These are the reference's *as-written* semantics, so
they pin the oracle's compat mode and the GPU compat mode bit for bit, and
pin the textbook mode wherever the two agree (1-qubit gates, measurement,
normalisation, sampling).

Usage:  python tests/golden/make_golden.py   (from the repository root)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib  # noqa: E402
from qcs_amd import circuits as C  # noqa: E402
from qcs_amd._native import libc  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")


def put_ops(store, key, ops):
    op, tq, cq, par = ops.arrays()
    store[key + "/op"] = op
    store[key + "/tq"] = tq
    store[key + "/cq"] = cq
    store[key + "/par"] = par


def main():
    ref = RefLib()
    L = ref.L
    cl = libc()
    store = {}

    # 1. two-qubit prepared state, then each controlled gate (as written)
    prep = C.Ops().add(C.RY, 0, -1, 0.7).add(C.RY, 1, -1, 1.9).add(C.RZ, 0, -1, 0.3)
    for name, extra in (("cnot01", C.Ops().add(C.CNOT, 1, 0)),
                        ("cnot10", C.Ops().add(C.CNOT, 0, 1)),
                        ("cphase10", C.Ops().add(C.CPHASE, 0, 1, 1.0))):
        ops = C.Ops()
        for a in zip(prep.op, prep.tq, prep.cq, prep.par):
            ops.add(*a)
        for a in zip(extra.op, extra.tq, extra.cq, extra.par):
            ops.add(*a)
        c = ref.create(2)
        ref.apply(c, ops)
        key = "circ/" + name
        put_ops(store, key, ops)
        store[key + "/n"] = np.int64(2)
        store[key + "/vec"] = ref.vector(c)
        store[key + "/scr"] = ref.scratch(c)
        L.qc_destroy(c)

    # 2. seeded random circuits, every gate kind (controlled gates as written)
    for n in (1, 2, 3, 4, 5, 6, 8, 10):
        for rep in range(2):
            seed = 1000 * n + rep
            ops = C.random_circuit(n, 40, seed)
            c = ref.create(n)
            ref.apply(c, ops)
            key = f"circ/rand_n{n}_s{seed}"
            put_ops(store, key, ops)
            store[key + "/n"] = np.int64(n)
            store[key + "/vec"] = ref.vector(c)
            store[key + "/scr"] = ref.scratch(c)
            L.qc_destroy(c)

    # 3. 1-qubit-only circuits (identical in both semantics)
    for n in (3, 7, 10):
        seed = 7000 + n
        ops = C.random_circuit(n, 60, seed, kinds=[C.H, C.X, C.Y, C.Z, C.P,
                                                   C.RX, C.RY, C.RZ])
        c = ref.create(n)
        ref.apply(c, ops)
        key = f"circ/oneq_n{n}_s{seed}"
        put_ops(store, key, ops)
        store[key + "/n"] = np.int64(n)
        store[key + "/vec"] = ref.vector(c)
        store[key + "/scr"] = ref.scratch(c)
        L.qc_destroy(c)

    # 4. layered circuit (the 20-qubit CPU workload's shape at small n), then shots
    for n, depth, seed in ((6, 3, 0), (8, 2, 5)):
        ops = C.layered(n, depth, seed)
        c = ref.create(n)
        ref.apply(c, ops)
        key = f"circ/layered_n{n}_d{depth}_s{seed}"
        put_ops(store, key, ops)
        store[key + "/n"] = np.int64(n)
        store[key + "/vec"] = ref.vector(c)
        store[key + "/scr"] = ref.scratch(c)
        hist = np.zeros(1 << n, np.int32)
        cl.srand(seed)
        L.qc_run_shots(c, 256, hist.ctypes.data_as(__import__("ctypes").POINTER(
            __import__("ctypes").c_int)))
        store[key + "/shots_seed"] = np.int64(seed)
        store[key + "/shots"] = np.int64(256)
        store[key + "/hist"] = hist
        L.qc_destroy(c)

    # 5. Grover (as written: stale-buffer oracle, unit-magnitude mean)
    for n, m in ((2, 3), (3, 5), (4, 9), (5, 17), (6, 42)):
        c = ref.create(n)
        L.qc_grover_search(c, m)
        key = f"grover/n{n}_m{m}"
        store[key + "/n"] = np.int64(n)
        store[key + "/m"] = np.int64(m)
        store[key + "/vec"] = ref.vector(c)
        store[key + "/scr"] = ref.scratch(c)
        store[key + "/num_gates"] = np.int64(L.qc_get_num_gates(c))
        L.qc_destroy(c)

    # 6. QFT on basis states (as written: the controlled phases are identities)
    for n, b in ((1, 1), (3, 0), (3, 1), (4, 6), (5, 19)):
        c = ref.create(n)
        for q in range(n):
            if (b >> q) & 1:
                L.qc_x(c, q)
        L.qc_quantum_fourier_transform(c)
        key = f"qft/n{n}_b{b}"
        store[key + "/n"] = np.int64(n)
        store[key + "/b"] = np.int64(b)
        store[key + "/vec"] = ref.vector(c)
        store[key + "/scr"] = ref.scratch(c)
        L.qc_destroy(c)

    # 7. measurement sequences under srand seeds (one rand() per measure)
    for n, seed in ((3, 11), (4, 12), (6, 13)):
        ops = C.random_circuit(n, 30, 500 + n, kinds=[C.H, C.RX, C.RY, C.RZ, C.P])
        c = ref.create(n)
        ref.apply(c, ops)
        cl.srand(seed)
        order = [(q * 5 + 1) % n for q in range(n)]
        outs = [L.qc_measure(c, q) for q in order]
        key = f"measure/n{n}_s{seed}"
        put_ops(store, key, ops)
        store[key + "/n"] = np.int64(n)
        store[key + "/seed"] = np.int64(seed)
        store[key + "/order"] = np.asarray(order, np.int64)
        store[key + "/outcomes"] = np.asarray(outs, np.int64)
        store[key + "/vec"] = ref.vector(c)
        L.qc_destroy(c)

    # 8. shot histograms under srand seeds
    for n, seed, shots in ((1, 3, 1000), (4, 21, 1024), (7, 22, 2000)):
        ops = C.random_circuit(n, 25, 900 + n, kinds=[C.H, C.RX, C.RY, C.RZ])
        c = ref.create(n)
        ref.apply(c, ops)
        hist = np.zeros(1 << n, np.int32)
        cl.srand(seed)
        L.qc_run_shots(c, shots, hist.ctypes.data_as(__import__("ctypes").POINTER(
            __import__("ctypes").c_int)))
        key = f"shots/n{n}_s{seed}"
        put_ops(store, key, ops)
        store[key + "/n"] = np.int64(n)
        store[key + "/seed"] = np.int64(seed)
        store[key + "/shots"] = np.int64(shots)
        store[key + "/vec"] = ref.vector(c)
        store[key + "/hist"] = hist
        L.qc_destroy(c)

    np.savez_compressed(OUT, **store)
    print(f"wrote {OUT}: {len(store)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
