"""Seeded synthetic circuits used by the tests, the smoke check and the bench.

A circuit is four parallel arrays ``(op, target, control, param)``; the op
codes are shared with the CPU oracle (``oracle/qcs_oracle.h``):

    H=0 X=1 Y=2 Z=3 P=4 RX=5 RY=6 RZ=7 CNOT=8 CPHASE=9 SWAP=10

Angles come from splitmix64(seed): draw z, take its top 53 bits as
u in [0, 1) and use 2*pi*u.  The reference has no angle generator, so this
one is defined here and used everywhere the same seeded circuit is needed.
"""
import math

import numpy as np

H, X, Y, Z, P, RX, RY, RZ, CNOT, CPHASE, SWAP = range(11)
NAMES = ["H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "CNOT", "CPHASE", "SWAP"]
_MASK = (1 << 64) - 1


def splitmix64(seed):
    """Infinite splitmix64 stream from ``seed``."""
    state = seed & _MASK
    while True:
        state = (state + 0x9E3779B97F4A7C15) & _MASK
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        yield z ^ (z >> 31)


def uniforms(seed, count):
    """``count`` uniforms in [0, 1) from splitmix64(seed)."""
    g = splitmix64(seed)
    return [(next(g) >> 11) * (1.0 / 9007199254740992.0) for _ in range(count)]


class Ops:
    """Growing gate list."""

    def __init__(self):
        self.op, self.tq, self.cq, self.par = [], [], [], []

    def add(self, op, t, c=-1, p=0.0):
        self.op.append(op)
        self.tq.append(t)
        self.cq.append(c)
        self.par.append(p)
        return self

    def arrays(self):
        return (np.asarray(self.op, np.int32), np.asarray(self.tq, np.int32),
                np.asarray(self.cq, np.int32), np.asarray(self.par, np.float64))

    def __len__(self):
        return len(self.op)


def layered(n, depth, seed, initial_h=True):
    """H on every qubit, then ``depth`` layers of Rz, Ry, Rz on every qubit
    (angles in that order from splitmix64(seed)) followed by CNOT(q, q+1) for
    even q, then for odd q.  Gate count n + depth*(3n + n - 1)."""
    ops = Ops()
    if initial_h:
        for q in range(n):
            ops.add(H, q)
    g = splitmix64(seed)
    two_pi = 2.0 * math.pi
    for _ in range(depth):
        for q in range(n):
            for kind in (RZ, RY, RZ):
                u = (next(g) >> 11) * (1.0 / 9007199254740992.0)
                ops.add(kind, q, -1, two_pi * u)
        for q in range(0, n - 1, 2):
            ops.add(CNOT, q + 1, q)
        for q in range(1, n - 1, 2):
            ops.add(CNOT, q + 1, q)
    return ops


def qft(n, swaps=False):
    """The reference's QFT order (src/qcs.c:443-457): H(i), then
    CPHASE(control j, target i, pi/2^(j-i)) for j > i; optional final swaps."""
    ops = Ops()
    for i in range(n):
        ops.add(H, i)
        for j in range(i + 1, n):
            ops.add(CPHASE, i, j, math.pi / float(1 << (j - i)))
    if swaps:
        for i in range(n // 2):
            ops.add(SWAP, i, n - 1 - i)
    return ops


def random_circuit(n, count, seed, kinds=None):
    """``count`` gates drawn from ``kinds`` (default: every 1- and 2-qubit
    kind) on random qubits, from splitmix64(seed)."""
    kinds = list(kinds) if kinds is not None else [H, X, Y, Z, P, RX, RY, RZ,
                                                  CNOT, CPHASE]
    if n < 2:
        kinds = [k for k in kinds if k < CNOT]
    g = splitmix64(seed)
    ops = Ops()
    for _ in range(count):
        k = kinds[next(g) % len(kinds)]
        t = next(g) % n
        ang = 2.0 * math.pi * ((next(g) >> 11) * (1.0 / 9007199254740992.0))
        if k >= CNOT:
            c = next(g) % (n - 1)
            if c >= t:
                c += 1
            ops.add(k, t, c, ang)
        else:
            ops.add(k, t, -1, ang)
    return ops


def gate_count_layered(n, depth):
    return n + depth * (3 * n + n - 1)


def splitmix64_array(seed, count):
    """The first ``count`` outputs of splitmix64(seed) as a uint64 array."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & _MASK) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def layered_arrays(n, depth, seed, initial_h=True):
    """``layered(n, depth, seed)`` as four numpy arrays (op, target, control,
    param), built without a Python loop per gate."""
    u = (splitmix64_array(seed, 3 * n * depth) >> np.uint64(11)).astype(np.float64)
    ang = (2.0 * math.pi) * (u * (1.0 / 9007199254740992.0))
    rot_op = np.tile(np.array([RZ, RY, RZ]), n)
    rot_q = np.repeat(np.arange(n), 3)
    cn_c = np.concatenate([np.arange(0, n - 1, 2), np.arange(1, n - 1, 2)])
    per = 3 * n + len(cn_c)
    op = np.empty(depth * per, np.int64)
    tq = np.empty(depth * per, np.int64)
    cq = np.empty(depth * per, np.int64)
    par = np.zeros(depth * per, np.float64)
    for d in range(depth):
        s = d * per
        op[s:s + 3 * n] = rot_op
        tq[s:s + 3 * n] = rot_q
        cq[s:s + 3 * n] = -1
        par[s:s + 3 * n] = ang[3 * n * d:3 * n * (d + 1)]
        op[s + 3 * n:s + per] = CNOT
        tq[s + 3 * n:s + per] = cn_c + 1
        cq[s + 3 * n:s + per] = cn_c
    if initial_h:
        op = np.concatenate([np.full(n, H), op])
        tq = np.concatenate([np.arange(n), tq])
        cq = np.concatenate([np.full(n, -1), cq])
        par = np.concatenate([np.zeros(n), par])
    return op, tq, cq, par


def batch_workload(count=4096, seed=3, qubits=(10, 20), depths=(5, 40)):
    """Independent layered circuits for the batched form: circuit i has
    n = qubits[0] + floor(u * (qubits[1] - qubits[0] + 1)) qubits and depth
    depths[0] + floor(u' * (depths[1] - depths[0] + 1)), angles from its own
    seed; u, u' and the seed are consecutive splitmix64(seed) outputs.  Even
    circuits run complex128 plain, odd ones fused complex64.  Returns a list
    of (n, depth, circuit_seed, fused)."""
    z = splitmix64_array(seed, 3 * count).reshape(count, 3)
    out = []
    nq = qubits[1] - qubits[0] + 1
    nd = depths[1] - depths[0] + 1
    for i in range(count):
        un = float(int(z[i, 0]) >> 11) * (1.0 / 9007199254740992.0)
        ud = float(int(z[i, 1]) >> 11) * (1.0 / 9007199254740992.0)
        out.append((qubits[0] + int(un * nq), depths[0] + int(ud * nd),
                    int(z[i, 2]), i % 2 == 1))
    return out
