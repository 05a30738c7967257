"""Independent numpy references used to pin the oracle (tests only).

Nothing here shares code with the oracle or the product: gate matrices are
the textbook closed forms, states are evolved with numpy tensor contractions
and, for small n, with explicitly built 2^n x 2^n Kronecker unitaries.
Qubit q is bit q of the basis index (qubit 0 least significant).
"""
import math

import numpy as np

from qcs_amd import circuits as C

SQ2 = 1.0 / math.sqrt(2.0)


def textbook(kind, a=0.0):
    """2x2 matrix of a gate kind (target part for CNOT / CPHASE)."""
    c, s = math.cos(a / 2), math.sin(a / 2)
    return {
        C.H: np.array([[SQ2, SQ2], [SQ2, -SQ2]], complex),
        C.X: np.array([[0, 1], [1, 0]], complex),
        C.Y: np.array([[0, -1j], [1j, 0]], complex),
        C.Z: np.array([[1, 0], [0, -1]], complex),
        C.P: np.array([[1, 0], [0, np.exp(1j * a)]], complex),
        C.RX: np.array([[c, -1j * s], [-1j * s, c]], complex),
        C.RY: np.array([[c, -s], [s, c]], complex),
        C.RZ: np.array([[np.exp(-0.5j * a), 0], [0, np.exp(0.5j * a)]], complex),
        C.CNOT: np.array([[0, 1], [1, 0]], complex),
        C.CPHASE: np.array([[1, 0], [0, np.exp(1j * a)]], complex),
    }[kind]


def m8(m2):
    """2x2 complex -> the 8-double {m00re, m00im, ...} layout."""
    return np.array([m2[0, 0].real, m2[0, 0].imag, m2[0, 1].real, m2[0, 1].imag,
                     m2[1, 0].real, m2[1, 0].imag, m2[1, 1].real, m2[1, 1].imag])


def apply_1q(psi, n, t, m):
    v = psi.reshape(1 << (n - t - 1), 2, 1 << t)
    return np.einsum("ij,ajb->aib", m, v).reshape(-1)


def apply_cu(psi, n, c, t, m):
    out = psi.copy()
    idx = np.arange(1 << n)
    sel = ((idx >> c) & 1 == 1) & ((idx >> t) & 1 == 0)
    i0 = idx[sel]
    i1 = i0 | (1 << t)
    v0, v1 = psi[i0], psi[i1]
    out[i0] = m[0, 0] * v0 + m[0, 1] * v1
    out[i1] = m[1, 0] * v0 + m[1, 1] * v1
    return out


def apply_swap(psi, n, a, b):
    idx = np.arange(1 << n)
    ba, bb = (idx >> a) & 1, (idx >> b) & 1
    perm = idx ^ ((ba ^ bb) << a) ^ ((ba ^ bb) << b)
    return psi[perm]


def simulate(n, ops, psi=None):
    """Textbook evolution of ``psi`` (default |0...0>) under ``ops``."""
    if psi is None:
        psi = np.zeros(1 << n, complex)
        psi[0] = 1
    for op, t, c, p in zip(ops.op, ops.tq, ops.cq, ops.par):
        if op in (C.CNOT, C.CPHASE):
            psi = apply_cu(psi, n, int(c), int(t), textbook(op, p))
        elif op == C.SWAP:
            psi = apply_swap(psi, n, int(t), int(c))
        else:
            psi = apply_1q(psi, n, int(t), textbook(op, p))
    return psi


def embed(n, t, m, c=None):
    """Dense 2^n x 2^n operator: m on qubit t (controlled on c if given)."""
    def kron_list(mats):
        out = np.array([[1.0 + 0j]])
        for q in reversed(range(n)):  # highest qubit is the leftmost factor
            out = np.kron(out, mats.get(q, np.eye(2)))
        return out
    if c is None:
        return kron_list({t: m})
    p0 = np.array([[1, 0], [0, 0]], complex)
    p1 = np.array([[0, 0], [0, 1]], complex)
    return kron_list({c: p0}) + kron_list({c: p1, t: m})


def dense_unitary(n, ops):
    U = np.eye(1 << n, dtype=complex)
    for op, t, c, p in zip(ops.op, ops.tq, ops.cq, ops.par):
        if op == C.SWAP:
            G = np.eye(1 << n, dtype=complex)[apply_swap(np.arange(1 << n), n, int(t), int(c))]
        elif op in (C.CNOT, C.CPHASE):
            G = embed(n, int(t), textbook(op, p), int(c))
        else:
            G = embed(n, int(t), textbook(op, p))
        U = G @ U
    return U


def bitrev(x, n):
    return int(format(x, f"0{n}b")[::-1], 2) if n > 0 else 0


def qft_closed_form(n, b, swaps=False):
    """The reference QFT order on |b>: amplitude at y is
    exp(2 pi i rev(b) y / N)/sqrt(N); with final swaps y -> rev(y)."""
    N = 1 << n
    y = np.arange(N)
    if swaps:
        y = np.array([bitrev(int(v), n) for v in y])
    return np.exp(2j * np.pi * bitrev(b, n) * y / N) / math.sqrt(N)


def grover_closed_form(n, m, k):
    """Textbook Grover from H^n|0>: marked sin((2k+1)t), others
    cos((2k+1)t)/sqrt(N-1), t = asin(1/sqrt(N))."""
    N = 1 << n
    th = math.asin(1.0 / math.sqrt(N))
    psi = np.full(N, math.cos((2 * k + 1) * th) / math.sqrt(N - 1), complex)
    psi[m] = math.sin((2 * k + 1) * th)
    return psi
