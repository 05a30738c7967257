"""ctypes bindings of the CPU checkers -- TEST INFRASTRUCTURE ONLY.

``OracleState`` wraps liboracle.so (the C restatement in qcs_oracle.c).
``RefLib`` wraps oracle/_ref/libqcsref*.so, the reference compiled from its
own sources (present only where the reference tree was available at build
time; tests that need it skip otherwise).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

c_int, c_long, c_double, c_void_p = (ctypes.c_int, ctypes.c_long,
                                     ctypes.c_double, ctypes.c_void_p)
PD = ctypes.POINTER(c_double)
PI = ctypes.POINTER(c_int)
PL = ctypes.POINTER(c_long)

_lib = None


def _dp(a):
    return a.ctypes.data_as(PD)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(ORACLE_LIB)
        protos = {
            "oq_create": (c_void_p, [c_int, c_int]),
            "oq_destroy": (None, [c_void_p]),
            "oq_size": (c_long, [c_void_p]),
            "oq_vector": (PD, [c_void_p]),
            "oq_scratch": (PD, [c_void_p]),
            "oq_apply_1q": (None, [c_void_p, c_int, PD]),
            "oq_apply_cu": (None, [c_void_p, c_int, c_int, PD]),
            "oq_apply_swap": (None, [c_void_p, c_int, c_int]),
            "oq_phase_flip": (None, [c_void_p, c_long]),
            "oq_diffusion": (None, [c_void_p]),
            "oq_grover_iterations": (c_long, [c_int, c_int]),
            "oq_normalize": (c_double, [c_void_p]),
            "oq_norm_sq": (c_double, [c_void_p]),
            "oq_measure": (c_int, [c_void_p, c_int, c_double]),
            "oq_sample": (None, [c_void_p, c_int, PD, PL]),
            "oq_probability": (c_double, [c_void_p, c_long]),
            "oq_argmax": (c_long, [c_void_p]),
            "oq_gate": (None, [c_int, c_double, PD]),
            "oq_run": (None, [c_void_p, c_int, PI, PI, PI, PD]),
            "oq_qft": (None, [c_void_p]),
            "oq_grover": (None, [c_void_p, c_long, c_long]),
            "oq_apply_1q_omp": (None, [c_void_p, c_int, PD, c_int]),
            "oq_apply_cu_omp": (None, [c_void_p, c_int, c_int, PD, c_int]),
            "oq_run_omp": (None, [c_void_p, c_int, PI, PI, PI, PD, c_int]),
            "oq_set_threads": (None, [c_int]),
        }
        for name, (res, args) in protos.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def gate(kind, angle=0.0):
    m = np.zeros(8, np.float64)
    lib().oq_gate(kind, angle, _dp(m))
    return m


class OracleState:
    """Double-buffered FP64 state of the oracle (compat or textbook)."""

    def __init__(self, n, compat=False):
        self.L = lib()
        self.n = n
        self.s = self.L.oq_create(n, 1 if compat else 0)
        if not self.s:
            raise MemoryError(f"oracle state of {n} qubits")

    def close(self):
        if self.s:
            self.L.oq_destroy(self.s)
            self.s = None

    def __del__(self):
        self.close()

    def vector(self):
        size = 1 << self.n
        buf = np.ctypeslib.as_array(self.L.oq_vector(self.s), shape=(2 * size,))
        return buf.copy().view(np.complex128)

    def scratch(self):
        size = 1 << self.n
        buf = np.ctypeslib.as_array(self.L.oq_scratch(self.s), shape=(2 * size,))
        return buf.copy().view(np.complex128)

    def set_vector(self, amps):
        size = 1 << self.n
        buf = np.ctypeslib.as_array(self.L.oq_vector(self.s), shape=(2 * size,))
        buf[:] = np.asarray(amps, np.complex128).view(np.float64)

    def run(self, ops):
        op, tq, cq, par = ops.arrays()
        self.L.oq_run(self.s, len(op), op.ctypes.data_as(PI), tq.ctypes.data_as(PI),
                      cq.ctypes.data_as(PI), _dp(par))

    def run_omp(self, ops, racy=False):
        """ops with the OpenMP loops (race-fixed unless racy)"""
        op, tq, cq, par = ops.arrays()
        self.L.oq_run_omp(self.s, len(op), op.ctypes.data_as(PI), tq.ctypes.data_as(PI),
                          cq.ctypes.data_as(PI), _dp(par), 1 if racy else 0)

    def apply_1q(self, t, m):
        m = np.ascontiguousarray(m, np.float64)
        self.L.oq_apply_1q(self.s, t, _dp(m))

    def apply_cu(self, c, t, m):
        m = np.ascontiguousarray(m, np.float64)
        self.L.oq_apply_cu(self.s, c, t, _dp(m))

    def phase_flip(self, i):
        self.L.oq_phase_flip(self.s, i)

    def diffusion(self):
        self.L.oq_diffusion(self.s)

    def normalize(self):
        return self.L.oq_normalize(self.s)

    def norm_sq(self):
        return self.L.oq_norm_sq(self.s)

    def measure(self, q, u):
        return self.L.oq_measure(self.s, q, u)

    def sample(self, u):
        u = np.ascontiguousarray(u, np.float64)
        out = np.empty(len(u), np.int64)
        self.L.oq_sample(self.s, len(u), _dp(u), out.ctypes.data_as(PL))
        return out

    def probability(self, i):
        return self.L.oq_probability(self.s, i)

    def argmax(self):
        return self.L.oq_argmax(self.s)

    def qft(self):
        self.L.oq_qft(self.s)

    def grover(self, marked, iterations):
        self.L.oq_grover(self.s, marked, iterations)


class RefLib:
    """The reference library compiled from its own sources (oracle/_ref)."""

    def __init__(self, variant="libqcsref.so"):
        path = os.path.join(REF_DIR, variant)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        Q = c_void_p
        protos = {
            "qc_create": (Q, [c_int]), "qc_destroy": (None, [Q]),
            "qc_h": (None, [Q, c_int]), "qc_x": (None, [Q, c_int]),
            "qc_y": (None, [Q, c_int]), "qc_z": (None, [Q, c_int]),
            "qc_cnot": (None, [Q, c_int, c_int]),
            "qc_phase": (None, [Q, c_int, c_double]),
            "qc_rx": (None, [Q, c_int, c_double]),
            "qc_ry": (None, [Q, c_int, c_double]),
            "qc_rz": (None, [Q, c_int, c_double]),
            "qc_cphase": (None, [Q, c_int, c_int, c_double]),
            "qc_measure": (c_int, [Q, c_int]),
            "qc_run_shots": (None, [Q, c_int, PI]),
            "qc_get_probability": (c_double, [Q, c_int]),
            "qc_find_most_likely_state": (c_int, [Q]),
            "qc_grover_search": (None, [Q, c_int]),
            "qc_quantum_fourier_transform": (None, [Q]),
            "qc_get_num_gates": (c_int, [Q]),
            "q_state_normalize": (None, [c_void_p]),
            "qcsref_vector": (PD, [Q]), "qcsref_scratch": (PD, [Q]),
            "qcsref_size": (c_long, [Q]),
        }
        for name, (res, args) in protos.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        self.L = L

    def create(self, n):
        return self.L.qc_create(n)

    def vector(self, c):
        size = self.L.qcsref_size(c)
        return np.ctypeslib.as_array(self.L.qcsref_vector(c),
                                     shape=(2 * size,)).copy().view(np.complex128)

    def scratch(self, c):
        size = self.L.qcsref_size(c)
        return np.ctypeslib.as_array(self.L.qcsref_scratch(c),
                                     shape=(2 * size,)).copy().view(np.complex128)

    def apply(self, c, ops):
        L = self.L
        for op, t, k, p in zip(ops.op, ops.tq, ops.cq, ops.par):
            t, k, p = int(t), int(k), float(p)
            if op == 0: L.qc_h(c, t)
            elif op == 1: L.qc_x(c, t)
            elif op == 2: L.qc_y(c, t)
            elif op == 3: L.qc_z(c, t)
            elif op == 4: L.qc_phase(c, t, p)
            elif op == 5: L.qc_rx(c, t, p)
            elif op == 6: L.qc_ry(c, t, p)
            elif op == 7: L.qc_rz(c, t, p)
            elif op == 8: L.qc_cnot(c, k, t)
            elif op == 9: L.qc_cphase(c, k, t, p)
            else: raise ValueError(f"op {op} has no reference counterpart")
