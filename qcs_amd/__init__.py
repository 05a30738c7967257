"""qcs_amd -- MI355X (gfx950) state-vector engine behind the QCS C89 API.

The product is native: ``libqcs.so`` (C89, ``include/qcs.h``) over
``libqcs_hip.so`` (HIP kernels, ``include/qcs_hip.h``).  This module only
binds them with ctypes so Python tests and the bench can drive the same
``qc_*`` calls a C program makes.
"""
import ctypes

import numpy as np

from . import circuits
from ._native import (BATCH_GRAPH, BATCH_KEEP_STATES, BATCH_REPLAN, C64, C128, DEFERRED,
                      FUSE_SPLIT_BF16, FUSE_TILES, FUSED_FP32, FUSE_K, HIP_COMPAT, K_1Q,
                      K_COMPOSE, K_COUNT, K_CU, K_FUSED, K_OTHER, K_REDUCE,
                      PLAIN_FP32, PLAIN_FP64, REFERENCE_COMPAT,
                      RENORM_AT_BARRIER, TILE_MATH_FP32, TILE_MATH_SPLIT_BF16,
                      TUNE_TILE_DIRECT, TUNE_TILE_GRID, TUNE_TILE_K4_FORM, TUNE_TILE_MATH, TUNE_TILE_THREADS,
                      TILE_K4_CHAIN, TILE_K4_GAUSS, TILE_K4_HALVES,
                      HipBatchCircuit, HipBlock, HipGate, HipPass, libc, load)

__all__ = ["Circuit", "Batch", "load", "libc", "circuits", "PLAIN_FP64", "PLAIN_FP32",
           "DEFERRED", "BATCH_GRAPH", "BATCH_KEEP_STATES", "BATCH_REPLAN",
           "HipBatchCircuit",
           "FUSED_FP32", "REFERENCE_COMPAT", "RENORM_AT_BARRIER", "FUSE_K",
           "FUSE_TILES", "FUSE_SPLIT_BF16", "HipPass", "TUNE_TILE_THREADS", "TUNE_TILE_GRID",
           "TUNE_TILE_MATH", "TILE_MATH_FP32", "TILE_MATH_SPLIT_BF16",
           "TUNE_TILE_K4_FORM", "TUNE_TILE_DIRECT", "TILE_K4_CHAIN", "TILE_K4_HALVES", "TILE_K4_GAUSS",
           "C128", "C64", "HIP_COMPAT", "HipGate", "HipBlock", "K_1Q", "K_CU",
           "K_FUSED", "K_COMPOSE", "K_REDUCE", "K_OTHER", "K_COUNT"]


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Circuit:
    """One ``t_q_circuit`` (device-resident state + gate history)."""

    def __init__(self, num_qubits, options=PLAIN_FP64):
        _, self.lib = load()
        self.n = num_qubits
        self.options = options
        self._c = self.lib.qc_create_ex(num_qubits, options)
        if not self._c:
            raise RuntimeError(f"qc_create_ex({num_qubits}, {options:#x}) failed "
                               "(no HIP device or out of device memory)")
        L = self.lib
        self._dispatch = {
            circuits.H: lambda t, c, p: L.qc_h(self._c, t),
            circuits.X: lambda t, c, p: L.qc_x(self._c, t),
            circuits.Y: lambda t, c, p: L.qc_y(self._c, t),
            circuits.Z: lambda t, c, p: L.qc_z(self._c, t),
            circuits.P: lambda t, c, p: L.qc_phase(self._c, t, p),
            circuits.RX: lambda t, c, p: L.qc_rx(self._c, t, p),
            circuits.RY: lambda t, c, p: L.qc_ry(self._c, t, p),
            circuits.RZ: lambda t, c, p: L.qc_rz(self._c, t, p),
            circuits.CNOT: lambda t, c, p: L.qc_cnot(self._c, c, t),
            circuits.CPHASE: lambda t, c, p: L.qc_cphase(self._c, c, t, p),
            circuits.SWAP: lambda t, c, p: L.qc_swap(self._c, t, c),
        }

    # lifetime -----------------------------------------------------------------
    def close(self):
        if self._c:
            self.lib.qc_destroy(self._c)
            self._c = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._c

    @property
    def device_state(self):
        """qcs_hip_state * of this circuit (flushes recorded gates)."""
        return self.lib.qc_get_device_state(self._c)

    # gates ----------------------------------------------------------------------
    def apply(self, ops):
        """Apply a circuits.Ops list through the qc_* entry points."""
        d = self._dispatch
        for op, t, c, p in zip(ops.op, ops.tq, ops.cq, ops.par):
            d[op](int(t), int(c), float(p))

    def __getattr__(self, name):
        # qc_<name>(circuit, ...) passthrough, e.g. c.h(0), c.cnot(0, 1)
        if name.startswith("_") or name in ("lib", "n", "options"):
            raise AttributeError(name)
        fn = getattr(self.lib, "qc_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda *a: fn(self._c, *a)

    # observation -----------------------------------------------------------------
    def amplitudes(self, offset=0, count=None):
        if count is None:
            count = (1 << self.n) - offset
        out = np.empty(2 * count, np.float64)
        rc = self.lib.qc_get_amplitudes(self._c, offset, count, _dptr(out))
        if rc != 0:
            raise RuntimeError(f"qc_get_amplitudes failed ({rc})")
        return out.view(np.complex128)

    def set_amplitudes(self, amps, offset=0):
        a = np.ascontiguousarray(np.asarray(amps, np.complex128)).view(np.float64)
        rc = self.lib.qc_set_amplitudes(self._c, offset, a.size // 2, _dptr(a))
        if rc != 0:
            raise RuntimeError(f"qc_set_amplitudes failed ({rc})")

    def run_shots_hist(self, shots):
        hist = np.zeros(1 << self.n, np.int32)
        self.lib.qc_run_shots(self._c, shots,
                              hist.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return hist

    def sample_indices(self, shots):
        idx = np.full(shots, -2, np.int64)
        self.lib.qc_sample(self._c, shots,
                           idx.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        return idx

    def measure_all_list(self):
        res = (ctypes.c_int * max(self.n, 1))()
        self.lib.qc_measure_all(self._c, res)
        return list(res)[: self.n]

    def compare_with(self, other):
        mx = ctypes.c_double(0.0)
        omf = self.lib.qc_compare(self._c, other._c, ctypes.byref(mx))
        return omf, mx.value


class Batch:
    """One ``t_q_batch``: runs QCS_DEFERRED circuits in one device launch."""

    def __init__(self, options=0):
        _, self.lib = load()
        self._b = self.lib.qc_batch_create(options)
        if not self._b:
            raise RuntimeError(f"qc_batch_create({options:#x}) failed (no HIP device?)")

    def close(self):
        if self._b:
            self.lib.qc_batch_destroy(self._b)
            self._b = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _handles(circuits_):
        arr = (ctypes.c_void_p * max(len(circuits_), 1))()
        for i, c in enumerate(circuits_):
            arr[i] = c.handle
        return arr

    def sample(self, circuits_, shots):
        """int64 array (len(circuits), shots) of sampled indices."""
        out = np.full((len(circuits_), shots), -2, np.int64)
        rc = self.lib.qc_batch_sample(self._b, self._handles(circuits_), len(circuits_),
                                      shots, out.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        if rc != 0:
            raise RuntimeError(f"qc_batch_sample failed ({rc})")
        return out

    def run_shots(self, circuits_, shots):
        hists = [np.full(1 << c.n, -1, np.int32) for c in circuits_]
        ptrs = (ctypes.POINTER(ctypes.c_int) * max(len(hists), 1))()
        for i, h in enumerate(hists):
            ptrs[i] = h.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        rc = self.lib.qc_batch_run_shots(self._b, self._handles(circuits_), len(circuits_),
                                         shots, ptrs)
        if rc != 0:
            raise RuntimeError(f"qc_batch_run_shots failed ({rc})")
        return hists

    def amplitudes(self, i, n, offset=0, count=None):
        if count is None:
            count = (1 << n) - offset
        out = np.empty(2 * count, np.float64)
        rc = self.lib.qc_batch_get_amplitudes(self._b, i, offset, count, _dptr(out))
        if rc != 0:
            raise RuntimeError(f"qc_batch_get_amplitudes failed ({rc})")
        return out.view(np.complex128)

    def last_timing(self):
        k, t = ctypes.c_double(), ctypes.c_double()
        replayed = self.lib.qc_batch_last_timing(self._b, ctypes.byref(k), ctypes.byref(t))
        return k.value, t.value, bool(replayed)

    def set_plan(self, fused, num_qubits, tile_bits=0, fuse_k=0):
        rc = self.lib.qc_batch_set_plan(self._b, int(bool(fused)), num_qubits, tile_bits, fuse_k)
        if rc != 0:
            raise ValueError(f"qc_batch_set_plan rejected ({rc})")

    def set_threads(self, threads):
        rc = self.lib.qc_batch_set_threads(self._b, threads)
        if rc != 0:
            raise ValueError(f"qc_batch_set_threads rejected ({rc})")

    def autotune(self, circuits_):
        rc = self.lib.qc_batch_autotune(self._b, self._handles(circuits_), len(circuits_))
        if rc != 0:
            raise RuntimeError(f"qc_batch_autotune failed ({rc})")
