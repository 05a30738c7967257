"""ctypes bindings of the two native libraries of the package.

``libqcs_hip.so`` is the gfx950 backend (C ABI of ``include/qcs_hip.h``);
``libqcs.so`` is the C89 host library implementing ``include/qcs.h`` on top
of it.  Both are built in place by ``make -C qcs_amd``.  Loading fails with
an exception when either is missing: there is no Python or CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_HERE, "libqcs_hip.so")
QCS_LIB_PATH = os.path.join(_HERE, "libqcs.so")

c_int, c_long, c_double, c_void_p = (ctypes.c_int, ctypes.c_long,
                                     ctypes.c_double, ctypes.c_void_p)
P = ctypes.POINTER

# qcs_hip.h constants
C128, C64 = 0, 1
HIP_COMPAT = 1
G_1Q, G_CU, G_SWAP = 0, 1, 2
K_1Q, K_CU, K_FUSED, K_COMPOSE, K_REDUCE, K_OTHER, K_COUNT = range(7)
MAX_FUSE = 5

# qcs.h options
PLAIN_FP64 = 0x0
FUSED_FP32 = 0x1
PLAIN_FP32 = 0x2
REFERENCE_COMPAT = 0x4
RENORM_AT_BARRIER = 0x8


def FUSE_K(k):
    return k << 8


class HipGate(ctypes.Structure):
    _fields_ = [("kind", c_int), ("target", c_int), ("control", c_int),
                ("reserved", c_int), ("m", c_double * 8)]


class HipBlock(ctypes.Structure):
    _fields_ = [("k", c_int), ("first_gate", c_int), ("num_gates", c_int),
                ("reserved", c_int), ("targets", c_int * 8)]


class HipPass(ctypes.Structure):
    _fields_ = [("tile_bits", c_int), ("first_block", c_int), ("num_blocks", c_int),
                ("reserved", c_int), ("bits", c_int * 16)]


TILE_BITS, TILE_LOW_BITS = 13, 4
TUNE_TILE_THREADS, TUNE_TILE_GRID, TUNE_TILE_MATH, TUNE_TILE_K4_FORM = 0, 1, 2, 3
TUNE_TILE_DIRECT = 4
TILE_K4_CHAIN, TILE_K4_HALVES, TILE_K4_GAUSS = 0, 1, 2
TILE_MATH_FP32, TILE_MATH_SPLIT_BF16 = 0, 1
FUSE_TILES = 0x10
FUSE_SPLIT_BF16 = 0x40
DEFERRED = 0x20

# batched circuits (qcs_hip.h / qcs.h)
BATCH_MAX_QUBITS = 22
BATCH_C128_TILE_BITS, BATCH_C64_TILE_BITS = 13, 14
HIP_BATCH_KEEP_STATES, HIP_BATCH_GRAPH = 1, 2
TUNE_BATCH_THREADS = 0
BATCH_GRAPH, BATCH_KEEP_STATES, BATCH_REPLAN = 0x1, 0x2, 0x4


class HipBatchCircuit(ctypes.Structure):
    _fields_ = [("num_qubits", c_int), ("precision", c_int), ("tile_bits", c_int),
                ("first_pass", c_int), ("num_passes", c_int),
                ("reserved", c_int * 3)]

_HIP_PROTOS = {
    "qcs_hip_last_error": (ctypes.c_char_p, []),
    "qcs_hip_device_available": (c_int, []),
    "qcs_hip_device_sync": (c_int, []),
    "qcs_hip_device_name": (c_int, [ctypes.c_char_p, c_int]),
    "qcs_hip_state_create": (c_int, [P(c_void_p), c_int, c_int, c_int]),
    "qcs_hip_state_destroy": (None, [c_void_p]),
    "qcs_hip_state_num_qubits": (c_int, [c_void_p]),
    "qcs_hip_state_precision": (c_int, [c_void_p]),
    "qcs_hip_state_flags": (c_int, [c_void_p]),
    "qcs_hip_state_set_basis": (c_int, [c_void_p, c_long]),
    "qcs_hip_sync": (c_int, [c_void_p]),
    "qcs_hip_read": (c_int, [c_void_p, c_long, c_long, P(c_double)]),
    "qcs_hip_write": (c_int, [c_void_p, c_long, c_long, P(c_double)]),
    "qcs_hip_read_scratch": (c_int, [c_void_p, c_long, c_long, P(c_double)]),
    "qcs_hip_apply_1q": (c_int, [c_void_p, c_int, P(c_double)]),
    "qcs_hip_apply_cu": (c_int, [c_void_p, c_int, c_int, P(c_double)]),
    "qcs_hip_apply_swap": (c_int, [c_void_p, c_int, c_int]),
    "qcs_hip_phase_flip": (c_int, [c_void_p, c_long]),
    "qcs_hip_diffusion": (c_int, [c_void_p]),
    "qcs_hip_grover": (c_int, [c_void_p, c_long, c_long]),
    "qcs_hip_norm_sq": (c_int, [c_void_p, P(c_double)]),
    "qcs_hip_normalize": (c_int, [c_void_p, P(c_double)]),
    "qcs_hip_qubit_probs": (c_int, [c_void_p, c_int, P(c_double), P(c_double)]),
    "qcs_hip_measure": (c_int, [c_void_p, c_int, c_double, P(c_int)]),
    "qcs_hip_measure_seq": (c_int, [c_void_p, c_int, P(c_int), P(c_double), P(c_int)]),
    "qcs_hip_probability": (c_int, [c_void_p, c_long, P(c_double)]),
    "qcs_hip_argmax": (c_int, [c_void_p, P(c_long)]),
    "qcs_hip_sample": (c_int, [c_void_p, c_long, P(c_double), P(c_long)]),
    "qcs_hip_sample_noalloc": (c_int, [c_void_p, c_long, c_void_p, c_void_p]),
    "qcs_hip_apply_fused": (c_int, [c_void_p, P(HipBlock), c_int, P(HipGate), c_int]),
    "qcs_hip_apply_unitary": (c_int, [c_void_p, c_int, P(c_int), P(c_double)]),
    "qcs_hip_compare": (c_int, [c_void_p, c_void_p, P(c_double), P(c_double)]),
    "qcs_hip_apply_passes": (c_int, [c_void_p, P(HipPass), c_int, P(HipBlock), c_int,
                                     P(HipGate), c_int]),
    "qcs_hip_apply_passes_normalize": (c_int, [c_void_p, P(HipPass), c_int, P(HipBlock),
                                               c_int, P(HipGate), c_int]),
    "qcs_hip_set_tuning": (c_int, [c_void_p, c_int, c_int]),
    "qcs_hip_get_tuning": (c_int, [c_void_p, c_int, P(c_int)]),
    "qcs_hip_timing_enable": (c_int, [c_void_p, c_int]),
    "qcs_hip_timing_read": (c_int, [c_void_p, c_int, P(c_double), P(c_long),
                                    P(c_double), P(c_double)]),
    "qcs_hip_timing_moved": (c_int, [c_void_p, c_int, P(c_double)]),
    "qcs_hip_batch_create": (c_int, [P(c_void_p), c_int]),
    "qcs_hip_batch_destroy": (None, [c_void_p]),
    "qcs_hip_batch_set_tuning": (c_int, [c_void_p, c_int, c_int]),
    "qcs_hip_batch_run": (c_int, [c_void_p, P(HipBatchCircuit), c_int, P(HipPass), c_int,
                                  P(HipBlock), c_int, P(HipGate), c_int, c_long,
                                  P(c_double), P(c_long)]),
    "qcs_hip_batch_rerun": (c_int, [c_void_p, c_long, P(c_double), P(c_long)]),
    "qcs_hip_batch_read_state": (c_int, [c_void_p, c_int, c_long, c_long, P(c_double)]),
    "qcs_hip_batch_last_timing": (c_int, [c_void_p, P(c_double), P(c_double), P(c_int)]),
}

_Q = c_void_p  # t_q_circuit *
_QCS_PROTOS = {
    "qc_create": (_Q, [c_int]),
    "qc_destroy": (None, [_Q]),
    "qc_h": (None, [_Q, c_int]),
    "qc_x": (None, [_Q, c_int]),
    "qc_y": (None, [_Q, c_int]),
    "qc_z": (None, [_Q, c_int]),
    "qc_cnot": (None, [_Q, c_int, c_int]),
    "qc_phase": (None, [_Q, c_int, c_double]),
    "qc_rx": (None, [_Q, c_int, c_double]),
    "qc_ry": (None, [_Q, c_int, c_double]),
    "qc_rz": (None, [_Q, c_int, c_double]),
    "qc_barrier": (None, [_Q]),
    "qc_reset": (None, [_Q, c_int]),
    "qc_measure": (c_int, [_Q, c_int]),
    "qc_measure_all": (None, [_Q, P(c_int)]),
    "qc_run": (None, [_Q]),
    "qc_run_shots": (None, [_Q, c_int, P(c_int)]),
    "qc_find_most_likely_state": (c_int, [_Q]),
    "qc_get_probability": (c_double, [_Q, c_int]),
    "qc_print_state": (None, [_Q, c_int]),
    "qc_print_circuit": (None, [_Q]),
    "qc_grover_search": (None, [_Q, c_int]),
    "qc_quantum_fourier_transform": (None, [_Q]),
    "qc_bernstein_vazirani": (None, [_Q, c_int]),
    "qc_ghz_state": (None, [_Q]),
    "qc_get_num_qubits": (c_int, [_Q]),
    "qc_get_num_gates": (c_int, [_Q]),
    "qc_optimize": (None, [_Q]),
    "qc_cphase": (None, [_Q, c_int, c_int, c_double]),
    "qc_swap": (None, [_Q, c_int, c_int]),
    "qc_create_ex": (_Q, [c_int, c_int]),
    "qc_get_options": (c_int, [_Q]),
    "qc_flush": (None, [_Q]),
    "qc_last_error": (c_int, [_Q]),
    "qc_get_probability_l": (c_double, [_Q, c_long]),
    "qc_find_most_likely_state_l": (c_long, [_Q]),
    "qc_grover_search_l": (None, [_Q, c_long]),
    "qc_grover_search_iters": (None, [_Q, c_long, c_long]),
    "qc_sample": (None, [_Q, c_int, P(c_long)]),
    "qc_get_amplitudes": (c_int, [_Q, c_long, c_long, P(c_double)]),
    "qc_set_amplitudes": (c_int, [_Q, c_long, c_long, P(c_double)]),
    "qc_norm_sq": (c_double, [_Q]),
    "qc_compare": (c_double, [_Q, _Q, P(c_double)]),
    "qc_get_num_fused_blocks": (c_long, [_Q]),
    "qc_get_num_fused_passes": (c_long, [_Q]),
    "qc_get_device_state": (c_void_p, [_Q]),
    "qc_set_basis_state": (None, [_Q, c_long]),
    "qcs_fusion_plan": (c_int, [P(HipGate), c_int, c_int, c_int, P(P(HipBlock)),
                                P(c_int), P(P(HipGate))]),
    "qcs_pass_plan": (c_int, [P(HipGate), c_int, c_int, c_int, c_int, c_int, c_int,
                              P(P(HipPass)), P(c_int), P(P(HipBlock)), P(c_int),
                              P(P(HipGate))]),
    "qcs_gate_pass_plan": (c_int, [P(HipGate), c_int, c_int, c_int, c_int,
                                   P(P(HipPass)), P(c_int)]),
    "qcs_gate_is_exact": (c_int, [P(HipGate)]),
    "qcs_gate_hoist_exact": (c_int, [P(HipGate), c_int, c_int, P(P(HipGate))]),
    "qc_batch_create": (c_void_p, [c_int]),
    "qc_batch_destroy": (None, [c_void_p]),
    "qc_batch_sample": (c_int, [c_void_p, P(_Q), c_int, c_int, P(c_long)]),
    "qc_batch_run_shots": (c_int, [c_void_p, P(_Q), c_int, c_int, P(P(c_int))]),
    "qc_batch_get_amplitudes": (c_int, [c_void_p, c_int, c_long, c_long, P(c_double)]),
    "qc_batch_last_timing": (c_int, [c_void_p, P(c_double), P(c_double)]),
    "qc_batch_set_plan": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "qc_batch_set_threads": (c_int, [c_void_p, c_int]),
    "qc_autotune": (c_int, [_Q]),
    "qc_batch_autotune": (c_int, [c_void_p, P(_Q), c_int]),
}

_hip = None
_qcs = None


def _bind(lib, protos):
    for name, (res, args) in protos.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load():
    """Load (once) and return ``(hip_lib, qcs_lib)``; raises when unbuilt."""
    global _hip, _qcs
    if _qcs is not None:
        return _hip, _qcs
    for p in (HIP_LIB_PATH, QCS_LIB_PATH):
        if not os.path.exists(p):
            raise RuntimeError(
                f"{os.path.basename(p)} is not built; run `make -C qcs_amd` "
                "(or __graft_entry__.build()) first")
    hip = ctypes.CDLL(HIP_LIB_PATH)
    qcs = ctypes.CDLL(QCS_LIB_PATH)
    _bind(hip, _HIP_PROTOS)
    _bind(qcs, _QCS_PROTOS)
    _hip, _qcs = hip, qcs
    return hip, qcs


def libc():
    """The process's C library (srand/rand feed qc_measure / qc_run_shots)."""
    lib = ctypes.CDLL(None)
    lib.srand.argtypes = [ctypes.c_uint]
    lib.srand.restype = None
    lib.rand.argtypes = []
    lib.rand.restype = c_int
    return lib
