"""qcs_hip_sample_noalloc against qcs_hip_sample, through the C ABI.

The no-allocation entry point keeps its chunk sums in one device array per
GPU and takes caller-owned device buffers; its samples must equal the
allocating entry point's on the same state and uniforms, call after call,
including a call on a smaller state after a larger one (the array keeps the
previous call's contents).  Device buffers come from the HIP runtime the
backend already loaded (hipMalloc / hipMemcpy through ctypes)."""
import ctypes

import numpy as np
import pytest

import qcs_amd
from qcs_amd import circuits as C

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
EINVAL = -1  # QCS_HIP_EINVAL


@pytest.fixture(scope="module")
def rt(qlib):
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    return lib


class DevBuf:
    def __init__(self, rt, nbytes):
        self.rt = rt
        self.p = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, arr):
        assert arr.nbytes <= self.nbytes
        assert self.rt.hipMemcpy(self.p, arr.ctypes.data, arr.nbytes, H2D) == 0

    def get(self, arr):
        assert self.rt.hipMemcpy(arr.ctypes.data, self.p, arr.nbytes, D2H) == 0
        return arr

    def free(self):
        self.rt.hipFree(self.p)


def prepared(n, precision_opt, seed):
    c = qcs_amd.Circuit(n, precision_opt)
    c.apply(C.layered(n, 2, seed))
    c.flush()
    return c


def both(hip, rt, c, shots, seed):
    st = c.device_state
    u = np.random.default_rng(seed).random(shots)
    u[:3] = [0.0, 0.5, 1.0]  # edge uniforms: first index, middle, dropped
    want = np.full(shots, -2, np.int64)
    assert hip.qcs_hip_sample(st, shots, u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              want.ctypes.data_as(ctypes.POINTER(ctypes.c_long))) == 0
    du, di = DevBuf(rt, u.nbytes), DevBuf(rt, want.nbytes)
    du.put(u)
    di.put(np.full(shots, -3, np.int64))
    rc = hip.qcs_hip_sample_noalloc(st, shots, du.p, di.p)
    assert rc == 0, hip.qcs_hip_last_error()
    assert hip.qcs_hip_sync(st) == 0
    got = di.get(np.empty(shots, np.int64))
    du.free()
    di.free()
    return got, want


def test_noalloc_matches_sample_two_calls(qlib, rt):
    hip, _ = qlib
    # a larger state, then a smaller one: the second call reads no stale sums
    with prepared(22, qcs_amd.PLAIN_FP64, 5) as big:
        got, want = both(hip, rt, big, 5000, 1)
    assert np.array_equal(got, want)
    assert want[2] == -1 and np.all(want[3:] >= 0)
    with prepared(14, qcs_amd.PLAIN_FP64, 6) as small:
        got, want = both(hip, rt, small, 700, 2)
    assert np.array_equal(got, want)
    # complex64 and a state smaller than one chunk
    with prepared(11, qcs_amd.PLAIN_FP32, 7) as tiny:
        got, want = both(hip, rt, tiny, 300, 3)
    assert np.array_equal(got, want)


def test_noalloc_rejects_what_it_cannot_do(qlib, rt):
    hip, _ = qlib
    with prepared(8, qcs_amd.PLAIN_FP64, 1) as c:
        st = c.device_state
        d = DevBuf(rt, 64)
        host = np.zeros(8)
        # host memory the device cannot address, and null buffers
        assert hip.qcs_hip_sample_noalloc(st, 8, host.ctypes.data, d.p) == EINVAL
        assert b"not device-accessible" in hip.qcs_hip_last_error()
        assert hip.qcs_hip_sample_noalloc(st, 8, None, d.p) == EINVAL
        assert hip.qcs_hip_sample_noalloc(st, 0, None, None) == 0  # no shots, no work
        d.free()
    # beyond the scratch array's size (31 qubits, complex64: 16 GiB)
    with qcs_amd.Circuit(31, qcs_amd.PLAIN_FP32) as c:
        d = DevBuf(rt, 64)
        assert hip.qcs_hip_sample_noalloc(c.device_state, 8, d.p, d.p) == EINVAL
        assert b"exceed" in hip.qcs_hip_last_error()
        d.free()
