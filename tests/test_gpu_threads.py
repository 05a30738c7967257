"""Separate circuits used from separate host threads at once (each owns its
HIP stream and buffers; the library documents this as supported).

Four threads each build their own circuits -- plain complex128, fused tile
programs (FP32 and split-bf16, whose kernels are set up once per process on
first use) and a deferred batch -- and their results must equal the same work
done on one thread, bit for bit (the kernels are deterministic for a given
circuit).  ctypes releases the GIL around every library call, so the threads
really do call into the library concurrently."""
import threading

import numpy as np
import pytest

import qcs_amd
from qcs_amd import circuits as C

pytestmark = pytest.mark.gpu


def work(seed):
    """amplitudes of a plain, two fused and one batched circuit"""
    n = 16
    ops = C.layered(n, 3, seed)
    out = {}
    with qcs_amd.Circuit(n) as p:
        p.apply(ops)
        out["plain"] = p.amplitudes()
    for name, opt in (("tiles", qcs_amd.FUSE_TILES), ("split", qcs_amd.FUSE_SPLIT_BF16)):
        with qcs_amd.Circuit(n, qcs_amd.FUSED_FP32 | opt) as f:
            f.apply(ops)
            out[name] = f.amplitudes()
    circs = []
    for k in range(3):
        c = qcs_amd.Circuit(10 + k, qcs_amd.DEFERRED)
        c.apply(C.layered(10 + k, 2, seed + k))
        circs.append(c)
    b = qcs_amd.Batch(qcs_amd.BATCH_KEEP_STATES)
    b.sample(circs, 0)
    out["batch"] = [b.amplitudes(i, 10 + i) for i in range(3)]
    b.close()
    for c in circs:
        c.close()
    return out


def test_circuits_on_separate_threads(qlib):
    seeds = [11, 12, 13, 14]
    want = {s: work(s) for s in seeds}
    got, errors = {}, []

    def run(s):
        try:
            got[s] = work(s)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(s,)) for s in seeds]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for s in seeds:
        for name in ("plain", "tiles", "split"):
            assert np.array_equal(got[s][name], want[s][name]), (s, name)
        for a, b in zip(got[s]["batch"], want[s]["batch"]):
            assert np.array_equal(a, b), s
