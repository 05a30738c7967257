"""The autotuner: it picks a fusion width and kernel variants for a fused
circuit without changing its state, picks batch plans without changing the
batch's results, and caches both per device."""
import os

import numpy as np
import pytest

import qcs_amd
from qcs_amd import circuits as C

pytestmark = pytest.mark.gpu


@pytest.fixture
def cache(tmp_path, monkeypatch):
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("QCS_TUNE_CACHE", str(path))
    return path


def test_fused_autotune_keeps_state_and_caches(qlib, cache):
    n = 14
    with qcs_amd.Circuit(n, qcs_amd.FUSED_FP32) as c:
        c.apply(C.layered(n, 1, 9))
        before = c.amplitudes()  # flushed
        assert 1 <= c.autotune() <= 5
        # tuning ran identity blocks only: the state is unchanged
        assert np.array_equal(c.amplitudes(), before)
        lines = cache.read_text().splitlines()
        assert len(lines) == 5 and all(" fused 14 " in l for l in lines)
    # with recorded gates the width follows them; results stay in tolerance
    ops = C.layered(n, 4, 3)
    with qcs_amd.Circuit(n, qcs_amd.FUSED_FP32) as c, qcs_amd.Circuit(n) as exact:
        c.apply(ops)
        k = c.autotune()
        assert 1 <= k <= 5
        exact.apply(ops)
        omf, _ = c.compare_with(exact)
        assert omf <= 1e-6
        # the second tuning read the cache
        assert len(cache.read_text().splitlines()) == 5


def test_batch_autotune_keeps_results(qlib, cl, cache):
    specs = [(15, C.layered(15, 3, 1), 0), (14, C.layered(14, 4, 2), 0),
             (16, C.layered(16, 3, 3), qcs_amd.FUSED_FP32),
             (12, C.layered(12, 3, 4), qcs_amd.FUSED_FP32)]
    def run(tune):
        circs = []
        for n, ops, opt in specs:
            c = qcs_amd.Circuit(n, qcs_amd.DEFERRED | opt)
            c.apply(ops)
            circs.append(c)
        b = qcs_amd.Batch(qcs_amd.BATCH_KEEP_STATES)
        if tune:
            b.autotune(circs)
        cl.srand(6)
        idx = b.sample(circs, 64)
        amps = [b.amplitudes(i, n) for i, (n, _, _) in enumerate(specs)]
        for c in circs:
            c.close()
        b.close()
        return idx, amps
    idx0, a0 = run(False)
    idx1, a1 = run(True)
    text = cache.read_text()
    assert " batch 0 14 " in text and " batch 0 15 " in text and " batch 1 16 " in text
    assert "batchthreads" in text
    for i, (n, _, opt) in enumerate(specs):
        if opt == 0:
            assert np.array_equal(a0[i], a1[i]) and np.array_equal(idx0[i], idx1[i])
        else:
            f = abs(np.vdot(a0[i], a1[i])) ** 2
            assert 1 - f <= 1e-6
    # cached: a second tuning adds nothing
    nlines = len(cache.read_text().splitlines())
    run(True)
    assert len(cache.read_text().splitlines()) == nlines


def tuning(hip, state, knob):
    import ctypes
    v = ctypes.c_int()
    assert hip.qcs_hip_get_tuning(state, knob, ctypes.byref(v)) == 0
    return v.value


def test_tile_autotune_times_the_pass_plan_and_caches(qlib, cache):
    # tile programs: fusion width, work-group size and grid are timed on the
    # circuit's own pass plan (identity blocks) and cached per gate structure
    hip, _ = qlib
    n = 18
    opts = qcs_amd.FUSED_FP32 | qcs_amd.FUSE_TILES
    with qcs_amd.Circuit(n, opts) as c, qcs_amd.Circuit(n) as exact:
        ops = C.layered(n, 6, 12)
        c.apply(ops)
        k = c.autotune()
        assert k in (3, 4)
        lines = cache.read_text().splitlines()
        assert len(lines) == 1 and f" tiles {n} " in lines[0]
        _, _, _, hsh, ck, cth, cgrid, cform, ms = lines[0].split()
        assert int(ck) == k and float(ms) > 0
        st = c.device_state  # flushes with the tuned plan
        assert tuning(hip, st, qcs_amd.TUNE_TILE_THREADS) == int(cth) in (512, 1024)
        assert tuning(hip, st, qcs_amd.TUNE_TILE_GRID) == int(cgrid)
        assert tuning(hip, st, qcs_amd.TUNE_TILE_K4_FORM) == int(cform) in (0, 1, 2)
        exact.apply(ops)
        omf, _ = c.compare_with(exact)
        assert omf <= 1e-6
        assert c.get_num_fused_passes() < c.get_num_fused_blocks()
    # the same structure with other angles: answered from the cache
    with qcs_amd.Circuit(n, opts) as c2, qcs_amd.Circuit(n) as exact:
        ops = C.layered(n, 6, 99)
        c2.apply(ops)
        assert c2.autotune() == k
        assert len(cache.read_text().splitlines()) == 1
        assert tuning(hip, c2.device_state, qcs_amd.TUNE_TILE_THREADS) == int(cth)
        exact.apply(ops)
        assert c2.compare_with(exact)[0] <= 1e-6
