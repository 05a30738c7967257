"""Fused precision form on the GPU.

* qcs_hip_apply_unitary: a random dense k-qubit unitary (k = 1..5) on a
  complex64 state goes through the FP32 matrix-core kernel (n >= 12) or the
  FP32 VALU kernel (n < 12); compared with numpy in FP64: max-abs <= 2e-6
  on a unit-norm state.  complex128 states take the FP64 VALU kernel:
  <= 1e-12.
* Fused circuits vs the plain complex128 run of the same circuit on the same
  GPU: 1 - F <= 1e-6 (the precision form's tolerance).
"""
import ctypes
import math

import numpy as np
import pytest

import helpers as HP
import qcs_amd
from qcs_amd import circuits as C

K4_FORMS = (qcs_amd.TILE_K4_CHAIN, qcs_amd.TILE_K4_HALVES, qcs_amd.TILE_K4_GAUSS)

pytestmark = pytest.mark.gpu


def rand_unitary(k, rng):
    z = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def rand_state(n, rng):
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return v / np.linalg.norm(v)


def apply_dense(psi, n, targets, U):
    """numpy FP64 reference: U on ascending `targets` (targets[0] = U's LSB)."""
    k = len(targets)
    out = psi.copy()
    idx = np.arange(1 << n)
    mask = sum(1 << t for t in targets)
    base = idx[(idx & mask) == 0]
    offs = np.array([sum(((j >> q) & 1) << targets[q] for q in range(k))
                     for j in range(1 << k)])
    X = psi[base[None, :] + offs[:, None]]
    out[base[None, :] + offs[:, None]] = U @ X
    return out


def hip_apply_unitary(hip, state, targets, U):
    t = (ctypes.c_int * len(targets))(*targets)
    u = np.ascontiguousarray(U, np.complex128).view(np.float64)
    rc = hip.qcs_hip_apply_unitary(state, len(targets), t,
                                   u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, hip.qcs_hip_last_error()


TARGET_SETS = [[0], [5], [13], [0, 1], [3, 12], [0, 1, 2], [1, 6, 11],
               [10, 11, 12, 13], [0, 2, 4, 6], [0, 1, 2, 3, 4], [2, 5, 7, 9, 13],
               [9, 10, 11, 12, 13], [0, 7, 8, 12, 13]]


@pytest.mark.parametrize("n", [12, 14])
@pytest.mark.parametrize("targets", TARGET_SETS, ids=lambda t: "t" + "_".join(map(str, t)))
def test_mfma_unitary_vs_numpy(qlib, n, targets):
    if max(targets) >= n:
        pytest.skip("target beyond n")
    hip, _ = qlib
    rng = np.random.default_rng(1000 * n + sum(targets))
    psi = rand_state(n, rng)
    U = rand_unitary(len(targets), rng)
    with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as c:
        c.set_amplitudes(psi)
        hip_apply_unitary(hip, c.device_state, targets, U)
        got = c.amplitudes()
    ref = apply_dense(psi.astype(np.complex64).astype(np.complex128), n, targets, U)
    err = np.max(np.abs(got - ref))
    assert err < 2e-6, err


@pytest.mark.parametrize("n,targets", [(5, [0, 1, 2, 3, 4]), (8, [1, 3, 6]),
                                       (11, [0, 10]), (3, [2])])
def test_valu_fp32_unitary_small_n(qlib, n, targets):
    hip, _ = qlib
    rng = np.random.default_rng(n)
    psi = rand_state(n, rng)
    U = rand_unitary(len(targets), rng)
    with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as c:
        c.set_amplitudes(psi)
        hip_apply_unitary(hip, c.device_state, targets, U)
        got = c.amplitudes()
    assert np.max(np.abs(got - apply_dense(psi, n, targets, U))) < 2e-6


@pytest.mark.parametrize("n,targets", [(6, [0, 1, 2, 3, 4]), (13, [2, 7, 12]),
                                       (14, [0, 13])])
def test_fp64_unitary(qlib, n, targets):
    hip, _ = qlib
    rng = np.random.default_rng(7 * n)
    psi = rand_state(n, rng)
    U = rand_unitary(len(targets), rng)
    with qcs_amd.Circuit(n) as c:
        c.set_amplitudes(psi)
        hip_apply_unitary(hip, c.device_state, targets, U)
        got = c.amplitudes()
    np.testing.assert_allclose(got, apply_dense(psi, n, targets, U), rtol=0, atol=1e-12)


@pytest.mark.parametrize("n,depth,k", [(12, 3, 5), (13, 4, 4), (16, 5, 5),
                                       (18, 4, 3), (20, 6, 5), (22, 3, 5)])
def test_fused_layered_vs_plain_fp64(qlib, n, depth, k):
    print(f"[fused] layered n={n} depth={depth} k={k}", flush=True)
    ops = C.layered(n, depth, 2)
    with qcs_amd.Circuit(n) as ref, \
            qcs_amd.Circuit(n, qcs_amd.FUSED_FP32 | qcs_amd.FUSE_K(k)) as f:
        ref.apply(ops)
        f.apply(ops)
        omf, mx = f.compare_with(ref)
        assert omf <= 1e-6, omf
        assert mx < 1e-5, mx
        assert 0 < f.get_num_fused_blocks() < len(ops)
        assert abs(f.norm_sq() - 1.0) < 1e-6


@pytest.mark.parametrize("n", [4, 9, 12, 15])
def test_fused_random_circuits_vs_oracle_and_numpy(qlib, n):
    kinds = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
    ops = C.random_circuit(n, 120, 4242 + n, kinds=kinds)
    ref = HP.simulate(n, ops)
    with qcs_amd.Circuit(n, qcs_amd.FUSED_FP32) as f:
        f.apply(ops)
        got = f.amplitudes()
    ip = np.vdot(ref, got)
    assert 1 - abs(ip) ** 2 / (np.vdot(got, got).real) <= 1e-6
    assert np.max(np.abs(got - ref)) < 1e-5


@pytest.mark.parametrize("n,threads", [(16, 512), (17, 1024), (20, 512), (20, 1024)])
def test_tile_programs_equal_single_block_passes(qlib, n, threads):
    """Tile programs apply the same composed blocks, in the same order, as
    one pass per block.  They run them in other instruction forms (16x16x4
    halves, three real products per complex product at 1024 threads) than
    the one-block pass, which rounds differently, so the states agree to
    FP32 rounding: a block applied out of order, to the wrong bits or with a
    wrong operand would differ by O(1)."""
    from test_fusion_plan import pass_plan, to_gates  # noqa: F401
    hip, _ = qlib
    print(f"[fused] tile programs n={n} threads={threads}", flush=True)
    kinds = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
    ops = C.random_circuit(n, 300, 31337 + n, kinds=kinds)
    passes, blocks, ordered = pass_plan(ops, n)
    P = (qcs_amd.HipPass * len(passes))(*passes)
    B = (qcs_amd.HipBlock * len(blocks))(*blocks)
    G = (qcs_amd.HipGate * len(ordered))(*ordered)
    psi = rand_state(n, np.random.default_rng(n))
    with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as a, \
            qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as b:
        a.set_amplitudes(psi)
        b.set_amplitudes(psi)
        assert hip.qcs_hip_set_tuning(a.device_state, qcs_amd.TUNE_TILE_THREADS, threads) == 0
        rc = hip.qcs_hip_apply_passes(a.device_state, P, len(passes), B, len(blocks),
                                      G, len(ordered))
        assert rc == 0, hip.qcs_hip_last_error()
        rc = hip.qcs_hip_apply_fused(b.device_state, B, len(blocks), G, len(ordered))
        assert rc == 0, hip.qcs_hip_last_error()
        got, want = a.amplitudes(), b.amplitudes()
    assert np.linalg.norm(got - want) <= 1e-5
    ref = HP.simulate(n, ops, psi.astype(np.complex64).astype(complex))
    assert 1 - abs(np.vdot(ref, got)) ** 2 / np.vdot(got, got).real <= 1e-6


@pytest.mark.parametrize("n,seed", [(16, 1), (19, 2), (22, 3)])
def test_tile_k4_halves_form(qlib, n, seed):
    """The pipelined forms of k = 4 blocks (two 16-column halves on the
    16x16x4 instruction; smaller blocks padded to four targets; a block's
    second half stored behind the next block's first half, whose B operands
    are read ahead), with four real products per complex product or three
    (Gauss), apply the same blocks as the 32x32 chain: the states agree to
    FP32 rounding and stay within 1 - F <= 1e-6 of the FP64 simulation.  A
    block read ahead of what it depends on, a half stored to the wrong
    columns, a block padded on a wrong target or a wrong sign or operand in
    the three-product combination would differ by O(1)."""
    from test_fusion_plan import pass_plan
    hip, _ = qlib
    print(f"[fused] k4 halves n={n}", flush=True)
    kinds = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
    ops = C.random_circuit(n, 400, 777 * seed + n, kinds=kinds)
    for o in (ops, C.layered(n, 6, seed)):
        passes, blocks, ordered = pass_plan(o, n)
        assert any(b.k == 4 for b in blocks)
        if o is ops:  # the random circuit also has smaller blocks to pad
            assert any(b.k < 4 for b in blocks)
        P = (qcs_amd.HipPass * len(passes))(*passes)
        B = (qcs_amd.HipBlock * len(blocks))(*blocks)
        G = (qcs_amd.HipGate * len(ordered))(*ordered)
        psi = rand_state(n, np.random.default_rng(n + seed))
        ref = HP.simulate(n, o, psi.astype(np.complex64).astype(complex))
        got = {}
        for form in K4_FORMS:
            with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as st:
                st.set_amplitudes(psi)
                assert hip.qcs_hip_set_tuning(st.device_state, qcs_amd.TUNE_TILE_K4_FORM,
                                              form) == 0
                rc = hip.qcs_hip_apply_passes(st.device_state, P, len(passes), B,
                                              len(blocks), G, len(ordered))
                assert rc == 0, hip.qcs_hip_last_error()
                got[form] = st.amplitudes()
        for form in (qcs_amd.TILE_K4_HALVES, qcs_amd.TILE_K4_GAUSS):
            x = got[form]
            assert np.linalg.norm(x - got[qcs_amd.TILE_K4_CHAIN]) <= 1e-5, form
            assert 1 - abs(np.vdot(ref, x)) ** 2 / np.vdot(x, x).real <= 1e-6, form
    # through the qc_* API: a tile-program circuit with the form set
    ops = C.layered(n, 8, seed + 10)
    for form in (qcs_amd.TILE_K4_HALVES, qcs_amd.TILE_K4_GAUSS):
        with qcs_amd.Circuit(n) as ref, \
                qcs_amd.Circuit(n, qcs_amd.FUSED_FP32 | qcs_amd.FUSE_TILES) as f:
            assert hip.qcs_hip_set_tuning(f.device_state, qcs_amd.TUNE_TILE_K4_FORM,
                                          form) == 0
            ref.apply(ops)
            f.apply(ops)
            omf, mx = f.compare_with(ref)
            assert omf <= 1e-6 and mx < 1e-5, (form, omf, mx)
    with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as a:
        assert hip.qcs_hip_set_tuning(a.device_state, qcs_amd.TUNE_TILE_K4_FORM, 3) == -1


def direct_passes(passes, blocks):
    """Passes whose first and last blocks (before padding) leave qubits 0..4
    to the lanes: the ones the direct form can take."""
    def off_low(b):
        return not set(range(5)) & set(b.targets[:b.k])
    return [p for p in passes if p.num_blocks > 0 and off_low(blocks[p.first_block])
            and off_low(blocks[p.first_block + p.num_blocks - 1])]


@pytest.mark.parametrize("n,seed", [(16, 1), (19, 2), (22, 3)])
def test_tile_direct_passes_bit_identical(qlib, n, seed):
    """Direct passes of the three-product form (the first block's operands
    loaded from HBM into registers, the last block's results stored from
    registers, on passes whose first and last blocks leave qubits 0..4 to
    the lanes) compute the blocks of the LDS path: the states agree to FP32
    rounding with the direct form on and off (the two may put a different
    target on the lanes' row bit, so the products sum in another order) and
    stay within 1 - F <= 1e-6 of the FP64 simulation.  An operand or result
    at a wrong address, row or half would differ by O(1)."""
    from test_fusion_plan import pass_plan
    hip, _ = qlib
    print(f"[fused] tile direct n={n}", flush=True)
    kinds = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
    for i, o in enumerate((C.layered(n, 6, seed),
                           C.random_circuit(n, 400, 31 * seed + n, kinds=kinds))):
        passes, blocks, ordered = pass_plan(o, n)
        if i == 0:  # the layered circuit has passes the direct form takes
            assert direct_passes(passes, blocks)
        P = (qcs_amd.HipPass * len(passes))(*passes)
        B = (qcs_amd.HipBlock * len(blocks))(*blocks)
        G = (qcs_amd.HipGate * len(ordered))(*ordered)
        psi = rand_state(n, np.random.default_rng(3 * n + seed))
        ref = HP.simulate(n, o, psi.astype(np.complex64).astype(complex))
        got = []
        for on in (1, 0):
            with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as st:
                st.set_amplitudes(psi)
                assert hip.qcs_hip_set_tuning(st.device_state, qcs_amd.TUNE_TILE_K4_FORM,
                                              qcs_amd.TILE_K4_GAUSS) == 0
                assert hip.qcs_hip_set_tuning(st.device_state, qcs_amd.TUNE_TILE_DIRECT,
                                              on) == 0
                rc = hip.qcs_hip_apply_passes(st.device_state, P, len(passes), B,
                                              len(blocks), G, len(ordered))
                assert rc == 0, hip.qcs_hip_last_error()
                got.append(st.amplitudes())
        assert np.linalg.norm(got[0] - got[1]) <= 1e-5, (n, i, len(direct_passes(passes, blocks)))
        for x in got:
            assert 1 - abs(np.vdot(ref, x)) ** 2 / np.vdot(x, x).real <= 1e-6
    with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as a:
        assert hip.qcs_hip_set_tuning(a.device_state, qcs_amd.TUNE_TILE_DIRECT, 2) == -1
        v = ctypes.c_int()
        assert hip.qcs_hip_get_tuning(a.device_state, qcs_amd.TUNE_TILE_DIRECT,
                                      ctypes.byref(v)) == 0 and v.value == 1


@pytest.mark.parametrize("n,seed", [(16, 4), (20, 5), (22, 6), (24, 7)])
def test_apply_passes_normalize_equals_apply_then_normalize(qlib, n, seed):
    """qcs_hip_apply_passes_normalize (three-product form: the pass before
    the last sums |x|^2 of what it writes and the last pass scales its first
    block by 1/sqrt(norm_sq)) leaves the state of qcs_hip_apply_passes
    followed by qcs_hip_normalize: the two scale factors differ by the norm
    change of one FP32 pass (~1e-7), so the states agree to 1e-6, and a state
    of norm^2 2.89 comes out of norm 1 within 1e-5.  Covered: the last passes
    direct or through LDS, a one-pass plan (the last pass sums the norm, the
    rescale follows), a grid with more work-groups than the reduction buffer
    has partial slots and the chain form (the separate norm and rescale), a
    zero state (left as it is, q_state_normalize's norm_sq > 1e-12 rule)."""
    from test_fusion_plan import pass_plan
    hip, _ = qlib
    print(f"[fused] apply+normalize n={n}", flush=True)
    o = C.layered(n, 5, seed)
    passes, blocks, ordered = pass_plan(o, n)
    assert len(passes) >= 2
    psi = 1.7 * rand_state(n, np.random.default_rng(5 * n + seed))
    cases = [(qcs_amd.TILE_K4_GAUSS, 1, len(passes), 0),
             (qcs_amd.TILE_K4_GAUSS, 0, len(passes), 0),
             (qcs_amd.TILE_K4_GAUSS, 1, 1, 0),
             (qcs_amd.TILE_K4_CHAIN, 1, len(passes), 0)]
    if (1 << (n - 13)) > 1024:
        # grid 1024 launches 2048 work-groups on direct passes (two fit a
        # CU), more than the reduction buffer's 1024 partial slots: the
        # separate norm pass runs
        cases.append((qcs_amd.TILE_K4_GAUSS, 1, len(passes), 1024))
    for form, direct, npass, grid in cases:
        # a one-pass plan: the first pass's blocks only
        nblk = sum(p.num_blocks for p in passes[:npass])
        ngate = sum(b.num_gates for b in blocks[:nblk])
        P = (qcs_amd.HipPass * npass)(*passes[:npass])
        B = (qcs_amd.HipBlock * nblk)(*blocks[:nblk])
        G = (qcs_amd.HipGate * ngate)(*ordered[:ngate])
        got = []
        for fused in (True, False):
            with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as st:
                st.set_amplitudes(psi)
                assert hip.qcs_hip_set_tuning(st.device_state, qcs_amd.TUNE_TILE_K4_FORM,
                                              form) == 0
                assert hip.qcs_hip_set_tuning(st.device_state, qcs_amd.TUNE_TILE_DIRECT,
                                              direct) == 0
                if grid:
                    assert hip.qcs_hip_set_tuning(st.device_state, qcs_amd.TUNE_TILE_GRID,
                                                  grid) == 0
                if fused:
                    rc = hip.qcs_hip_apply_passes_normalize(st.device_state, P, npass,
                                                            B, nblk, G, ngate)
                else:
                    rc = hip.qcs_hip_apply_passes(st.device_state, P, npass, B, nblk, G,
                                                  ngate)
                    rc = rc or hip.qcs_hip_normalize(st.device_state, None)
                assert rc == 0, hip.qcs_hip_last_error()
                got.append(st.amplitudes())
        assert abs(np.vdot(got[0], got[0]).real - 1.0) <= 1e-5, (form, direct, npass, grid)
        assert np.linalg.norm(got[0] - got[1]) <= 1e-6, (form, direct, npass, grid)
    P = (qcs_amd.HipPass * len(passes))(*passes)
    B = (qcs_amd.HipBlock * len(blocks))(*blocks)
    G = (qcs_amd.HipGate * len(ordered))(*ordered)
    with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as st:
        st.set_amplitudes(np.zeros(1 << n, complex))
        assert hip.qcs_hip_apply_passes_normalize(st.device_state, P, len(passes), B,
                                                  len(blocks), G, len(ordered)) == 0
        assert not np.any(st.amplitudes())


def test_apply_passes_rejects_bad_plans(qlib):
    from test_fusion_plan import pass_plan
    hip, _ = qlib
    n = 16
    ops = C.layered(n, 2, 5)
    passes, blocks, ordered = pass_plan(ops, n)
    B = (qcs_amd.HipBlock * len(blocks))(*blocks)
    G = (qcs_amd.HipGate * len(ordered))(*ordered)
    for entry in (hip.qcs_hip_apply_passes, hip.qcs_hip_apply_passes_normalize):
        with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as a:
            bad = (qcs_amd.HipPass * len(passes))(*passes)
            bad[0].bits[0] = 1  # tile must start with bits 0..5
            assert entry(a.device_state, bad, len(passes), B, len(blocks), G,
                         len(ordered)) == -1
            bad = (qcs_amd.HipPass * len(passes))(*passes)
            bad[0].num_blocks += 1  # blocks not covered exactly once
            assert entry(a.device_state, bad, len(passes), B, len(blocks), G,
                         len(ordered)) == -1
        with qcs_amd.Circuit(n) as c128:
            good = (qcs_amd.HipPass * len(passes))(*passes)
            assert entry(c128.device_state, good, len(passes), B, len(blocks), G,
                         len(ordered)) == -5


@pytest.mark.parametrize("n,depth", [(16, 6), (20, 8), (24, 4)])
def test_fused_tile_mode_layered_vs_plain(qlib, n, depth):
    print(f"[fused] tile mode layered n={n}", flush=True)
    ops = C.layered(n, depth, 3)
    with qcs_amd.Circuit(n) as ref, \
            qcs_amd.Circuit(n, qcs_amd.FUSED_FP32 | qcs_amd.FUSE_TILES) as f:
        ref.apply(ops)
        f.apply(ops)
        omf, mx = f.compare_with(ref)
        assert omf <= 1e-6 and mx < 1e-5, (omf, mx)
        assert f.get_num_fused_passes() < f.get_num_fused_blocks()


@pytest.mark.parametrize("n,threads", [(16, 512), (17, 1024), (20, 1024)])
def test_split_bf16_tile_programs(qlib, n, threads):
    """Split-bf16 tile math applies the same blocks as the FP32 form within
    the rounding of the state to 16 significant bits per block: 1 - F
    against the FP64 simulation stays far below the form's 1e-6 (dropping a
    split product, e.g. the matrix's middle part, costs ~2^-8 per block and
    fails the 1e-8 bound here)."""
    from test_fusion_plan import pass_plan
    hip, _ = qlib
    print(f"[fused] split-bf16 tile programs n={n} threads={threads}", flush=True)
    kinds = [C.H, C.X, C.Y, C.Z, C.P, C.RX, C.RY, C.RZ, C.CNOT, C.CPHASE, C.SWAP]
    ops = C.random_circuit(n, 300, 4242 + n, kinds=kinds)
    passes, blocks, ordered = pass_plan(ops, n, low=4, max_blocks=15)
    P = (qcs_amd.HipPass * len(passes))(*passes)
    B = (qcs_amd.HipBlock * len(blocks))(*blocks)
    G = (qcs_amd.HipGate * len(ordered))(*ordered)
    psi = rand_state(n, np.random.default_rng(7 * n))
    with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as a, \
            qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as b:
        a.set_amplitudes(psi)
        b.set_amplitudes(psi)
        for st in (a, b):
            assert hip.qcs_hip_set_tuning(st.device_state, qcs_amd.TUNE_TILE_THREADS,
                                          threads) == 0
        assert hip.qcs_hip_set_tuning(a.device_state, qcs_amd.TUNE_TILE_MATH,
                                      qcs_amd.TILE_MATH_SPLIT_BF16) == 0
        for st in (a, b):
            rc = hip.qcs_hip_apply_passes(st.device_state, P, len(passes), B, len(blocks),
                                          G, len(ordered))
            assert rc == 0, hip.qcs_hip_last_error()
        got, fp32 = a.amplitudes(), b.amplitudes()
    ref = HP.simulate(n, ops, psi.astype(np.complex64).astype(complex))
    omf = 1 - abs(np.vdot(ref, got)) ** 2 / (np.vdot(got, got).real * np.vdot(ref, ref).real)
    omf32 = 1 - abs(np.vdot(ref, fp32)) ** 2 / (np.vdot(fp32, fp32).real
                                                 * np.vdot(ref, ref).real)
    print(f"[fused] split 1-F {omf:.3e} (FP32 {omf32:.3e}), {len(blocks)} blocks", flush=True)
    assert omf <= 1e-8
    assert np.max(np.abs(got - ref)) < 2e-5 * np.max(np.abs(ref)) * math.sqrt(len(blocks))


def test_split_bf16_pass_limit_and_options(qlib):
    from test_fusion_plan import pass_plan
    hip, qcs = qlib
    n = 16
    ops = C.random_circuit(12, 800, 5, kinds=[C.H, C.RX, C.RY, C.CNOT])
    passes, blocks, ordered = pass_plan(ops, n, low=4, max_blocks=22)
    assert max(p.num_blocks for p in passes) > 15  # the plan needs the FP32 limit
    P = (qcs_amd.HipPass * len(passes))(*passes)
    B = (qcs_amd.HipBlock * len(blocks))(*blocks)
    G = (qcs_amd.HipGate * len(ordered))(*ordered)
    with qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as a:
        assert hip.qcs_hip_set_tuning(a.device_state, qcs_amd.TUNE_TILE_MATH, 7) == -1
        assert hip.qcs_hip_set_tuning(a.device_state, qcs_amd.TUNE_TILE_MATH,
                                      qcs_amd.TILE_MATH_SPLIT_BF16) == 0
        assert hip.qcs_hip_apply_passes(a.device_state, P, len(passes), B, len(blocks),
                                        G, len(ordered)) == -1
    # the option belongs to the fused form
    assert not qcs.qc_create_ex(n, qcs_amd.FUSE_SPLIT_BF16)
    assert not qcs.qc_create_ex(n, qcs_amd.PLAIN_FP32 | qcs_amd.FUSE_SPLIT_BF16)


@pytest.mark.parametrize("n,depth", [(16, 6), (20, 8), (24, 4)])
def test_fused_split_mode_layered_vs_plain(qlib, n, depth):
    print(f"[fused] split-bf16 mode layered n={n}", flush=True)
    ops = C.layered(n, depth, 3)
    with qcs_amd.Circuit(n) as ref, \
            qcs_amd.Circuit(n, qcs_amd.FUSED_FP32 | qcs_amd.FUSE_SPLIT_BF16) as f:
        ref.apply(ops)
        f.apply(ops)
        omf, mx = f.compare_with(ref)
        print(f"[fused] split mode 1-F {omf:.3e} max-abs {mx:.3e}", flush=True)
        assert omf <= 1e-8 and mx < 1e-5, (omf, mx)
        assert f.get_num_fused_passes() < f.get_num_fused_blocks()


def test_fused_qft_with_swaps_closed_form(qlib):
    n, b = 14, 9999
    with qcs_amd.Circuit(n, qcs_amd.FUSED_FP32) as f:
        for q in range(n):
            if (b >> q) & 1:
                f.x(q)
        f.apply(C.qft(n, swaps=True))
        got = f.amplitudes()
    ref = HP.qft_closed_form(n, b, swaps=True)
    assert np.max(np.abs(got - ref)) < 1e-6
    assert 1 - abs(np.vdot(ref, got)) ** 2 <= 1e-6


def test_fused_observation_flushes_and_measures(qlib, cl):
    n = 13
    ops = C.layered(n, 2, 11)
    with qcs_amd.Circuit(n) as ref, qcs_amd.Circuit(n, qcs_amd.FUSED_FP32) as f:
        ref.apply(ops)
        f.apply(ops)
        assert f.find_most_likely_state() == ref.find_most_likely_state()
        cl.srand(17)
        a = [f.measure(q) for q in range(n)]
        cl.srand(17)
        b = [ref.measure(q) for q in range(n)]
        assert a == b
