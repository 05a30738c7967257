"""Full-size single-GPU cases, checked through properties that do not need
a CPU run of the whole state:

  * 28 qubits, complex128 plain: the depth-20 layered circuit followed by
    its inverse returns to |0...0>; the 28-qubit QFT (with final swaps) of a
    basis state |b>, b uniform from seed 1, equals the closed form
    exp(2 pi i rev(b) rev(y) / N) / sqrt(N) on sampled index ranges.
  * 33 qubits, complex128 (128 GiB): Grover for the marked index drawn from
    seed 4, truncated to 64 iterations, against the analytic amplitudes
    sin((2k+1) t), cos((2k+1) t)/sqrt(N-1); the complex64 run of the same
    iterations (64 GiB, beside it) reports 1 - F against it.
  * 34 qubits, complex64 fused (128 GiB): the depth-20 layered circuit with
    an FP64 norm + renormalisation at every layer (QCS_RENORM_AT_BARRIER),
    then its inverse; 1 - |<0|U^-1 U|0>|^2 bounds the accumulated FP32 error
    of both halves (a complex128 run at 34 qubits needs 256 GiB and does not
    fit beside it).
  * 30 and 34 qubits, complex64 tile programs (FP32 and split-bf16): the
    depth-20 layered circuit and its inverse in one flush each; the split
    form rounds the state to 16 significant bits per block, so its bound is
    the form's 1e-6 rather than the FP32 forms' 1e-9."""
import math

import numpy as np
import pytest

import helpers as HP
import qcs_amd
from qcs_amd import circuits as C

pytestmark = pytest.mark.gpu


def inverse(ops):
    inv = C.Ops()
    for op, t, c, p in reversed(list(zip(ops.op, ops.tq, ops.cq, ops.par))):
        if op in (C.RX, C.RY, C.RZ, C.P, C.CPHASE):
            inv.add(op, t, c, -p)
        elif op in (C.H, C.X, C.Y, C.Z, C.CNOT, C.SWAP):
            inv.add(op, t, c, p)
        else:
            raise ValueError(op)
    return inv


def uniform_index(seed, n):
    return int(C.uniforms(seed, 1)[0] * (1 << n))


def bitrev_array(y, n):
    r = np.zeros_like(y)
    for i in range(n):
        r |= ((y >> i) & 1) << (n - 1 - i)
    return r


def test_28q_layered_round_trip(qlib):
    n = 28
    print(f"[full] layered round trip n={n}", flush=True)
    ops = C.layered(n, 20, 1)
    assert len(ops) == 2248
    with qcs_amd.Circuit(n) as c:
        c.apply(ops)
        mid = c.norm_sq()
        c.apply(inverse(ops))
        a0 = c.amplitudes(0, 1 << 16)
        assert abs(mid - 1.0) < 1e-11
        assert abs(a0[0] - 1.0) < 1e-11
        assert np.max(np.abs(a0[1:])) < 1e-11
        assert abs(c.norm_sq() - 1.0) < 1e-11
        assert abs(c.get_probability_l(0) - 1.0) < 1e-11


def test_28q_qft_closed_form(qlib):
    n = 28
    b = uniform_index(1, n)
    print(f"[full] qft n={n} b={b}", flush=True)
    with qcs_amd.Circuit(n) as c:
        c.set_basis_state(b)
        c.apply(C.qft(n, swaps=True))
        N = 1 << n
        rb = HP.bitrev(b, n)
        for off in (0, 12345678, N - (1 << 16)):
            y = np.arange(off, off + (1 << 16), dtype=np.int64)
            ref = np.exp(2j * np.pi * ((rb * bitrev_array(y, n)) % N) / N) / math.sqrt(N)
            got = c.amplitudes(off, 1 << 16)
            assert np.max(np.abs(got - ref)) < 1e-13
        assert abs(c.norm_sq() - 1.0) < 1e-12


def test_33q_grover_truncated(qlib):
    n, k = 33, 64
    m = uniform_index(4, n)
    print(f"[full] grover n={n} marked={m} iterations={k}", flush=True)
    th = math.asin(1.0 / math.sqrt(2.0 ** n))
    want_m = math.sin((2 * k + 1) * th)
    want_o = math.cos((2 * k + 1) * th) / math.sqrt(2.0 ** n - 1)
    with qcs_amd.Circuit(n) as c, qcs_amd.Circuit(n, qcs_amd.PLAIN_FP32) as c32:
        c.grover_search_iters(m, k)
        c32.grover_search_iters(m, k)
        assert abs(c.get_probability_l(m) - want_m ** 2) < 1e-12
        am = c.amplitudes(m, 1)[0]
        assert abs(am - want_m) < 1e-12
        for off in (0, (m + 12345) % (1 << n), (1 << n) - 8):
            a = c.amplitudes(off, 8)
            idx = np.arange(off, off + 8)
            a = a[idx != m]
            assert np.max(np.abs(a - want_o)) < 1e-15
        assert abs(c.norm_sq() - 1.0) < 1e-10
        omf, mx = c32.compare_with(c)
        print(f"[full] grover n={n}: complex64 1-F {omf:.3e} max-abs {mx:.3e}", flush=True)
        assert omf <= 1e-6


def test_34q_fused_with_layer_renorm_round_trip(qlib):
    n, depth = 34, 20
    print(f"[full] fused layered round trip n={n}", flush=True)
    ops = C.layered(n, depth, 2)
    per = 4 * n - 1
    with qcs_amd.Circuit(n, qcs_amd.FUSED_FP32 | qcs_amd.RENORM_AT_BARRIER) as c:
        first = C.Ops()
        for i in range(n):
            first.add(ops.op[i], ops.tq[i], ops.cq[i], ops.par[i])
        c.apply(first)
        for d in range(depth):
            lay = C.Ops()
            for i in range(n + d * per, n + (d + 1) * per):
                lay.add(ops.op[i], ops.tq[i], ops.cq[i], ops.par[i])
            c.apply(lay)
            c.barrier()  # flush + FP64 norm and renormalisation
        assert abs(c.norm_sq() - 1.0) < 1e-6
        inv = inverse(ops)
        for d in range(depth):
            lay = C.Ops()
            for i in range(d * per, (d + 1) * per):
                lay.add(inv.op[i], inv.tq[i], inv.cq[i], inv.par[i])
            c.apply(lay)
            c.barrier()
        last = C.Ops()
        for i in range(depth * per, len(inv.op)):
            last.add(inv.op[i], inv.tq[i], inv.cq[i], inv.par[i])
        c.apply(last)
        p0 = c.get_probability_l(0)
        print(f"[full] fused n={n}: 1 - |<0|U^-1 U|0>|^2 = {1 - p0:.3e}, "
              f"blocks {c.get_num_fused_blocks()}", flush=True)
        assert 1.0 - p0 <= 2e-6


@pytest.mark.parametrize("n,mode", [(30, "tiles"), (30, "split"), (34, "split")])
def test_tile_programs_round_trip(qlib, n, mode):
    print(f"[full] tile programs round trip n={n} mode={mode}", flush=True)
    ops = C.layered(n, 20, 2)
    opt = qcs_amd.FUSE_TILES if mode == "tiles" else qcs_amd.FUSE_SPLIT_BF16
    with qcs_amd.Circuit(n, qcs_amd.FUSED_FP32 | opt) as c:
        c.apply(ops)
        c.flush()
        assert abs(c.norm_sq() - 1.0) < 1e-6
        c.apply(inverse(ops))
        a0 = c.amplitudes(0, 1)[0]
        # the weight outside |0...0>, summed in FP64 with a0 taken out: 1 - F
        # without the cancellation of 1 - |a0|^2 at FP32 resolution
        c.set_amplitudes(np.zeros(1, np.complex128), 0)
        rest = c.norm_sq()
        omf = rest / (rest + abs(a0) ** 2)
        print(f"[full] {mode} n={n}: 1 - |<0|U^-1 U|0>|^2 = {omf:.3e}, "
              f"blocks {c.get_num_fused_blocks()}, passes {c.get_num_fused_passes()}",
              flush=True)
        assert c.get_num_fused_passes() < c.get_num_fused_blocks()
        assert omf <= (1e-6 if mode == "split" else 1e-9)
