#!/usr/bin/env python3
"""Amplitude updates/s at 30 qubits: fused FP32 vs plain FP64, at 1-F <= 1e-6.

One step = one full run of the 30-qubit depth-20 layered circuit (H on every
qubit, then 20 layers of Rz-Ry-Rz on every qubit and a CNOT ladder; angles
from splitmix64(seed 2); 2,410 gates) on a complex64 state resident in HBM:
reset to |0...0>, record the gates through the qc_* API, flush (fusion plan,
FP64 composition of the fused unitaries on the device, tile programs that
apply several fused blocks per HBM pass on the FP32 matrix cores, FP64
renormalisation).  value = gates * 2^30 / step time.

Beside it, in the same process:
  * the plain FP64 form of the same circuit (complex128, one kernel per gate),
    for the speedup and for 1 - F / max-abs of the fused state;
  * the dominant kernel's roofline: average duration of the fused
    matrix-core kernel from HIP events recorded on the state's stream inside
    the timed steps; tile programs are bound by the FP32 matrix cores
    (algorithmic flops per launch = 8 * 2^k * 2^30 per block of the pass),
    one-block passes by HBM (2 * 2^30 * 8 bytes per launch);
  * the other fused modes on the same circuit (other_modes): one HBM pass
    per fused block (k <= 5), and split-bf16 tile programs;
  * the CPU baseline: the reference's own sequential library (oracle/_ref,
    built from its sources) on the 20-qubit depth-10 circuit, first gates
    only (a bounded sample), 1 core; beside it the reference's OpenMP build
    and the oracle's race-fixed OpenMP loop on the box's cores.

Other workloads (not the bench line): --workload batch (4096 circuits in one
persistent launch), --workload qft (the 28-qubit QFT with swaps).

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload ...]
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "amplitude updates/s at 30 qubits: fused FP32 vs plain FP64, at 1-F<=1e-6"
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0
# forms of 4-qubit blocks in FP32 tile programs (QCS_HIP_TILE_K4_*)
K4_FORMS = {"chain": 0, "halves": 1, "gauss": 2}


def parse():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--qubits", type=int, default=30)
    p.add_argument("--depth", type=int, default=20)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--fuse-k", type=int, default=0,
                   help="max qubits per fused block (default: 4 in tile programs, 5 per-block)")
    p.add_argument("--mode", choices=["tiles", "split", "blocks"], default="tiles",
                   help="tiles: several fused blocks per HBM pass; split: tiles whose "
                        "blocks run as split-bf16 products; blocks: one pass per block")
    p.add_argument("--tile-threads", type=int, default=1024)
    p.add_argument("--k4-form", choices=["chain", "halves", "gauss"], default=None,
                   help="form of 4-qubit blocks in FP32 tile programs (1024 threads): one "
                        "32x32 chain per wave, two pipelined 16-column halves, or those "
                        "halves with three real products per complex product; default: "
                        "the autotuner's choice (gauss with --no-tune)")
    p.add_argument("--no-plain", action="store_true", help="skip the plain FP64 run")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-alt", action="store_true",
                   help="skip the timing of the other fused modes")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--workload", choices=["fused30", "batch", "qft", "observe", "config5", "cpu"],
                   default="fused30",
                   help="fused30: the 30-qubit fused-vs-plain line (default); batch: "
                        "4096 independent 10-20 qubit circuits in one persistent launch; "
                        "qft: the 28-qubit QFT with swaps, plain complex128 and fused; "
                        "observe: the observation kernels on a 30-qubit complex128 state; "
                        "config5: 33-qubit Grover (complex128) and the 34-qubit fused circuit; "
                        "cpu: config 1 in full on every CPU build of the reference (no GPU)")
    p.add_argument("--batch-count", type=int, default=4096)
    p.add_argument("--qft-qubits", type=int, default=28,
                   help="qubits of the QFT workload (default 28)")
    p.add_argument("--obs-qubits", type=int, default=30,
                   help="qubits of the observe workload (default 30)")
    p.add_argument("--batch-threads", type=int, default=512)
    p.add_argument("--no-tune", action="store_true",
                   help="skip the autotuner (default plans and kernel variants)")
    p.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                   help="only run a plain device copy of GIB GiB (float32, out = in * 1) "
                        "three times: the known-size kernel that scales HBM byte "
                        "counters (FETCH_SIZE, WRITE_SIZE)")
    return p.parse_args()


def copy_probe(gib):
    import torch
    n = int(gib * 2**30) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n,
                      "launches": 3}), flush=True)


def timing(hip, state, cls):
    import ctypes
    ms, n, b, f = (ctypes.c_double(), ctypes.c_long(), ctypes.c_double(),
                   ctypes.c_double())
    rc = hip.qcs_hip_timing_read(state, cls, ctypes.byref(ms), ctypes.byref(n),
                                 ctypes.byref(b), ctypes.byref(f))
    if rc != 0:
        raise RuntimeError(hip.qcs_hip_last_error())
    return ms.value, n.value, b.value, f.value


def moved(hip, state, cls):
    """Bytes of the 128-byte lines the class's launches read and wrote."""
    import ctypes
    mv = ctypes.c_double()
    if hip.qcs_hip_timing_moved(state, cls, ctypes.byref(mv)) != 0:
        raise RuntimeError(hip.qcs_hip_last_error())
    return mv.value


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(seconds):
    """Reference sequential library on the 20-qubit depth-10 circuit, then its
    OpenMP build and the race-fixed OpenMP restatement."""
    from qcs_amd import circuits as C
    n = 20
    ops = C.layered(n, 10, 0)
    try:
        from oracle.oracle import RefLib
        ref = RefLib("libqcsref.so")
        kind = "reference"
        what = ("reference sequential build (its sources, -std=c89 -O3 "
                "-march=x86-64-v3 -mavx2 -mfma)")
        c = ref.create(n)
        apply_one = lambda i: ref.apply(c, _slice(ops, i))  # noqa: E731
        close = lambda: ref.L.qc_destroy(c)  # noqa: E731
    except (FileNotFoundError, OSError):
        from oracle.oracle import OracleState
        st = OracleState(n, compat=True)
        kind = "port"
        what = "C restatement of the reference's sequential loop (oracle)"
        apply_one = lambda i: st.run(_slice(ops, i))  # noqa: E731
        close = st.close
    done = 0
    t0 = time.perf_counter()
    while done < len(ops) and time.perf_counter() - t0 < 0.5 * seconds:
        apply_one(done)
        done += 1
    dt = time.perf_counter() - t0
    close()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    omp = {}
    try:
        # the reference's OpenMP build, as written (its shared loop index races)
        from oracle.oracle import RefLib
        refo = RefLib("libqcsref_omp.so")
        co = refo.create(n)
        k = 0
        t1 = time.perf_counter()
        while k < len(ops) and time.perf_counter() - t1 < 0.25 * seconds:
            refo.apply(co, _slice(ops, k))
            k += 1
        d1 = time.perf_counter() - t1
        refo.L.qc_destroy(co)
        omp["openmp_reference"] = {
            "value": k * float(1 << n) / d1, "unit": "amplitude updates/s",
            "cores": threads, "kind": "reference",
            "sample": f"first {k} gates, {d1:.1f} s, reference OpenMP build (racy as written)"}
    except (FileNotFoundError, OSError):
        pass
    # the oracle's OpenMP restatement of the 1-qubit loop with the race fixed
    from oracle.oracle import OracleState, gate as ogate
    st = OracleState(n)
    k = 0
    t2 = time.perf_counter()
    one_q = [i for i in range(len(ops)) if ops.op[i] < C.CNOT]
    while k < len(one_q) and time.perf_counter() - t2 < 0.25 * seconds:
        i = one_q[k]
        m = ogate(ops.op[i], ops.par[i])
        st.L.oq_apply_1q_omp(st.s, int(ops.tq[i]), m.ctypes.data_as(
            __import__("ctypes").POINTER(__import__("ctypes").c_double)), 0)
        k += 1
    d2 = time.perf_counter() - t2
    st.close()
    omp["openmp_race_fixed"] = {
        "value": k * float(1 << n) / d2, "unit": "amplitude updates/s",
        "cores": threads, "kind": "port",
        "sample": f"first {k} one-qubit gates, {d2:.1f} s, OpenMP restatement with a private index"}
    return {
        "value": done * float(1 << n) / dt,
        "unit": "amplitude updates/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"first {done} of {len(ops)} gates of the 20-qubit depth-10 "
                   f"layered circuit (seed 0), {dt:.1f} s, {what}"),
        "gates_per_s": done / dt,
        "cpu": cpu_model(),
        "nproc": os.cpu_count(),
        **omp,
    }


def _ref_leg(ref, n, ops, shots, budget_s, what, cores):
    """One reference build on config 1: apply the gates one call each until
    done or `budget_s` passed, then (when all ran) qc_run_shots after
    srand(0).  Amp-updates/s = gates * 2^n / time (the GPU's definition)."""
    import ctypes
    import qcs_amd
    c = ref.create(n)
    done, t0 = 0, time.perf_counter()
    while done < len(ops) and time.perf_counter() - t0 < budget_s:
        ref.apply(c, _slice(ops, done))
        done += 1
    dt = time.perf_counter() - t0
    leg = {"value": done * float(1 << n) / dt, "unit": "amplitude updates/s",
           "cores": cores, "kind": "reference", "gates_per_s": done / dt,
           "gates_run": done, "seconds": dt, "build": what}
    if done == len(ops):
        hist = (ctypes.c_int * (1 << n))()
        qcs_amd.libc().srand(0)
        t1 = time.perf_counter()
        ref.L.qc_run_shots(c, shots, hist)
        ts = time.perf_counter() - t1
        leg["shots_s"] = ts
        leg["shots_per_s"] = shots / ts
        leg["shots_landed"] = int(sum(hist))
        leg["sample"] = (f"config 1 in full: {len(ops)} gates in {dt:.1f} s, then {shots} "
                         f"shots after srand(0) in {ts:.2f} s ({leg['shots_landed']} landed)")
    else:
        leg["sample"] = (f"first {done} of {len(ops)} gates of config 1 within {dt:.1f} s "
                         "(the full circuit would take minutes)")
    ref.L.qc_destroy(c)
    return leg


def _omp_threads(k):
    """threads of every OpenMP loop in this process (one libgomp instance)"""
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(k))
    except OSError:
        pass


def bench_cpu(args):
    """Config 1 on the host cores (not a GPU line): the 20-qubit depth-10
    layered circuit, seed 0, 810 gates, then 1024 shots after srand(0), with
    every CPU build the reference's path has, each a leg of its own:
      * the reference's sequential build at its own flags (-std=c89 -O3
        -march=x86-64-v3 -mavx2 -mfma) and at -O2, 1 core, in full;
      * the reference's OpenMP build (as written: the 1q loop's shared index
        and the 2q loop's double write race) at 1 thread in full and at
        nproc threads on a bounded sample (its scaling is negative);
      * the oracle's race-fixed OpenMP restatement of both loops, textbook and
        as-written semantics, at 1 thread and at nproc, in full.
    The reference builds are oracle/_ref (compiled from its sources); without
    them only the oracle legs run."""
    from qcs_amd import circuits as C
    from oracle.oracle import OracleState
    import qcs_amd
    import numpy as np
    n, depth, shots = 20, 10, 1024
    ops = C.layered(n, depth, 0)
    nproc = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    legs = {}
    for lib, what, budget in (("libqcsref.so", "sequential, -O3 -march=x86-64-v3 -mavx2 -mfma", 120),
                              ("libqcsref_O2.so", "sequential, -O2", 120)):
        try:
            from oracle.oracle import RefLib
            ref = RefLib(lib)
        except (FileNotFoundError, OSError):
            continue
        key = "reference_" + ("O3" if "O2" not in lib else "O2")
        legs[key] = _ref_leg(ref, n, ops, shots, budget, what, 1)
        print(f"bench: cpu {key} {legs[key]['sample']}", file=sys.stderr, flush=True)
    try:
        from oracle.oracle import RefLib
        refo = RefLib("libqcsref_omp.so")
        for k, budget in ((1, 120), (nproc, 15)):
            _omp_threads(k)
            key = f"reference_openmp_{k}_threads"
            legs[key] = _ref_leg(refo, n, ops, shots, budget,
                                 "OpenMP build as written (racy), -O3", k)
            print(f"bench: cpu {key} {legs[key]['sample']}", file=sys.stderr, flush=True)
    except (FileNotFoundError, OSError):
        pass
    cl = qcs_amd.libc()
    for compat in (False, True):
        for k in (1, nproc):
            st = OracleState(n, compat=compat)
            st.L.oq_set_threads(k)
            t0 = time.perf_counter()
            st.run_omp(ops, racy=False)
            dt = time.perf_counter() - t0
            cl.srand(0)
            u = np.array([cl.rand() / 2147483647.0 for _ in range(shots)])
            t1 = time.perf_counter()
            idx = st.sample(u)
            ts = time.perf_counter() - t1
            st.close()
            sem = "as_written" if compat else "textbook"
            key = f"oracle_openmp_race_fixed_{sem}_{k}_threads"
            legs[key] = {"value": len(ops) * float(1 << n) / dt, "unit": "amplitude updates/s",
                         "cores": k, "kind": "port", "gates_per_s": len(ops) / dt,
                         "seconds": dt, "shots_s": ts, "shots_landed": int(np.sum(idx >= 0)),
                         "sample": f"config 1 in full ({len(ops)} gates, {shots} shots), "
                                   f"race-fixed OpenMP loops, {sem} controlled gate"}
            print(f"bench: cpu {key} {dt:.2f} s", file=sys.stderr, flush=True)
    _omp_threads(nproc)
    head = legs.get("reference_O3") or next(iter(legs.values()))
    print(json.dumps({
        "metric": "amplitude updates/s, config 1 (20 qubits, depth 10, 810 gates) on host cores",
        "value": head["value"], "unit": "amplitude updates/s",
        "n_gpus": 0, "steps": 1, "warmup": 0,
        "ms_per_step": head.get("seconds", 0.0) * 1e3, "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "20-qubit layered Rz-Ry-Rz + CNOT-ladder circuit, depth 10, "
                               "seed 0 (810 gates), then 1024 shots after srand(0)",
                   "qubits": n, "gates": len(ops), "shots": shots},
        "cpu": cpu_model(), "nproc": os.cpu_count(), "omp_threads": nproc,
        "gcc": _tool_version(["gcc", "--version"]),
        "glibc": platform.libc_ver()[1],
        "legs": legs,
    }), flush=True)


def _tool_version(cmd):
    try:
        return subprocess.run(cmd, capture_output=True, text=True,
                              timeout=10).stdout.splitlines()[0]
    except (OSError, subprocess.SubprocessError, IndexError):
        return None


def _slice(ops, i):
    from qcs_amd import circuits as C
    o = C.Ops()
    o.add(ops.op[i], ops.tq[i], ops.cq[i], ops.par[i])
    return o


class _Arrays:
    def __init__(self, a):
        self.op, self.tq, self.cq, self.par = (a[0].tolist(), a[1].tolist(),
                                               a[2].tolist(), a[3].tolist())


def _reflib():
    """the reference's sequential library built from its sources (oracle/_ref),
    or None where it was not built"""
    try:
        from oracle.oracle import RefLib
        return RefLib("libqcsref.so")
    except (FileNotFoundError, OSError, ImportError):
        return None


def ref_qft_baseline(n=18):
    """The reference's qc_quantum_fourier_transform (no swaps; its CPHASE as
    written) on a fresh n-qubit circuit, one call, 1 core."""
    ref = _reflib()
    if ref is None:
        return None
    c = ref.create(n)
    t0 = time.perf_counter()
    ref.L.qc_quantum_fourier_transform(c)
    dt = time.perf_counter() - t0
    g = ref.L.qc_get_num_gates(c)
    ref.L.qc_destroy(c)
    return {"value": g * float(1 << n) / dt, "unit": "amplitude updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"qc_quantum_fourier_transform on {n} qubits ({g} gates, no swaps), "
                      f"{dt:.1f} s, reference sequential build",
            "gates_per_s": g / dt, "cpu": cpu_model()}


def ref_observe_baseline(n=20, shots=1024, grover_n=16):
    """The reference's observation calls on an n-qubit state (H on every
    qubit), 1 core, as GB/s of the same algorithmic bytes as the GPU line."""
    ref = _reflib()
    if ref is None:
        return None
    import ctypes
    s = 16
    N = float(1 << n)
    c = ref.create(n)
    for q in range(n):
        ref.L.qc_h(c, q)
    out = {}

    def timed(name, fn, nbytes):
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        out[name] = {"ms": dt * 1e3, "gbs": nbytes / dt / 1e9}

    hist = (ctypes.c_int * (1 << n))()
    timed("argmax", lambda: ref.L.qc_find_most_likely_state(c), N * s)
    timed(f"sample_{shots}_shots", lambda: ref.L.qc_run_shots(c, shots, hist),
          N * s + 2 * N * 8 + shots * n * 8)
    timed("measure_qubit_7", lambda: ref.L.qc_measure(c, 7), 2 * N * s)
    ref.L.qc_destroy(c)
    g = ref.create(grover_n)
    t0 = time.perf_counter()
    ref.L.qc_grover_search(g, 12345 % (1 << grover_n))
    dt = time.perf_counter() - t0
    # the reference's iteration count, floor(pi/4 * sqrt(2^n)) (src/q_utils.c:126-134)
    rounds = int(math.pi / 4.0 * math.sqrt(float(1 << grover_n)))
    out[f"grover_round_{grover_n}q"] = {
        "ms": dt / rounds * 1e3, "gbs": 3 * float(1 << grover_n) * s * rounds / dt / 1e9}
    ref.L.qc_destroy(g)
    return {"value": out["argmax"]["gbs"], "unit": "GB/s", "cores": 1, "kind": "reference",
            "sample": f"{n}-qubit state (Grover: {grover_n} qubits, {rounds} rounds), one call each, "
                      "reference sequential build",
            "ops": out, "cpu": cpu_model()}


def bench_qft(args):
    """The 28-qubit QFT with final swaps (src/qcs.c:443-457 order: H(i), then
    CPHASE(j -> i, pi/2^(j-i)) for j > i; 406 + 14 gates) on a basis state
    |b>, b uniform from splitmix64 seed 1.  A step is: set |b>, apply the
    gates, wait.  Plain complex128 (value) and fused FP32 tile programs beside
    it; one 65536-amplitude range is checked against the closed form
    exp(2 pi i rev(b) rev(y) / N) / sqrt(N)."""
    import math
    import numpy as np
    import qcs_amd
    from qcs_amd import circuits as C

    hip, _ = qcs_amd.load()
    if not hip.qcs_hip_device_available():
        raise SystemExit("no HIP device")
    n = args.qft_qubits
    N = 1 << n
    b = int(C.uniforms(1, 1)[0] * N)
    ops = C.qft(n, swaps=True)
    G = len(ops)

    def rev(x):
        r = np.zeros_like(x)
        for i in range(n):
            r |= ((x >> i) & 1) << (n - 1 - i)
        return r

    def run(options):
        c = qcs_amd.Circuit(n, options)
        st = c.device_state
        for _ in range(args.warmup):
            c.set_basis_state(b)
            c.apply(ops)
            c.flush()
        hip.qcs_hip_timing_enable(st, 1)
        hip.qcs_hip_device_sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            c.set_basis_state(b)
            c.apply(ops)
            c.flush()
        return c, (time.perf_counter() - t0) / args.steps

    plain, tp = run(qcs_amd.PLAIN_FP64)
    st = plain.device_state
    m1, n1, b1, _ = timing(hip, st, qcs_amd.K_1Q)
    m2, n2, b2, _ = timing(hip, st, qcs_amd.K_CU)
    off = 12345678 % (N - (1 << 16))
    y = np.arange(off, off + (1 << 16), dtype=np.int64)
    ref = np.exp(2j * np.pi * ((int(rev(np.array([b]))[0]) * rev(y)) % N) / N) / math.sqrt(N)
    err = float(np.max(np.abs(plain.amplitudes(off, 1 << 16) - ref)))
    fused, tf = run(qcs_amd.FUSED_FP32 | qcs_amd.FUSE_TILES)
    omf, mx = fused.compare_with(plain)
    runs = args.warmup + args.steps  # the counters add up over every run
    blocks = fused.get_num_fused_blocks() // runs
    passes = fused.get_num_fused_passes() // runs
    fused.close()
    plain.close()
    print(f"bench: qft {n} q plain {tp * 1e3:.1f} ms, fused {tf * 1e3:.1f} ms per step",
          file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": f"amplitude updates/s, {n}-qubit QFT with swaps, complex128 plain",
        "value": G * N / tp,
        "unit": "amplitude updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": tp * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}-qubit QFT (reference gate order) with final swaps on "
                               f"|b>, b = {b} (splitmix64 seed 1)",
                   "qubits": n, "gates": G},
        "gates_per_s": G / tp,
        "closed_form_max_abs_err": err,
        "k_1q": {"launches": n1, "avg_ms": m1 / max(n1, 1),
                 "gbs": (b1 / max(n1, 1)) / (m1 / max(n1, 1) * 1e-3) / 1e9 if n1 else None},
        "k_cu_and_swap": {"launches": n2, "avg_ms": m2 / max(n2, 1),
                          "gbs_algorithmic": (b2 / max(n2, 1)) / (m2 / max(n2, 1) * 1e-3) / 1e9
                          if n2 else None},
        "fused_fp32_tiles": {"ms_per_step": tf * 1e3, "amp_updates_per_s": G * N / tf,
                             "fused_blocks": blocks, "hbm_passes": passes,
                             "one_minus_f": omf, "max_abs_err": mx},
        "cpu_baseline": None if args.no_cpu else ref_qft_baseline(),
    }), flush=True)


def bench_observe(args):
    """The observation side of the plain form on a 30-qubit complex128 state
    (16 GiB; H on every qubit, then one layered Rz-Ry-Rz + CNOT layer): norm
    reduction, renormalisation, argmax, one-qubit measurement, 1024-shot
    sampling and Grover rounds, each through the C ABI, timed with a device
    synchronisation around it (median of --steps runs), against the
    algorithmic bytes of each operation: N*s read for a reduction, 3*N*s for
    reduce + rescale, 2*N*s for a measurement (p0 on one half, collapse and
    rescale), N*s + 2*N*8 + shots*log2(N)*8 for sampling, 3*N*s per textbook
    Grover round (its phase flip touches one amplitude).  moved_bytes is what
    this implementation reads and writes where that differs: sampling never
    writes the CDF (chunk sums, then one 4096-amplitude chunk per shot).  A
    measurement of qubit 7 reads the bit-7 = 0 half for p0 (and the other half
    for p1 only on outcome 1), then reads the kept half and writes the whole
    state: 2*N*s moved for outcome 0, 2.5*N*s for outcome 1.  Before each
    timed measurement an untimed H on qubit 7 makes p0 = p1 = 1/2, and
    u = 0.25 or 0.75 picks the outcome; both are reported, and their mean.
    Two gate kernels that the other
    workloads' circuits do not reach are timed here too: the phase gate
    (N*s: half the state read and written) and, on a reference-compatible
    circuit, the out-of-place controlled gate (2*N*s)."""
    import ctypes
    import statistics
    import numpy as np
    import qcs_amd
    from qcs_amd import circuits as C

    hip, _ = qcs_amd.load()
    if not hip.qcs_hip_device_available():
        raise SystemExit("no HIP device")
    n, s = args.obs_qubits, 16
    N = float(1 << n)
    c = qcs_amd.Circuit(n)
    c.apply(C.layered(n, 1, args.seed))
    st = c.device_state
    d = ctypes.c_double()
    l = ctypes.c_long()
    o = ctypes.c_int()
    shots = 1024
    u = (ctypes.c_double * shots)(*np.random.default_rng(0).random(shots))
    idx = (ctypes.c_long * shots)()

    def ok(rc):
        if rc != 0:
            raise RuntimeError(hip.qcs_hip_last_error())

    amp = (ctypes.c_double * 2)()

    def perturb():
        # one amplitude changed (untimed), so the norm is not 1 and
        # normalize rescales the state
        hip.qcs_hip_read(st, 0, 1, amp)
        amp[0] *= 1.5
        ok(hip.qcs_hip_write(st, 0, 1, amp))

    r2 = math.sqrt(0.5)
    hm = (ctypes.c_double * 8)(r2, 0.0, r2, 0.0, r2, 0.0, -r2, 0.0)

    def balance_q7():
        # after a measurement qubit 7 is in a basis state: H makes both
        # outcomes equally likely again
        ok(hip.qcs_hip_apply_1q(st, 7, hm))

    # name: (call, algorithmic bytes, bytes this implementation moves, setup)
    chunk = 4096 * s  # a shot scans one 4096-amplitude chunk
    ops = {
        "norm_sq": (lambda: ok(hip.qcs_hip_norm_sq(st, ctypes.byref(d))), N * s, N * s, None),
        "normalize": (lambda: ok(hip.qcs_hip_normalize(st, None)), 3 * N * s, 3 * N * s,
                      perturb),
        "argmax": (lambda: ok(hip.qcs_hip_argmax(st, ctypes.byref(l))), N * s, N * s, None),
        "measure_qubit_7": (lambda: ok(hip.qcs_hip_measure(st, 7, 0.25, ctypes.byref(o))),
                            2 * N * s, 2 * N * s, balance_q7),
        "measure_qubit_7_outcome_1": (
            lambda: ok(hip.qcs_hip_measure(st, 7, 0.75, ctypes.byref(o))),
            2 * N * s, 2.5 * N * s, balance_q7),
        "sample_1024_shots": (lambda: ok(hip.qcs_hip_sample(st, shots, u, idx)),
                              N * s + 2 * N * 8 + shots * n * 8, N * s + shots * chunk, None),
        "grover_round": (lambda: ok(hip.qcs_hip_grover(st, 12345, 1)), 3 * N * s, 3 * N * s,
                         None),
        "probability": (lambda: ok(hip.qcs_hip_probability(st, 12345, ctypes.byref(d))),
                        2 * s, 2 * s, None),
        # the phase gate P(0.3) on qubit 7 (k_1q_phase: reads and writes only
        # the bit-7 = 1 half)
        "phase_gate_q7": (lambda: ok(hip.qcs_hip_apply_1q(st, 7, pm)), N * s, N * s, None),
    }
    pm = (ctypes.c_double * 8)(1.0, 0.0, 0.0, 0.0, 0.0, 0.0, math.cos(0.3), math.sin(0.3))
    out = {}

    def time_ops(state, table):
        for name, (fn, nbytes, moved, setup) in table.items():
            times = []
            for i in range(args.warmup + args.steps):
                if setup:
                    setup()
                hip.qcs_hip_device_sync()
                t0 = time.perf_counter()
                fn()
                ok(hip.qcs_hip_sync(state))
                if i >= args.warmup:
                    times.append(time.perf_counter() - t0)
            t = statistics.median(times)
            out[name] = {"ms": t * 1e3, "algorithmic_bytes": nbytes,
                         "gbs": nbytes / t / 1e9, "hbm_frac": nbytes / t / 1e9 / HBM_PEAK_GBS,
                         "moved_bytes": moved, "moved_gbs": moved / t / 1e9}
            print(f"bench: observe {name} {t * 1e3:.3f} ms", file=sys.stderr, flush=True)

    time_ops(st, ops)
    m0, m1 = out["measure_qubit_7"], out["measure_qubit_7_outcome_1"]
    out["measure_qubit_7_mean_of_outcomes"] = {
        "ms": (m0["ms"] + m1["ms"]) / 2, "algorithmic_bytes": 2 * N * s,
        "gbs": 2 * N * s / ((m0["ms"] + m1["ms"]) / 2e3) / 1e9,
        "moved_bytes": 2.25 * N * s}
    c.close()
    # the reference-compatible controlled gate (out of place over the whole
    # state, then a buffer swap: 2*N*s), on a compat circuit (two buffers)
    cc = qcs_amd.Circuit(n, qcs_amd.REFERENCE_COMPAT)
    cst = cc.device_state
    xm = (ctypes.c_double * 8)(0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0)
    time_ops(cst, {"cnot_compat_3_9": (lambda: ok(hip.qcs_hip_apply_cu(cst, 3, 9, xm)),
                                       2 * N * s, 2 * N * s, None)})
    cc.close()
    print(json.dumps({
        "metric": f"observation kernels on a {n}-qubit complex128 state (GB/s of algorithmic bytes)",
        "value": out["norm_sq"]["gbs"], "unit": "GB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": out["norm_sq"]["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}-qubit complex128 state after H on every qubit and one "
                               f"layered layer (seed {args.seed}); one call per timed run",
                   "qubits": n},
        "ops": out,
        "cpu_baseline": None if args.no_cpu else ref_observe_baseline(),
    }), flush=True)


def bench_config5(args):
    """The largest single-GPU states, one timed run each (after one warmup):
    (a) 33 qubits complex128 (128 GiB): Grover for the marked index from
    splitmix64 seed 4, truncated to 64 rounds, checked against the analytic
    amplitude sin((2k+1) asin(2^-n/2)); (b) 34 qubits complex64 (128 GiB),
    freed (a) first: the depth-20 layered circuit (2,734 gates) as fused FP32
    tile programs, one flush."""
    import math
    import qcs_amd
    from qcs_amd import circuits as C

    hip, _ = qcs_amd.load()
    if not hip.qcs_hip_device_available():
        raise SystemExit("no HIP device")
    out = {}
    n, k = 33, 64
    m = int(C.uniforms(4, 1)[0] * (1 << n))
    want = math.sin((2 * k + 1) * math.asin(1.0 / math.sqrt(2.0 ** n))) ** 2
    with qcs_amd.Circuit(n) as g:
        st = g.device_state
        for _ in range(2):  # warmup, then the timed run
            g.set_basis_state(0)
            hip.qcs_hip_device_sync()
            t0 = time.perf_counter()
            for q in range(n):
                g.h(q)
            g.flush()
            t_h = time.perf_counter() - t0
            t0 = time.perf_counter()
            # the rounds of qc_grover_search_iters, on the device
            if hip.qcs_hip_grover(st, m, k) != 0 or hip.qcs_hip_sync(st) != 0:
                raise RuntimeError(hip.qcs_hip_last_error())
            t_r = time.perf_counter() - t0
        p = g.get_probability_l(m)
        N, sb = float(1 << n), 16.0
        out["grover_33q_c128"] = {
            "h_layer_ms": t_h * 1e3, "rounds": k, "rounds_ms": t_r * 1e3,
            "ms_per_round": t_r / k * 1e3,
            # the reference's round: a sum, then a read+write update (3*N*s);
            # here the update also forms the next round's sum (2*N*s moved)
            "round_gbs": 3 * N * sb / (t_r / k) / 1e9,
            "round_moved_gbs": 2 * N * sb / (t_r / k) / 1e9,
            "round_moved_hbm_frac": 2 * N * sb / (t_r / k) / 1e9 / HBM_PEAK_GBS,
            "marked": m, "p_marked": p, "p_marked_analytic": want,
            "abs_err": abs(p - want)}
    t = t_h + t_r
    print(f"bench: config5 grover 33 q {t * 1e3:.0f} ms", file=sys.stderr, flush=True)
    n, depth = 34, 20
    ops = C.layered(n, depth, 2)
    G = len(ops)
    with qcs_amd.Circuit(n, qcs_amd.FUSED_FP32 | qcs_amd.FUSE_TILES) as f:
        times = []
        for _ in range(2):
            f.set_basis_state(0)
            hip.qcs_hip_device_sync()
            t0 = time.perf_counter()
            f.apply(ops)
            f.flush()
            times.append(time.perf_counter() - t0)
        t = times[-1]
        out["fused_34q_tiles"] = {
            "ms": t * 1e3, "gates": G, "amp_updates_per_s": G * float(1 << n) / t,
            "fused_blocks": f.get_num_fused_blocks() // 2,
            "hbm_passes": f.get_num_fused_passes() // 2, "norm_sq": f.norm_sq()}
    print(f"bench: config5 fused 34 q {t * 1e3:.0f} ms", file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "33-qubit complex128 Grover and 34-qubit fused FP32 (largest single-GPU states)",
        "value": out["fused_34q_tiles"]["amp_updates_per_s"], "unit": "amplitude updates/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": out["fused_34q_tiles"]["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (Grover), f32 (fused)",
        "data": "synthetic",
        "config": {"workload": "33 q Grover, marked from splitmix64 seed 4, 64 rounds; 34 q "
                               "layered depth 20 (seed 2) as fused FP32 tile programs"},
        **out}), flush=True)


def bench_batch(args):
    """4096 independent layered circuits (n ~ U[10,20], depth ~ U[5,40],
    splitmix64 seed 3; even circuits complex128 plain, odd fused complex64),
    256 shots each, one qc_batch_sample call per step (host planning, one
    persistent device launch replayed from a hipGraph after the first step,
    sampled indices back on the host).  Timed steps plan from scratch
    (QCS_BATCH_REPLAN); the reuse of an unchanged batch's plan (new uniforms
    only) is timed beside them as rerun_ms_per_step."""
    import qcs_amd
    from qcs_amd import circuits as C

    hip, _ = qcs_amd.load()
    if not hip.qcs_hip_device_available():
        raise SystemExit("no HIP device")
    shots = 256
    work = C.batch_workload(args.batch_count, 3)
    t0 = time.perf_counter()
    circs = []
    amp_updates = {False: 0.0, True: 0.0}
    gates = 0
    for n, d, seed, fused in work:
        c = qcs_amd.Circuit(n, qcs_amd.DEFERRED | (qcs_amd.FUSED_FP32 if fused else 0))
        c.apply(_Arrays(C.layered_arrays(n, d, seed)))
        g = C.gate_count_layered(n, d)
        gates += g
        amp_updates[fused] += g * float(1 << n)
        circs.append(c)
    t_record = time.perf_counter() - t0
    print(f"batch: recorded {len(circs)} circuits, {gates} gates in {t_record:.1f} s",
          file=sys.stderr, flush=True)
    cl = qcs_amd.libc()
    cl.srand(3)
    # one run without a graph times the batch kernel alone (HIP events)
    b0 = qcs_amd.Batch()
    b0.set_threads(args.batch_threads)
    if not args.no_tune:
        b0.autotune(circs)
    t0 = time.perf_counter()
    b0.sample(circs, shots)
    first_s = time.perf_counter() - t0
    kernel_ms, first_dev_ms, _ = b0.last_timing()
    b0.close()
    print(f"batch: first run {first_s:.2f} s, device {first_dev_ms:.1f} ms, "
          f"kernel {kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    b = qcs_amd.Batch(qcs_amd.BATCH_GRAPH | qcs_amd.BATCH_REPLAN)
    b.set_threads(args.batch_threads)
    if not args.no_tune:
        t0 = time.perf_counter()
        b.autotune(circs)
        print(f"batch: autotuned in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
              flush=True)
    for _ in range(max(args.warmup, 1)):
        b.sample(circs, shots)
    hip.qcs_hip_device_sync()
    totals = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        b.sample(circs, shots)
        totals.append(b.last_timing()[1])
    hip.qcs_hip_device_sync()
    T = (time.perf_counter() - t1) / args.steps
    b.close()
    # the same circuits sampled again with the plan of the previous run
    br = qcs_amd.Batch(qcs_amd.BATCH_GRAPH)
    br.set_threads(args.batch_threads)
    if not args.no_tune:
        br.autotune(circs)
    for _ in range(2):
        br.sample(circs, shots)
    hip.qcs_hip_device_sync()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        br.sample(circs, shots)
    hip.qcs_hip_device_sync()
    T_rerun = (time.perf_counter() - t1) / args.steps
    br.close()
    print(f"batch: {T * 1e3:.1f} ms per planned step, {T_rerun * 1e3:.1f} ms per rerun",
          file=sys.stderr, flush=True)
    total_amp = amp_updates[False] + amp_updates[True]
    dev_ms = sum(totals) / len(totals)
    # the complex128 part alone (it sets the launch time), for its roofline
    c128 = [c for c, w in zip(circs, work) if not w[3]]
    bp = qcs_amd.Batch()
    bp.set_threads(args.batch_threads)
    if not args.no_tune:
        bp.autotune(c128)
    cl.srand(3)
    bp.sample(c128, shots)
    c128_kernel_ms = bp.last_timing()[0]
    bp.close()
    # FP64 VALU floor of the complex128 circuits: every amplitude of a gate's
    # sweep takes 6 FP64 operations for H, RY (real matrices: 4 multiplies, 2
    # adds) and RZ (diagonal: one complex multiply), none for CNOT (an
    # exchange); at one operation per lane per clock, 64 per CU per clock
    # (an FMA would count 2 flops: 78.6 TFLOP/s), i.e. 39.3e12 per second
    valu_peak = 64.0 * 256 * 2.4e9
    ops64 = sum((n + 3 * n * d) * 6.0 * float(1 << n) for n, d, _, f in work if not f)
    floor_ms = ops64 / valu_peak * 1e3
    print(f"batch: complex128 part alone {c128_kernel_ms:.1f} ms (floor {floor_ms:.1f} ms)",
          file=sys.stderr, flush=True)
    parity = batch_parity(circs, work, shots, args)
    print(f"batch: parity {parity}", file=sys.stderr, flush=True)
    result = {
        "metric": "amplitude updates/s, 4096 batched circuits (10-20 qubits), one persistent launch",
        "value": total_amp / T,
        "unit": "amplitude updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": T * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": ("4096 independent layered Rz-Ry-Rz + CNOT-ladder circuits, "
                         "n ~ U[10,20], depth ~ U[5,40] (splitmix64 seed 3), even "
                         "circuits complex128 plain, odd fused complex64, 256 shots each"),
            "circuits": len(circs),
            "gates": gates,
            "shots_per_circuit": shots,
            "threads_per_workgroup": args.batch_threads,
        },
        "device_ms_per_step": dev_ms,
        "host_ms_per_step": T * 1e3 - dev_ms,
        "rerun_ms_per_step": T_rerun * 1e3,
        "kernel_ms": kernel_ms,
        "device_amp_updates_per_s_kernel": total_amp / (kernel_ms * 1e-3),
        "device_amp_updates_per_s": total_amp / (dev_ms * 1e-3),
        "amp_updates_c128": amp_updates[False],
        "amp_updates_c64": amp_updates[True],
        "circuits_per_s": len(circs) / T,
        "record_s": t_record,
        "roofline": {
            "bound": "valu", "kernel": "k_batch, the complex128 circuits alone (FP64 VALU)",
            "achieved": c128_kernel_ms, "peak": floor_ms, "unit": "ms",
            "frac": floor_ms / c128_kernel_ms, "traffic": None,
            "fp64_ops": ops64, "fp64_ops_peak_per_s": valu_peak,
            "ops_per_amplitude": {"H": 6, "RY": 6, "RZ": 6, "CNOT": 0},
        },
        "parity": parity,
        "cpu_baseline": None if args.no_cpu else ref_batch_baseline(work, shots, args.cpu_seconds),
    }
    for c in circs:
        c.close()
    print(json.dumps(result), flush=True)


def batch_parity(circs, work, shots, args):
    """One more run with every final state kept: each fused complex64
    circuit against a complex128 twin batch of the same circuits and stream
    (1 - F, and the samples that differ: an FP32 state moves a uniform
    across a CDF boundary only when it lies within the two CDFs' difference
    of it), and a spread sample of the complex128 circuits against circuits
    of their own run gate by gate (amplitudes and samples bit for bit)."""
    import numpy as np
    import qcs_amd
    from qcs_amd import circuits as C
    cl = qcs_amd.libc()
    b = qcs_amd.Batch(qcs_amd.BATCH_KEEP_STATES)
    b.set_threads(args.batch_threads)
    cl.srand(3)
    idx = b.sample(circs, shots)
    twins = []
    for n, d, seed, fused in work:
        c = qcs_amd.Circuit(n, qcs_amd.DEFERRED)
        c.apply(_Arrays(C.layered_arrays(n, d, seed)))
        twins.append(c)
    t = qcs_amd.Batch(qcs_amd.BATCH_KEEP_STATES)
    cl.srand(3)
    tidx = t.sample(twins, shots)
    worst, differ, total = 0.0, 0, 0
    for i, (n, d, seed, fused) in enumerate(work):
        if not fused:
            continue
        a, r = b.amplitudes(i, n), t.amplitudes(i, n)
        f = 1.0 - abs(np.vdot(a, r)) ** 2 / (np.vdot(a, a).real * np.vdot(r, r).real)
        worst = max(worst, f)
        differ += int(np.sum(idx[i] != tidx[i]))
        total += shots
    # every 64th complex128 circuit (and the largest), run on a state of its own
    plain = [i for i, w in enumerate(work) if not w[3]]
    pick = sorted(set(plain[::64]) | {max(plain, key=lambda i: (work[i][0], work[i][1]))})
    checked, equal = 0, 0
    cl.srand(3)
    for i, (n, d, seed, fused) in enumerate(work):
        if i not in pick:
            for _ in range(shots):
                cl.rand()
            continue
        with qcs_amd.Circuit(n) as c:
            c.apply(_Arrays(C.layered_arrays(n, d, seed)))
            own = c.sample_indices(shots)
            ok = np.array_equal(c.amplitudes(), b.amplitudes(i, n)) and \
                np.array_equal(own, idx[i])
        checked += 1
        equal += int(ok)
        if i == pick[-1]:
            break
    for c in twins:
        c.close()
    b.close()
    t.close()
    return {"c64_worst_one_minus_f": worst, "c64_samples_differing_from_c128": differ,
            "c64_samples": total,
            "c128_circuits_checked_bit_exact": checked, "c128_circuits_equal": equal}


def ref_batch_baseline(work, shots, seconds):
    """The reference's own path per circuit (qc_create, the gates one call
    each, qc_run_shots), sequential build, 1 core, on the workload's circuits
    in order, skipping those that would not finish in the time left."""
    import ctypes
    ref = _reflib()
    if ref is None:
        return None
    from qcs_amd import circuits as C
    import qcs_amd
    qcs_amd.libc().srand(3)
    done, amp, skipped, t0 = 0, 0.0, 0, time.perf_counter()
    ns_per_amp = 30e-9  # the reference's measured rate on these hosts, for the skip test
    for n, d, seed, fused in work:
        left = seconds - (time.perf_counter() - t0)
        g = C.gate_count_layered(n, d)
        if left <= 0:
            break
        if g * float(1 << n) * ns_per_amp > left:
            skipped += 1
            continue
        c = ref.create(n)
        ref.apply(c, _Arrays(C.layered_arrays(n, d, seed)))
        hist = (ctypes.c_int * (1 << n))()
        ref.L.qc_run_shots(c, shots, hist)
        ref.L.qc_destroy(c)
        done += 1
        amp += g * float(1 << n)
    dt = time.perf_counter() - t0
    return {"value": amp / dt, "unit": "amplitude updates/s", "cores": 1, "kind": "reference",
            "circuits_per_s": done / dt,
            "sample": f"{done} circuits of the workload in order ({skipped} skipped as too "
                      f"long for the {seconds:.0f} s budget), gates + qc_run_shots({shots}) "
                      "each, reference sequential build"}


def run_mode(hip, ops, n, mode, args, steps, warmup, tune):
    """Time `steps` fused runs of `ops` in one fused mode; returns the open
    circuit (its final state) and the figures."""
    import qcs_amd
    G, N = len(ops), float(1 << n)
    opts = qcs_amd.FUSED_FP32
    if args.fuse_k:
        opts |= qcs_amd.FUSE_K(args.fuse_k)
    if mode == "tiles":
        opts |= qcs_amd.FUSE_TILES
    if mode == "split":
        opts |= qcs_amd.FUSE_SPLIT_BF16
    tiled = mode in ("tiles", "split")
    import ctypes
    fused = qcs_amd.Circuit(n, opts)
    hip.qcs_hip_set_tuning(fused.device_state, qcs_amd.TUNE_TILE_THREADS, args.tile_threads)
    if args.k4_form:
        hip.qcs_hip_set_tuning(fused.device_state, qcs_amd.TUNE_TILE_K4_FORM,
                               K4_FORMS[args.k4_form])
    fuse_k = args.fuse_k or (4 if tiled else 5)
    autotuned = tune and not args.fuse_k and not args.k4_form
    if autotuned:
        # one pass per block: kernel variant per (n, k) and the fusion width;
        # tile programs: fusion width, work-group size and grid timed on this
        # circuit's own pass plan (cached per device and gate structure)
        fused.apply(ops)
        t0 = time.perf_counter()
        fuse_k = fused.autotune()
        tune_s = time.perf_counter() - t0
        if fuse_k < 1:
            raise RuntimeError(f"qc_autotune failed ({fuse_k})")
        fused.flush()
    knob = ctypes.c_int()
    hip.qcs_hip_get_tuning(fused.device_state, qcs_amd.TUNE_TILE_THREADS, ctypes.byref(knob))
    tile_threads = knob.value
    hip.qcs_hip_get_tuning(fused.device_state, qcs_amd.TUNE_TILE_GRID, ctypes.byref(knob))
    tile_grid = knob.value
    hip.qcs_hip_get_tuning(fused.device_state, qcs_amd.TUNE_TILE_K4_FORM, ctypes.byref(knob))
    k4_form = {v: k for k, v in K4_FORMS.items()}[knob.value]

    def step():
        fused.set_basis_state(0)
        fused.apply(ops)
        fused.flush()

    for _ in range(warmup):
        step()
    state = fused.device_state
    hip.qcs_hip_device_sync()
    hip.qcs_hip_timing_enable(state, 1)
    b0 = fused.get_num_fused_blocks()
    p0 = fused.get_num_fused_passes()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    hip.qcs_hip_device_sync()
    T = (time.perf_counter() - t0) / steps
    f_ms, f_n, _, f_flops = timing(hip, state, qcs_amd.K_FUSED)
    c_ms, _, _, _ = timing(hip, state, qcs_amd.K_COMPOSE)
    hip.qcs_hip_timing_enable(state, 0)
    avg_ms = f_ms / max(f_n, 1)
    flops_per_launch = f_flops / max(f_n, 1)
    r = {
        "mode": mode,
        "value": G * N / T,
        "ms_per_step": T * 1e3,
        "gates_per_s": G / T,
        "fusion_k": fuse_k,
        "autotuned": autotuned,
        "autotune_s": tune_s if autotuned else None,
        "tile_threads": tile_threads,
        "tile_grid": tile_grid,
        "k4_form": k4_form,
        "fused_blocks": (fused.get_num_fused_blocks() - b0) // steps,
        "hbm_passes": (fused.get_num_fused_passes() - p0) // steps,
        "fused_kernel_ms_per_step": f_ms / steps,
        "compose_ms_per_step": c_ms / steps,
        "launches": f_n,
        "avg_launch_ms": avg_ms,
        "hbm_gbs": 2.0 * N * 8.0 / (avg_ms * 1e-3) / 1e9,
        "tflops": flops_per_launch / (avg_ms * 1e-3) / 1e12,
    }
    return fused, r


MODE_TEXT = {
    "blocks": ("on FP32 MFMA, one HBM pass each", "k_fused_mfma (complex64 fused k-qubit pass)"),
    "tiles": ("on FP32 MFMA, several per HBM pass (tile programs)",
              "k_tile_program (complex64, fused blocks per tile pass, FP32 MFMA)"),
    "split": ("on bf16 MFMA as split products (matrix 3 parts, state 2), several per HBM "
              "pass (tile programs)",
              "k_tile_program (complex64, fused blocks per tile pass, split-bf16 MFMA)"),
}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the state-vector path here is single-GPU by design", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload == "cpu":
        bench_cpu(args)
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "qft":
        bench_qft(args)
        return
    if args.workload == "observe":
        bench_observe(args)
        return
    if args.workload == "config5":
        bench_config5(args)
        return
    import qcs_amd
    from qcs_amd import circuits as C

    hip, _ = qcs_amd.load()
    if not hip.qcs_hip_device_available():
        raise SystemExit("no HIP device")
    n, depth = args.qubits, args.depth
    ops = C.layered(n, depth, args.seed)
    G, N = len(ops), float(1 << n)

    fused, m = run_mode(hip, ops, n, args.mode, args, args.steps, args.warmup,
                        not args.no_tune)
    tiled = args.mode in ("tiles", "split")
    print(f"bench: {args.mode} {m['ms_per_step']:.1f} ms per step", file=sys.stderr,
          flush=True)
    result = {
        "metric": METRIC,
        "value": m["value"],
        "unit": "amplitude updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 state, split-bf16 MFMA products, f32 accumulation"
                  if args.mode == "split" else "f32"),
        "data": "synthetic",
        "config": {
            "workload": (f"{n}-qubit layered Rz-Ry-Rz + CNOT-ladder circuit, depth "
                         f"{depth}, angles splitmix64(seed {args.seed}), complex64 "
                         f"state ({(8 << n) / 2**30:.0f} GiB), gates fused into "
                         f"k<={m['fusion_k']} unitaries " + MODE_TEXT[args.mode][0]),
            "qubits": n,
            "depth": depth,
            "gates": G,
            "fusion_k": m["fusion_k"],
            "autotuned": m["autotuned"],
            "fusion_mode": args.mode,
            "tile_threads": m["tile_threads"],
            "tile_grid": m["tile_grid"],
            "k4_form": m["k4_form"],
            "autotune_s": m["autotune_s"],
            "fused_blocks": m["fused_blocks"],
            "hbm_passes": m["hbm_passes"],
        },
        "gates_per_s": m["gates_per_s"],
        "fused_kernel_ms_per_step": m["fused_kernel_ms_per_step"],
        "compose_ms_per_step": m["compose_ms_per_step"],
    }
    # the dominant kernel's roofline: one-block passes stream the state once
    # per block (HBM-bound); tile programs apply several blocks per pass and
    # are bound by the matrix cores
    if tiled:
        peak = BF16_MFMA_PEAK_TFLOPS / 4.0 if args.mode == "split" else FP32_MFMA_PEAK_TFLOPS
        result["roofline"] = {
            "bound": "mfma", "kernel": MODE_TEXT[args.mode][1],
            "achieved": m["tflops"], "peak": peak, "unit": "TFLOP/s",
            "frac": m["tflops"] / peak, "traffic": None,
            "launches": m["launches"], "avg_launch_ms": m["avg_launch_ms"],
            "hbm_gbs": m["hbm_gbs"], "hbm_frac": m["hbm_gbs"] / HBM_PEAK_GBS,
        }
    else:
        result["roofline"] = {
            "bound": "hbm", "kernel": MODE_TEXT[args.mode][1],
            "achieved": m["hbm_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": m["hbm_gbs"] / HBM_PEAK_GBS, "traffic": None,
            "launches": m["launches"], "avg_launch_ms": m["avg_launch_ms"],
            "mfma_tflops": m["tflops"], "mfma_frac": m["tflops"] / FP32_MFMA_PEAK_TFLOPS,
        }
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            t = json.load(fh).get("k_tile_program" if tiled else "k_fused_mfma")
        if t:
            # measured by rocprofv3 in a separate run of this workload, not live
            result["roofline"]["traffic"] = t.get("bytes_per_launch")
            result["roofline"]["traffic_source"] = (
                "profiles/traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE of a separate run "
                "of this workload, scaled by a copy of known size")

    plain = None
    if not args.no_plain:
        plain = qcs_amd.Circuit(n)
        pstate = plain.device_state
        hip.qcs_hip_timing_enable(pstate, 1)
        hip.qcs_hip_device_sync()
        t1 = time.perf_counter()
        plain.apply(ops)
        plain.flush()
        Tp = time.perf_counter() - t1
        m1, n1, b1, _ = timing(hip, pstate, qcs_amd.K_1Q)
        m2, n2, b2, _ = timing(hip, pstate, qcs_amd.K_CU)
        mv1, mv2 = moved(hip, pstate, qcs_amd.K_1Q), moved(hip, pstate, qcs_amd.K_CU)
        omf, mx = fused.compare_with(plain)
        result["plain_fp64"] = {
            "ms": Tp * 1e3,
            "amp_updates_per_s": G * N / Tp,
            "gates_per_s": G / Tp,
            "k_1q_avg_ms": m1 / max(n1, 1),
            "k_1q_gbs": (b1 / max(n1, 1)) / (m1 / max(n1, 1) * 1e-3) / 1e9 if n1 else None,
            "k_cu_avg_ms": m2 / max(n2, 1),
            "k_cu_gbs": (b2 / max(n2, 1)) / (m2 / max(n2, 1) * 1e-3) / 1e9 if n2 else None,
        }
        result["plain_fp64"]["k_1q_hbm_frac"] = (result["plain_fp64"]["k_1q_gbs"] or 0) / HBM_PEAK_GBS
        # the plain form as a whole: algorithmic bytes of every gate over the
        # gate kernels' time, and over the step's wall time
        if n1 + n2:
            gk = (b1 + b2) / ((m1 + m2) * 1e-3) / 1e9
            result["plain_fp64"]["gate_kernels_gbs"] = gk
            result["plain_fp64"]["gate_kernels_hbm_frac"] = gk / HBM_PEAK_GBS
            result["plain_fp64"]["step_hbm_frac"] = (b1 + b2) / Tp / 1e9 / HBM_PEAK_GBS
        # what HBM moved: whole 128-byte lines (a controlled gate on qubit 0,
        # 1 or 2 reads and writes every line, twice its algorithmic bytes)
        for key, mv, ms_, n_ in (("k_1q", mv1, m1, n1), ("k_cu", mv2, m2, n2)):
            gbs = mv / (ms_ * 1e-3) / 1e9 if n_ else None
            result["plain_fp64"][key + "_moved_gbs"] = gbs
            result["plain_fp64"][key + "_moved_hbm_frac"] = (gbs or 0) / HBM_PEAK_GBS
        result["speedup_vs_plain_fp64"] = Tp / m["ms_per_step"] * 1e3
        result["one_minus_f"] = omf
        result["max_abs_err"] = mx
    fused.close()

    # the other fused modes on the same circuit, for comparison (not `value`):
    # one pass per block, and split-bf16 tile programs (bf16 matrix cores,
    # the state rounded to 16 significant bits per block)
    if not args.no_alt:
        alt = {}
        for mode in ("tiles", "blocks", "split"):
            if mode == args.mode:
                continue
            f2, m2 = run_mode(hip, ops, n, mode, args, min(args.steps, 3), 1,
                              not args.no_tune)
            if plain is not None:
                m2["one_minus_f"], m2["max_abs_err"] = f2.compare_with(plain)
            f2.close()
            alt[mode] = {k: m2[k] for k in ("value", "ms_per_step", "fused_blocks",
                                            "hbm_passes", "avg_launch_ms", "hbm_gbs",
                                            "tflops", "fusion_k")}
            for k in ("one_minus_f", "max_abs_err"):
                if k in m2:
                    alt[mode][k] = m2[k]
            print(f"bench: {mode} {m2['ms_per_step']:.1f} ms per step", file=sys.stderr,
                  flush=True)
        result["other_modes"] = alt
    if plain is not None:
        plain.close()

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
