"""qc_optimize on the host (no device needed: QCS_DEFERRED circuits record
without a state): the linear reduction gives the history the reference's
restart-from-the-start loop (src/qcs.c:650-698) gives."""
import ctypes
import random

import pytest

import qcs_amd


def reference_optimize(hist):
    """The reference's algorithm, restated: remove the first adjacent pair of
    identical H/X/Y/Z (same target) or CNOT (same control and target), then
    start again from the beginning."""
    h = list(hist)
    i = 0
    while i < len(h) - 1:
        (a, ta, ca), (b, tb, cb) = h[i], h[i + 1]
        one = ta == tb and a == b and a in ("H", "X", "Y", "Z")
        cx = a == "CNOT" and b == "CNOT" and ta == tb and ca == cb
        if one or cx:
            del h[i:i + 2]
            i = 0
        else:
            i += 1
    return h


def sequence_line(c, capfd):
    libc = ctypes.CDLL(None)
    capfd.readouterr()
    c.print_circuit()
    libc.fflush(None)
    out = capfd.readouterr().out
    return [l for l in out.splitlines() if l.startswith("GATE SEQUENCE:")][0]


def fmt(h):
    parts = []
    for name, t, c in h:
        parts.append(f"{name}({c},{t})" if name == "CNOT" else f"{name}({t})")
    return "GATE SEQUENCE: " + " ".join(parts)


@pytest.mark.parametrize("seed", range(6))
def test_optimize_matches_reference_loop(seed, capfd):
    rnd = random.Random(seed)
    n = 3
    c = qcs_amd.Circuit(n, qcs_amd.DEFERRED)
    hist = []
    for _ in range(60):
        k = rnd.choice(["H", "X", "Y", "Z", "CNOT", "CNOT", "RZ"])
        t = rnd.randrange(n)
        if k == "CNOT":
            ctl = rnd.choice([q for q in range(n) if q != t])
            c.cnot(ctl, t)
            hist.append(("CNOT", t, ctl))
        elif k == "RZ":
            c.rz(t, 0.25)
            hist.append(("RZ", t, -1))
        else:
            getattr(c, k.lower())(t)
            hist.append((k, t, -1))
    want = reference_optimize(hist)
    c.optimize()
    assert c.get_num_gates() == len(want)
    assert sequence_line(c, capfd).rstrip() == fmt(want).rstrip()
    c.close()
