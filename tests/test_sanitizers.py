"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

tests/c/sanitize_host.c is built with gcc -fsanitize=address,undefined
together with the oracle (oracle/qcs_oracle.c) and a sanitized copy of the
host library (qcs_amd/src/*.c, linked against the uninstrumented
libqcs_hip.so), and run: any out-of-bounds access, use after free,
misaligned or overflowing arithmetic in the oracle, the planners or the
qc_* entry points stops it with a report.  No GPU is used (without a device
every qc_* call on a circuit takes its no-device path).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "qcs_amd")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
       "-fno-omit-frame-pointer", "-g", "-O1"]


def _have_libqcs_hip():
    return os.path.exists(os.path.join(LIBDIR, "libqcs_hip.so"))


@pytest.fixture(scope="module")
def sanitized_exe(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    if not _have_libqcs_hip():
        pytest.skip("libqcs_hip.so not built")
    out = tmp_path_factory.mktemp("san")
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "qcs_amd", "src"),
           "-I" + os.path.join(ROOT, "oracle")]
    srcs = sorted(os.path.join(ROOT, "qcs_amd", "src", f)
                  for f in os.listdir(os.path.join(ROOT, "qcs_amd", "src")) if f.endswith(".c"))
    lib = str(out / "libqcs_san.so")
    subprocess.run(["gcc", "-std=c89", "-fPIC", "-shared", *SAN, *inc, *srcs, "-o", lib,
                    "-L" + LIBDIR, "-lqcs_hip", "-Wl,-rpath," + LIBDIR, "-lm"],
                   check=True, capture_output=True, text=True)
    exe = str(out / "sanitize_host")
    subprocess.run(["gcc", "-std=c99", "-fopenmp", *SAN, *inc,
                    os.path.join(ROOT, "tests", "c", "sanitize_host.c"),
                    os.path.join(ROOT, "oracle", "qcs_oracle.c"), "-o", exe,
                    "-L" + str(out), "-lqcs_san", "-Wl,-rpath," + str(out),
                    "-L" + LIBDIR, "-lqcs_hip", "-Wl,-rpath," + LIBDIR, "-lm"],
                   check=True, capture_output=True, text=True)
    return exe


def test_host_code_clean_under_asan_ubsan(sanitized_exe):
    env = dict(os.environ)
    # the HIP runtime behind libqcs_hip.so is not instrumented and keeps
    # allocations until exit: leak checking is off, everything else on
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["OMP_NUM_THREADS"] = "2"
    r = subprocess.run([sanitized_exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sanitize ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
