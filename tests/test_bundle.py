"""The single-header form of the host library (qcs_amd/bundle.py): the
generated qcs.h builds C89 programs with -pedantic -Werror (its private
names cannot clash with the program's), links from several translation
units, and a host-only program built from it prints what the same program
linked with libqcs.so prints.  On a GPU the drop-in program built from it
passes every check."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "qcs_amd")
CDIR = os.path.join(ROOT, "tests", "c")
CFLAGS = ["gcc", "-std=c89", "-pedantic", "-Wall", "-Wextra", "-Werror"]
LINK_HIP = ["-L" + LIBDIR, "-lqcs_hip", "-Wl,-rpath," + LIBDIR, "-lpthread", "-lm"]


def make_bundle(tmp_path):
    subprocess.run([sys.executable, os.path.join(LIBDIR, "bundle.py"), "--out",
                    str(tmp_path / "qcs.h")], check=True, capture_output=True, text=True)
    return str(tmp_path)


def cc(args):
    r = subprocess.run(args, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r


def drop_in_from_bundle(tmp_path):
    inc = make_bundle(tmp_path)
    exe = str(tmp_path / "drop_in_bundle")
    # drop_in.c includes system headers before qcs.h: the POSIX level the
    # implementation needs comes from the command line
    cc(CFLAGS + ["-DQCS_GPU_HIP", "-DQCS_IMPLEMENTATION", "-D_POSIX_C_SOURCE=200112L",
                 "-I" + inc, os.path.join(CDIR, "drop_in.c"), "-o", exe] + LINK_HIP)
    return exe


def test_bundle_builds_drop_in_program(tmp_path):
    assert os.path.exists(drop_in_from_bundle(tmp_path))


def test_bundle_header_only_part_is_c89(tmp_path):
    inc = make_bundle(tmp_path)
    obj = str(tmp_path / "other.o")
    cc(CFLAGS + ["-I" + inc, "-c", os.path.join(CDIR, "bundle_other.c"), "-o", obj])


def test_bundle_rejects_other_modes(tmp_path):
    inc = make_bundle(tmp_path)
    src = tmp_path / "m.c"
    src.write_text('#define QCS_GPU_OPENCL\n#include "qcs.h"\nint main(void) { return 0; }\n')
    r = subprocess.run(CFLAGS + ["-I" + inc, "-c", str(src), "-o", str(tmp_path / "m.o")],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "HIP build" in r.stderr


def test_bundle_matches_library_host_only(tmp_path):
    inc = make_bundle(tmp_path)
    exe_b, exe_l = str(tmp_path / "host_bundle"), str(tmp_path / "host_lib")
    # the implementation in bundle_host.c only; bundle_other.c includes the API
    main_o, other_o = str(tmp_path / "host.o"), str(tmp_path / "other.o")
    cc(CFLAGS + ["-DQCS_IMPLEMENTATION", "-I" + inc, "-c", os.path.join(CDIR, "bundle_host.c"),
                 "-o", main_o])
    cc(CFLAGS + ["-I" + inc, "-c", os.path.join(CDIR, "bundle_other.c"), "-o", other_o])
    cc(["gcc", main_o, other_o, "-o", exe_b] + LINK_HIP)
    cc(CFLAGS + ["-I" + os.path.join(ROOT, "include"), os.path.join(CDIR, "bundle_host.c"),
                 os.path.join(CDIR, "bundle_other.c"), "-o", exe_l, "-L" + LIBDIR, "-lqcs",
                 "-Wl,-rpath," + LIBDIR, "-lm"])
    out_b = subprocess.run([exe_b], capture_output=True, text=True, timeout=60)
    out_l = subprocess.run([exe_l], capture_output=True, text=True, timeout=60)
    assert out_b.returncode == 0 and out_l.returncode == 0, out_b.stderr + out_l.stderr
    assert out_b.stdout == out_l.stdout
    assert out_b.stdout.startswith("gates 14 qubits 6 other 14 options 32\n")


@pytest.mark.gpu
def test_bundle_drop_in_program_runs(tmp_path, qlib):
    exe = drop_in_from_bundle(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("ok ") == 12
