"""A C89 program against include/qcs.h, built with gcc -std=c89 -pedantic and
linked with -lqcs: it compiles here, and on a GPU every check passes."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "drop_in.c")
LIBDIR = os.path.join(ROOT, "qcs_amd")


def build(tmp_path):
    exe = str(tmp_path / "drop_in")
    subprocess.run(["gcc", "-std=c89", "-pedantic", "-Wall", "-Werror",
                    "-I" + os.path.join(ROOT, "include"), SRC, "-o", exe,
                    "-L" + LIBDIR, "-lqcs", "-Wl,-rpath," + LIBDIR, "-lm"],
                   check=True, capture_output=True, text=True)
    return exe


def test_c89_program_builds(tmp_path):
    assert os.path.exists(build(tmp_path))


@pytest.mark.gpu
def test_c89_program_runs(tmp_path, qlib):
    exe = build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("ok ") == 12
