"""include/bmg.h from plain C: tests/c/bmg_harness.c is compiled with gcc
as strict C99 (-pedantic -Werror) and linked against libbmg.so.  Without a
GPU it checks the struct layouts and the error convention; with one it also
solves a 2D Poisson problem through bmg_poisson_view on storage laid out
like the reference's, verified by a residual computed in the C program."""
import os
import subprocess

import pytest

import basilisk_amd as bmg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_harness(tmp_path):
    exe = str(tmp_path / "bmg_harness")
    libdir = os.path.dirname(bmg.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O1",
                    "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c", "bmg_harness.c"),
                    "-o", exe, "-L", libdir, "-lbmg", "-Wl,-rpath," + libdir, "-lm"], check=True)
    return exe


def test_c_harness_layouts(tmp_path):
    exe = build_harness(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert "layouts ok" in out.stdout


@pytest.mark.gpu
def test_c_harness_solve(tmp_path):
    exe = build_harness(tmp_path)
    print("case c harness solve", flush=True)
    out = subprocess.run([exe, "solve"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "solve ok" in out.stdout
