"""A C program drives the library through include/rigidfibers.h alone
(tests/c_client/ref_client.c): it compiles and links with gcc against the
built librigidfibers.so (CPU), and on the GPU its two timesteps of the
reference's M=100 scene reproduce the golden positions and orientations of
tests/reference/100_numeric_gmres (runner/cuda_runner.py:165-215)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(ROOT, "rigidfibers_amd")


def _build(tmp_path):
    exe = tmp_path / "ref_client"
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "c_client", "ref_client.c"), "-L", LIBDIR, "-lrigidfibers",
                    f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", str(exe)], check=True)
    return exe


def test_c_client_builds_against_the_header(tmp_path):
    exe = _build(tmp_path)
    assert exe.exists()
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 2 and "usage" in out.stderr  # argument check only: no device call


@pytest.mark.gpu
def test_c_client_reproduces_golden_steps(tmp_path):
    exe = _build(tmp_path)
    ref = os.path.join(HERE, "golden", "reference", "100_numeric_gmres")
    out = subprocess.run([str(exe), os.path.join(HERE, "golden", "XcT_ref100.in"), "2",
                          os.path.join(ref, "1_POS.ref"), os.path.join(ref, "1_ORIENT.ref")],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
