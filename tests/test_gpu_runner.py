"""The drop-in command line on the GPU: fibers_runner.py run / validate /
benchmark as the reference's runner drives them (runner/cuda_runner.py:154-215
for run and validate, :217-339 for benchmark; the dumps and saves of
src/simulation.cu:530-822, performance.out of src/performance.cc:110-134),
each as its own process on the reference's M = 100 inputs."""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import GOLDEN, ref100, ref_config

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIG = os.path.join(GOLDEN, "config.ini")
FIBERS = os.path.join(GOLDEN, "XcT_ref100.in")


def _runner(*args, timeout=300):
    cmd = [sys.executable, "-u", os.path.join(ROOT, "fibers_runner.py"), *args]
    print("$", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    print(p.stdout[-3000:], p.stderr[-3000:], flush=True)
    return p


def _read_state(path):
    """M, then "x y z" / "tx ty tz" per fiber (src/simulation.cu:530-566)."""
    with open(path) as fh:
        M = int(fh.readline())
        vals = np.array([float(v) for v in fh.read().split()]).reshape(M, 2, 3)
    return vals[:, 0, :].ravel(), vals[:, 1, :].ravel()


@pytest.mark.parametrize("solver", ["--gmres", "--magma"])
@pytest.mark.parametrize("integration", ["--numerical", "--analytical"])
def test_validate_command(tmp_path, solver, integration):
    """`cuda validate` in the four variants of the reference's golden runs:
    exit status 0, every comparison within 1e-5, and the dumps of both steps
    under the CUDA names."""
    p = _runner("cuda", "validate", solver, integration, "--D2", CONFIG, FIBERS, "--output-dir", str(tmp_path))
    assert p.returncode == 0
    assert "Validation passed" in p.stdout
    assert p.stdout.count("Maximum Delta") == 12  # 6 fields x 2 steps
    for step in (0, 1):
        for name in ("b_vector", "x_vector", "t_vel", "r_vel", "positions", "orientations"):
            assert (tmp_path / f"{name}_{step}.out").stat().st_size > 0
    assert (tmp_path / "current.map").exists()


@pytest.mark.parametrize("solver,name", [("--gmres", "gmres"), ("--magma", "direct")])
def test_run_command_saves_the_trajectory(tmp_path, solver, name):
    """`cuda run` over the configuration's 10 steps writes XcT_res_00001 ..
    XcT_res_00010 and TcR_res_* (STATE/VELOCITY_SAVE_INTERVAL = 1); the last
    state file equals a Simulation stepped the same way, to the files' 8
    decimals."""
    from rigidfibers_amd.simulation import Simulation

    p = _runner("cuda", "run", solver, "--numerical", "--D2", CONFIG, FIBERS, "--output-dir", str(tmp_path))
    assert p.returncode == 0
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("XcT_res_"))
    assert files == [f"XcT_res_{t:05d}.out" for t in range(1, 11)]
    assert len([f for f in os.listdir(tmp_path) if f.startswith("TcR_res_")]) == 10
    M, pos, ori = ref100()
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    with Simulation(params, pos, ori, solver=name, integration="numerical") as sim:
        for t in range(10):
            sim.step(t)
            if t == 0:
                p1, o1 = sim.state()
                v1, w1 = sim.velocities()
        p10, o10 = sim.state()
    for path, (a, b) in ((tmp_path / "XcT_res_00001.out", (p1, o1)), (tmp_path / "XcT_res_00010.out", (p10, o10)),
                         (tmp_path / "TcR_res_00001.out", (v1, w1))):
        fa, fb = _read_state(path)
        assert np.abs(fa - a).max() <= 5.1e-9 and np.abs(fb - b).max() <= 5.1e-9, path.name


def test_benchmark_command(tmp_path):
    """`cuda benchmark` over two sizes writes performance.out (name,time for
    the four phases of the last size) and results.csv (one row per size, the
    mean and standard deviation of every phase and of the total)."""
    p = _runner("cuda", "benchmark", "--gmres", "--numerical", "--D2", CONFIG, "--sizes", "200,100",
                "--output-dir", str(tmp_path))
    assert p.returncode == 0
    with open(tmp_path / "performance.out") as fh:
        rows = [line.strip().split(",") for line in fh if line.strip()]
    assert rows[0] == ["name", "time"]
    assert [r[0] for r in rows[1:]] == ["assemble_system", "solve_system", "update_velocities", "update_fibers"]
    assert all(float(r[1]) > 0 for r in rows[1:])
    with open(tmp_path / "results.csv") as fh:
        lines = [line.rstrip("\n").split("\t") for line in fh]
    head, body = lines[0], lines[1:]
    assert head[0] == "X" and "total" in head and "assemble_system_std" in head
    assert [int(r[0]) for r in body] == [100, 200]
    col = head.index("total")
    assert all(float(r[col]) > 0 for r in body)
