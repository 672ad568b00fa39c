"""Precision form of the timestep against the FP64 oracle: the mixed BF16 /
FP16 MFMA LUs with FP64 iterative refinement (and the FP32 LU beside them),
analytical inner integrals, on balls at the number density of the M = 5000
ball of radius 35.

  * 10 steps at M = 250: positions and orientations within 1e-6 (relative,
    norm-wise) of the oracle's trajectory -- the accuracy target of the mixed
    forms;
  * M = 5000 (n = 75 000): the rows of A and b of three groups of target
    fibers against the oracle's rows, the FP64 residual of the refined
    solution on those rows (in the GPU's FP32 system and in the oracle's FP64
    one), the velocities of the sampled fibers, and the state after two steps
    against the FP32 LU's.

The oracle's solve is its LU (sgesv semantics, fortran/solve_system.incl:3);
over the 10-step trajectory it uses its GMRES at a relative residual of
1e-13, checked against its LU on the first step's system.
"""
import numpy as np
import pytest

from helpers import ref_config, rel_err

pytestmark = pytest.mark.gpu

BALL_DENSITY = 5000 / (4.0 / 3.0 * np.pi * 35.0 ** 3)


def _ball(M, seed):
    from rigidfibers_amd.scenes import ball

    return ball(M, (M / (4.0 / 3.0 * np.pi) / BALL_DENSITY) ** (1 / 3), min_dist=0.4, seed=seed)


def _params(M):
    p = ref_config(validate=False)
    p.NUMBER_OF_FIBERS = M
    return p


@pytest.fixture(scope="module")
def oracle_trajectory(oracle64):
    """The oracle's 10 analytical steps at M = 250 (positions and orientations
    after each step)."""
    from oracle.pyoracle import make_params

    M = 250
    pos, ori = _ball(M, seed=250)
    p = make_params(M, threads=8)
    A, b = oracle64.assemble(p, pos, ori, True)
    x_lu = oracle64.solve_direct(A, b)
    x_kr, info = oracle64.solve_gmres(A, b, 30, 2000, 1e-13)
    assert rel_err(x_kr, x_lu) < 1e-11, "the oracle's tight GMRES is not its LU solution"
    rec = []
    oracle64.simulate(p, pos, ori, 10, solver="gmres", analytical=True, gmres=(30, 2000, 1e-13), record=rec)
    return M, pos, ori, rec


@pytest.mark.parametrize("solver", ["mixed-bf16", "mixed-fp16", "direct"])
def test_lu_ten_steps_against_oracle(oracle_trajectory, solver):
    from rigidfibers_amd.simulation import Simulation

    M, pos, ori, rec = oracle_trajectory
    worst_p = worst_o = 0.0
    with Simulation(_params(M), pos, ori, solver=solver, integration="analytical",
                    refinement_tolerance=1e-10) as sim:
        for t in range(10):
            sim.step(t)
            st = sim.solver_stats()
            assert st["converged"], (t, st)
            p, o = sim.state()
            worst_p = max(worst_p, rel_err(p, rec[t]["pos"]))
            worst_o = max(worst_o, rel_err(o, rec[t]["ori"]))
    print(f"{solver} M={M} 10 steps: positions {worst_p:.2e}, orientations {worst_o:.2e} against the FP64 oracle")
    assert worst_p <= 1e-6 and worst_o <= 1e-6


@pytest.fixture(scope="module")
def m5000_fp32_lu_states():
    """The plain FP32 LU's state after 1 and 2 steps of the M = 5000 ball."""
    from rigidfibers_amd.scenes import ball
    from rigidfibers_amd.simulation import Simulation

    pos, ori = ball(5000, 35.0, min_dist=0.4, seed=2026)
    out = []
    with Simulation(_params(5000), pos, ori, solver="direct", integration="analytical",
                    refinement_tolerance=1e-10) as sim:
        for t in range(2):
            sim.step(t)
            out.append(sim.state())
    return pos, ori, out


@pytest.mark.parametrize("solver", ["mixed-bf16", "mixed-fp16"])
def test_m5000_ball_mixed_lu_against_oracle(oracle64, oracle32, m5000_fp32_lu_states, solver):
    from oracle.pyoracle import make_params
    from rigidfibers_amd.simulation import Simulation

    pos, ori, fp32_states = m5000_fp32_lu_states
    M = 5000
    p = make_params(M, threads=16)
    with Simulation(_params(M), pos, ori, solver=solver, integration="analytical",
                    refinement_tolerance=1e-10) as sim:
        sim.assemble_system()
        if solver == "mixed-bf16":  # the assembly does not depend on the solver: check A once
            A, b, _ = sim.system(with_matrix=True)
        sim.solve_system()
        st = sim.solver_stats()
        _, b, x = sim.system()
        sim.update_velocities()
        v, w = sim.velocities()
        sim.update_fibers(True)
        sim.advance()
        s1 = sim.state()
        sim.step(1)
        s2 = sim.state()
    assert st["converged"] and st["relative_residual"] <= 1e-10
    bnorm = np.linalg.norm(b)
    for i0 in (0, 2499, 4997):
        i1 = i0 + 3
        rows = slice(15 * i0, 15 * i1)
        Ar, br = oracle64.assemble_rows(p, pos, ori, i0, i1, True)
        r_oracle = np.linalg.norm(br - Ar @ x) / np.linalg.norm(br)
        vr, wr = oracle64.velocities_rows(p, pos, ori, x, i0, i1, True)
        ev, ew = rel_err(v[3 * i0:3 * i1], vr), rel_err(w[3 * i0:3 * i1], wr)
        msg = f"{solver} rows {i0}..{i1}: residual in the oracle's system {r_oracle:.1e}, v {ev:.1e}, w {ew:.1e}"
        if solver == "mixed-bf16":
            A32, b32 = oracle32.assemble_rows(p, pos, ori, i0, i1, True)
            Ag = np.asarray(A[rows], dtype=np.float64)
            ea, eb, ra, rb = rel_err(A[rows], Ar), rel_err(b[rows], br), rel_err(A32, Ar), rel_err(b32, br)
            r_own = np.linalg.norm(b[rows] - Ag @ x) / bnorm
            msg += (f", A {ea:.1e} (FP32 oracle {ra:.1e}), b {eb:.1e} (FP32 oracle {rb:.1e}), "
                    f"residual in the GPU's FP32 system {r_own:.1e}")
        print(msg)
        if solver == "mixed-bf16":
            # FP64 recurrences: no FP32 blow-up, entries at FP32 rounding; b sums
            # the FP32-rounded contributions of 5000 sources with cancellation
            assert ea < 2e-6 and ea < max(2e-6, 3 * ra) and eb < max(2e-6, 3 * rb)
            assert r_own <= 1e-10
        assert r_oracle <= 1e-5
        assert ev < 1e-6 and ew < 1e-6
    if solver == "mixed-bf16":
        del A
    for k, (ps, os_) in enumerate((s1, s2)):
        ep, eo = rel_err(ps, fp32_states[k][0]), rel_err(os_, fp32_states[k][1])
        print(f"{solver} step {k + 1}: positions {ep:.1e}, orientations {eo:.1e} against the FP32 LU")
        assert ep <= 1e-6 and eo <= 1e-6
