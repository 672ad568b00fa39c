"""GPU parity: the HIP path (through the C ABI) against the reference's golden
vectors and the FP64 oracle.

Tolerances are norm-wise relative errors (max |gpu - ref| / max |ref|):
  * 1e-5 for the state (positions, orientations) -- the FP32 drop-in bar;
  * 1e-5 for b, x and the velocities (FP32 assembly of a system with
    condition number ~4.5; observed errors are ~1e-7).
"""
import numpy as np
import pytest

from helpers import FIELDS, VARIANTS, golden, random_scene, ref100, ref_config, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _sim(params, pos, ori, **kw):
    from rigidfibers_amd.simulation import Simulation

    return Simulation(params, pos, ori, **kw)


def _variant_kwargs(variant):
    integration = "analytical" if variant.startswith("analytic") else "numerical"
    solver = "direct" if variant.endswith("direct") else "gmres"
    return dict(solver=solver, integration=integration)


def _run_recorded(sim, steps):
    out = []
    for t in range(steps):
        sim.assemble_system()
        _, b, _ = sim.system()
        sim.solve_system()
        _, _, x = sim.system()
        sim.update_velocities()
        v, w = sim.velocities()
        sim.update_fibers(t == 0)
        p, o = sim.next_state()
        sim.advance()
        out.append({"BVec": b, "XVec": x, "TRANSVel": v, "ROTVel": w, "POS": p, "ORIENT": o})
    return out


@pytest.mark.parametrize("variant", VARIANTS)
def test_golden_vectors(variant):
    """Two validated timesteps of the reference's M=100 scene, every dumped
    quantity against tests/reference/100_<variant> (runner/cuda_runner.py:165-215)."""
    M, pos, ori = ref100()
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    with _sim(params, pos, ori, **_variant_kwargs(variant)) as sim:
        rec = _run_recorded(sim, 2)
        stats = sim.solver_stats()
    worst = {}
    for step in range(2):
        for f in FIELDS:
            e = rel_err(rec[step][f], golden(variant, step, f))
            worst[f"{step}_{f}"] = e
    print(variant, {k: f"{v:.1e}" for k, v in worst.items()}, stats)
    for k, e in worst.items():
        assert e < TOL, f"{variant} {k}: relative error {e:.2e}"


def test_teacher_forced_second_step():
    """Step 1 started from the golden step-0 state (multistep history set
    through rf_set_state) reproduces the golden step-1 vectors."""
    variant = "numeric_direct"
    M, pos0, ori0 = ref100()
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    p1, o1 = golden(variant, 0, "POS"), golden(variant, 0, "ORIENT")
    v0, w0 = golden(variant, 0, "TRANSVel"), golden(variant, 0, "ROTVel")
    with _sim(params, pos0, ori0, **_variant_kwargs(variant)) as sim:
        sim.set_state(p1, o1, previous=(pos0, ori0, v0, w0))
        sim.step(1)
        p, o = sim.state()
        v, w = sim.velocities()
    for f, a in (("POS", p), ("ORIENT", o), ("TRANSVel", v), ("ROTVel", w)):
        assert rel_err(a, golden(variant, 1, f)) < TOL, f


@pytest.mark.parametrize("N,nqi,analytical", [(5, 8, False), (5, 8, True), (1, 8, False), (3, 4, True),
                                              (8, 16, False), (8, 8, True), (2, 16, True), (6, 4, False)])
def test_assembly_matches_oracle(oracle64, oracle32, N, nqi, analytical):
    """A entrywise and b against the FP64 oracle for other expansion orders and
    quadrature sizes (random 20-fiber scene with near pairs).  The bound is the
    larger of 2e-6 and 3x the error the reference's own FP32 arithmetic (the
    FP32 build of the oracle) makes on the same scene: near-touching fibers
    amplify FP32 rounding in the numerical integrals, and the reference's FP32
    analytical recurrences can lose every digit (the path here uses FP64)."""
    from oracle.pyoracle import make_params

    M = 20
    pos, ori = random_scene(M, side=3.0, seed=N * 100 + nqi, min_dist=0.3)
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    params.NUMBER_OF_TERMS_IN_FORCE_EXPANSION = N
    params.NUMBER_OF_QUADRATURE_INTERVALS = nqi
    with _sim(params, pos, ori, solver="gmres", integration="analytical" if analytical else "numerical") as sim:
        sim.assemble_system()
        A, b, _ = sim.system(with_matrix=True)
    p = make_params(M, N=N, nqi=nqi)
    A_ref, b_ref = oracle64.assemble(p, pos, ori, analytical)
    A_32, b_32 = oracle32.assemble(p, pos, ori, analytical)
    ea, eb = rel_err(A, A_ref), rel_err(b, b_ref)
    ra, rb = rel_err(A_32, A_ref), rel_err(b_32, b_ref)
    print(f"N={N} nqi={nqi} analytical={analytical}: A {ea:.1e} (ref FP32 {ra:.1e}) b {eb:.1e} (ref FP32 {rb:.1e})")
    assert ea < max(2e-6, 3 * ra) and eb < max(2e-6, 3 * rb)
    if analytical:  # FP64 recurrences: no FP32 blow-up
        assert ea < 2e-6 and eb < 2e-6


@pytest.mark.parametrize("operator", ["stored", "pairwise"])
@pytest.mark.parametrize("N,nqi,analytical", [(5, 8, False), (5, 8, True), (3, 4, True), (8, 16, False),
                                              (1, 8, True)])
def test_velocities_match_oracle(oracle64, N, nqi, analytical, operator):
    """Velocities from the GPU's own solution against the oracle's velocity
    routine fed the same solution (isolates the velocity update), for the
    operator stored by the theta assembly and for the pairwise kernel."""
    from oracle.pyoracle import make_params

    M = 40
    pos, ori = random_scene(M, side=5.0, seed=11, min_dist=0.4)
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    params.NUMBER_OF_TERMS_IN_FORCE_EXPANSION = N
    params.NUMBER_OF_QUADRATURE_INTERVALS = nqi
    params.GMRES_TOLERANCE = 1e-10
    with _sim(params, pos, ori, solver="gmres", integration="analytical" if analytical else "numerical",
              velocity_operator=operator) as sim:
        sim.assemble_system()
        sim.solve_system()
        _, _, x = sim.system()
        sim.update_velocities()
        v, w = sim.velocities()
    v_ref, w_ref = oracle64.velocities(make_params(M, N=N, nqi=nqi), pos, ori, x, analytical)
    ev, ew = rel_err(v, v_ref), rel_err(w, w_ref)
    print(f"N={N} nqi={nqi} analytical={analytical} {operator}: v {ev:.1e} w {ew:.1e}")
    assert ev < 1e-6 and ew < 1e-6


def test_stored_velocities_follow_state_changes(oracle64):
    """After rf_set_state the stored velocity operator belongs to the old state:
    rf_update_velocities must not use it (pairwise fallback), and a new
    assembly makes it current again."""
    from oracle.pyoracle import make_params

    M = 30
    pos, ori = random_scene(M, side=4.5, seed=5, min_dist=0.4)
    pos2, ori2 = random_scene(M, side=4.5, seed=6, min_dist=0.4)
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-10
    with _sim(params, pos, ori) as sim:
        sim.assemble_system()
        sim.solve_system()
        _, _, x = sim.system()
        sim.set_state(pos2, ori2)
        sim.update_velocities()  # operator of (pos, ori) is stale: pairwise for (pos2, ori2)
        v, w = sim.velocities()
        v_ref, w_ref = oracle64.velocities(make_params(M), pos2, ori2, x, False)
        assert rel_err(v, v_ref) < 1e-6 and rel_err(w, w_ref) < 1e-6
        sim.assemble_system()
        sim.solve_system()
        _, _, x2 = sim.system()
        sim.update_velocities()
        v, w = sim.velocities()
    v_ref, w_ref = oracle64.velocities(make_params(M), pos2, ori2, x2, False)
    assert rel_err(v, v_ref) < 1e-6 and rel_err(w, w_ref) < 1e-6


@pytest.mark.parametrize("analytical", [False, True])
def test_m2000_sampled_rows_against_oracle(oracle64, oracle32, analytical):
    """M = 2000 fibers in a cube of side 40 (n = 30 000): the rows of A and b of
    three groups of target fibers against the FP64 oracle's rows (bound: the
    larger of 2e-6 and 3x the FP32 oracle's error on the same rows), the
    residual of the GPU's GMRES solution on those rows, and the velocities of
    the sampled fibers from the GPU's solution against the oracle's (bound: the
    larger of 1e-6 and 3x the FP32 oracle's error from the same solution; the
    rotational velocity is a small difference of FP32-stored sums)."""
    from oracle.pyoracle import make_params
    from rigidfibers_amd.scenes import cube

    M = 2000
    pos, ori = cube(M, 40.0, min_dist=0.4, seed=2000)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-6
    with _sim(params, pos, ori, integration="analytical" if analytical else "numerical") as sim:
        sim.assemble_system()
        A, b, _ = sim.system(with_matrix=True)
        sim.solve_system()
        st = sim.solver_stats()
        _, _, x = sim.system()
        sim.update_velocities()
        v, w = sim.velocities()
    assert st["converged"]
    p = make_params(M)
    bnorm = np.linalg.norm(b)
    for i0 in (0, 997, 1997):
        i1 = i0 + 3
        rows = slice(15 * i0, 15 * i1)
        Ar, br = oracle64.assemble_rows(p, pos, ori, i0, i1, analytical)
        A32, b32 = oracle32.assemble_rows(p, pos, ori, i0, i1, analytical)
        ea, eb = rel_err(A[rows], Ar), rel_err(b[rows], br)
        ra, rb = rel_err(A32, Ar), rel_err(b32, br)
        r = br - Ar @ x
        vr, wr = oracle64.velocities_rows(p, pos, ori, x, i0, i1, analytical)
        v32, w32 = oracle32.velocities_rows(p, pos, ori, x, i0, i1, analytical)
        ev, ew = rel_err(v[3 * i0:3 * i1], vr), rel_err(w[3 * i0:3 * i1], wr)
        rv, rw = rel_err(v32, vr), rel_err(w32, wr)
        print(f"rows {i0}..{i1}: A {ea:.1e} (FP32 oracle {ra:.1e}) b {eb:.1e} (FP32 oracle {rb:.1e}) "
              f"residual {np.linalg.norm(r) / bnorm:.1e} v {ev:.1e} (FP32 oracle {rv:.1e}) "
              f"w {ew:.1e} (FP32 oracle {rw:.1e})")
        assert ea < max(2e-6, 3 * ra) and eb < max(2e-6, 3 * rb)
        assert np.linalg.norm(r) <= 1e-5 * bnorm
        assert ev < max(1e-6, 3 * rv) and ew < max(1e-6, 3 * rw)


def test_single_fiber_known_answer():
    """M = 1: A = I, b = 0, x = 0, v = (0.5/d)((d+2) F + (d-2)(t.F) t), w = 0
    (src/kernels/reset_velocities.cu:22-38)."""
    t = np.array([0.3, -0.4, 0.5])
    t /= np.linalg.norm(t)
    params = ref_config()
    params.NUMBER_OF_FIBERS = 1
    with _sim(params, [1.0, 2.0, 3.0], t) as sim:
        sim.step(0)
        A, b, x = sim.system(with_matrix=True)
        v, w = sim.velocities()
        p, o = sim.state()
    d = -np.log(0.01 ** 2 * np.e)
    F = np.array([0.0, 0.0, -1.0])
    v_ref = (0.5 / d) * ((d + 2) * F + (d - 2) * (t @ F) * t)
    assert np.array_equal(A, np.eye(15, dtype=np.float32))
    assert np.all(b == 0) and np.all(x == 0)
    assert np.allclose(v, v_ref, rtol=1e-12, atol=1e-14) and np.all(w == 0)
    assert np.allclose(p, np.array([1.0, 2.0, 3.0]) + 0.1 * v_ref, rtol=1e-12)
    assert np.allclose(o, t, rtol=1e-12)


def test_ring_symmetry():
    """16 vertical fibers on a ring (the reference's tests/XcT_ring16.in): by
    symmetry every fiber has the same speed and the same vertical velocity, to
    the precision of the ring's 8-decimal coordinates and the FP32 system."""
    import os

    from helpers import GOLDEN
    from rigidfibers_amd.config import read_fibers

    M, pos, ori = read_fibers(os.path.join(GOLDEN, "XcT_ring16.in"))
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-10
    with _sim(params, pos, ori) as sim:
        sim.step(0)
        v, w = sim.velocities()
    v = v.reshape(M, 3)
    speed = np.linalg.norm(v, axis=1)
    assert np.ptp(speed) < 1e-6 * speed.max()
    assert np.ptp(v[:, 2]) < 1e-6 * np.abs(v[:, 2]).max()


@pytest.mark.parametrize("solver,M", [("gmres", 60), ("bicgstab", 60), ("direct", 60), ("direct", 230),
                                      ("mixed-bf16", 60), ("mixed-fp16", 60), ("mixed-bf16", 230),
                                      ("mixed-fp16", 230), ("gmres-ir", 60), ("gmres-ir", 230)])
def test_solvers_agree_with_oracle(oracle64, solver, M):
    """x of each solver against the oracle's LU solution of the oracle's system
    (n = 900 and 3450: one and several 1024-column outer LU panels with ragged
    edges)."""
    from oracle.pyoracle import make_params

    pos, ori = random_scene(M, side=5.0 * (M / 60) ** (1 / 3), seed=3, min_dist=0.4)
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-9
    params.BICGSTAB_TOLERANCE = 1e-9
    with _sim(params, pos, ori, solver=solver) as sim:
        sim.assemble_system()
        sim.solve_system()
        _, _, x = sim.system()
        st = sim.solver_stats()
    A, b = oracle64.assemble(make_params(M), pos, ori)
    x_ref = oracle64.solve_direct(A, b)
    print(solver, st, rel_err(x, x_ref))
    assert st["converged"]
    assert rel_err(x, x_ref) < 1e-6


def test_lu_modes_agree_across_panels():
    """n = 4500 (five 1024-column panels, the last one ragged, taken in
    groups of three): the FP32 LU and the BF16 / FP16 LUs with
    per-block-scaled packs give the same refined solution.  The first group
    updates its second and third panels' rows of the columns right of it
    (K = 1024, 2048), then the trailing matrix once (K = 3072); the next
    group reaches the end, so that update is not split: 3 timed launches.
    No solve needs the GMRES fallback."""
    M = 300
    pos, ori = random_scene(M, side=5.0 * (M / 60) ** (1 / 3), seed=8, min_dist=0.4)
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    out = {}
    for solver in ("direct", "mixed-bf16", "mixed-fp16"):
        with _sim(params, pos, ori, solver=solver, refinement_tolerance=1e-11) as sim:
            sim.step(0)
            _, _, x = sim.system()
            st, tm = sim.solver_stats(), sim.timings()
        print(solver, st, tm["gemm_launches"], tm["direct_fallback"])
        assert st["converged"] and st["relative_residual"] <= 1e-11
        assert tm["direct_fallback"] == 0
        out[solver] = (x, tm["gemm_launches"])
    n, W, G = 15 * M, 1024, 3
    GW = G * W
    launches = sum((G - 1) + 1 + (K0 + 2 * GW < n) for K0 in range(0, n, GW) if K0 + GW < n)
    assert launches == 3
    for solver, (x, nl) in out.items():
        assert nl == launches
        e = rel_err(x, out["direct"][0])
        print(f"{solver} against the FP32 LU: x {e:.1e}")
        assert e < 1e-9


@pytest.mark.parametrize("matrix_format", ["theta", "dense"])
def test_gmres_stops_on_the_cycle_estimate(matrix_format):
    """A restart cycle whose own residual estimate meets the tolerance ends
    the solve after its update of x (lib/viennacl/linalg/gmres.hpp:330-335);
    confirm_residual computes the true residual of that x first, as
    fortran/sgmres.f does with icntl(8) = 1.  The same x either way, one
    matvec more per solve with confirmation, and the estimate is the true
    residual to far below the tolerance."""
    M = 120
    pos, ori = random_scene(M, side=6.0, seed=23, min_dist=0.4)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-8
    steps = 3
    out = {}
    for confirm in (False, True):
        with _sim(params, pos, ori, solver="gmres", integration="analytical", matrix_format=matrix_format,
                  confirm_residual=confirm) as sim:
            res, matvecs = [], 0
            for t in range(steps):
                sim.step(t)
                st = sim.solver_stats()
                assert st["converged"] and st["iterations"] > 0
                res.append((st["iterations"], st["relative_residual"]))
                mv = sim.timings()["gemv_launches"]  # the step's own count
                # the steps' products, one true residual per restart (and per
                # warm start on the dense form), one per solve to confirm;
                # the products of steps after a cycle's end are not counted
                its, m = st["iterations"], params.GMRES_RESTART
                assert mv == its + (its - 1) // m + (t > 0 and matrix_format == "dense") + confirm
                matvecs += mv
            out[confirm] = (sim.state(), res, matvecs)
    (s0, r0, g0), (s1, r1, g1) = out[False], out[True]
    print(f"{matrix_format}: matvecs {g0} / {g1}, residuals (estimate / true) "
          + ", ".join(f"{a[1]:.6e} / {b[1]:.6e}" for a, b in zip(r0, r1)))
    assert np.array_equal(s0[0], s1[0]) and np.array_equal(s0[1], s1[1])
    assert [a[0] for a in r0] == [b[0] for b in r1]
    assert g1 - g0 == steps
    for (_, est), (_, true) in zip(r0, r1):
        assert est <= 1e-8 and true <= 1e-8
        assert abs(est - true) <= 1e-3 * 1e-8


def test_bicgstab_reports_its_recurrence_residual():
    """A converged BiCGStab solve reports its recurrence residual, as the
    reference's solver does (lib/viennacl/linalg/bicgstab.hpp:146-160);
    confirm_residual adds the true residual of the returned x (one matvec
    more per solve) and changes nothing else."""
    M = 120
    pos, ori = random_scene(M, side=6.0, seed=29, min_dist=0.4)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.BICGSTAB_TOLERANCE = 1e-8
    steps = 3
    out = {}
    for confirm in (False, True):
        with _sim(params, pos, ori, solver="bicgstab", confirm_residual=confirm) as sim:
            res, matvecs = [], []
            for t in range(steps):
                sim.step(t)
                st = sim.solver_stats()
                assert st["converged"] and st["iterations"] > 0
                res.append((st["iterations"], st["relative_residual"]))
                matvecs.append(sim.timings()["gemv_launches"])
            out[confirm] = (sim.state(), res, matvecs)
    (s0, r0, g0), (s1, r1, g1) = out[False], out[True]
    print("bicgstab: matvecs", g0, "/", g1, "residuals (recurrence / true)",
          ", ".join(f"{a[1]:.6e} / {b[1]:.6e}" for a, b in zip(r0, r1)))
    assert np.array_equal(s0[0], s1[0]) and np.array_equal(s0[1], s1[1])
    assert [a[0] for a in r0] == [b[0] for b in r1]
    assert [b - a for a, b in zip(g0, g1)] == [1] * steps
    for (_, rec), (_, true) in zip(r0, r1):
        assert rec <= 1e-8 and true <= 1e-8 * 10
        assert abs(rec - true) <= 1e-2 * 1e-8


@pytest.mark.parametrize("matrix_format", ["theta", "dense"])
def test_gmres_restart_cycles_and_iteration_cap(matrix_format):
    """Restart cycles of 3 steps, each enqueued at once with its control on
    the device: the solve reaches a tight true residual over many cycles
    (confirmed, so the reported residual is ||b - A x|| / ||b|| itself) and
    agrees with the FP64-refined LU solution to the FP32 assembly's rounding;
    capped at 5 steps, it stops after a full cycle and a 2-step one,
    unconverged, with the true residual of that x."""
    M = 40
    pos, ori = random_scene(M, side=5.0, seed=31, min_dist=0.4)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_RESTART = 3
    params.GMRES_TOLERANCE = 1e-11
    with _sim(params, pos, ori, solver="mixed-bf16", refinement_tolerance=1e-12) as sim:
        sim.step(0)
        _, _, x_ref = sim.system()
    with _sim(params, pos, ori, solver="gmres", matrix_format=matrix_format, confirm_residual=True) as sim:
        sim.step(0)
        _, _, x = sim.system()
        st, mv = sim.solver_stats(), sim.timings()["gemv_launches"]
    its = st["iterations"]
    print(f"{matrix_format}: {its} steps, {mv} products, residual {st['relative_residual']:.2e}, "
          f"x vs LU {rel_err(x, x_ref):.1e}")
    assert st["converged"] and st["relative_residual"] <= 1e-11 and its > 6
    assert mv == its + (its - 1) // 3 + 1  # steps, restarts, confirmation
    assert rel_err(x, x_ref) < 1e-5
    params.GMRES_MAX_ITERATIONS = 5
    with _sim(params, pos, ori, solver="gmres", matrix_format=matrix_format) as sim:
        sim.step(0)
        st, mv = sim.solver_stats(), sim.timings()["gemv_launches"]
    assert st["iterations"] == 5 and not st["converged"] and st["relative_residual"] > 1e-11
    assert mv == 5 + 2  # the products of the steps, and the true residual after each cycle


def test_gmres_ir_uses_the_fp16_copy():
    """GMRES-IR runs its inner iterations on the FP16 copy and reaches the
    outer tolerance on the FP32 system; its states follow plain GMRES."""
    M = 120
    pos, ori = random_scene(M, side=6.0, seed=17, min_dist=0.4)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-10
    params.GMRES_RESTART = 30
    out = {}
    for solver in ("gmres", "gmres-ir"):
        with _sim(params, pos, ori, solver=solver) as sim:
            for t in range(3):
                sim.step(t)
            st, tm = sim.solver_stats(), sim.timings()
            out[solver] = sim.state()
        assert st["converged"] and st["relative_residual"] <= 1e-10
        if solver == "gmres-ir":
            assert tm["gemv_half_launches"] > 0 and tm["gemv_launches"] >= 2
    ep, eo = rel_err(out["gmres-ir"][0], out["gmres"][0]), rel_err(out["gmres-ir"][1], out["gmres"][1])
    print(f"gmres-ir vs gmres after 3 steps: pos {ep:.1e} ori {eo:.1e}")
    assert ep < 1e-9 and eo < 1e-9


def test_ten_steps_against_oracle(oracle64):
    """Free-running 10 steps of the reference scene: positions/orientations
    within 1e-5 of the FP64 oracle (GMRES at 1e-8 vs oracle LU)."""
    from oracle.pyoracle import make_params

    M, pos, ori = ref100()
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-8
    with _sim(params, pos, ori) as sim:
        for t in range(10):
            sim.step(t)
        p, o = sim.state()
    p_ref, o_ref = oracle64.simulate(make_params(M), pos, ori, 10, solver="direct")
    ep, eo = rel_err(p, p_ref), rel_err(o, o_ref)
    print(f"10 steps: pos {ep:.1e} ori {eo:.1e}")
    assert ep < 1e-5 and eo < 1e-5


def test_cube_m100_ten_steps(oracle64):
    """M = 100 centres uniform in [-5, 5]^3 (min distance 0.4), numerical
    integrals, GMRES to 1e-6, 10 steps: the state against the FP64 oracle's
    direct solves."""
    from oracle.pyoracle import make_params
    from rigidfibers_amd.scenes import cube

    M = 100
    pos, ori = cube(M, 10.0, min_dist=0.4, seed=100)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-6
    with _sim(params, pos, ori, solver="gmres", integration="numerical") as sim:
        for t in range(10):
            sim.step(t)
            assert sim.solver_stats()["converged"]
        p, o = sim.state()
    p_ref, o_ref = oracle64.simulate(make_params(M), pos, ori, 10, solver="direct")
    ep, eo = rel_err(p, p_ref), rel_err(o, o_ref)
    print(f"cube M=100, 10 steps: pos {ep:.1e} ori {eo:.1e}")
    assert ep < 1e-5 and eo < 1e-5


@pytest.mark.parametrize("analytical", [False, True])
def test_cube_m2000_gmres_and_lu_trajectories(analytical):
    """M = 2000 in a cube of side 40 (n = 30 000), 20 steps: GMRES to 1e-8 and
    the FP32 LU with FP64 refinement (its factors from the expanded theta
    form) follow the same trajectory."""
    from rigidfibers_amd.scenes import cube

    M = 2000
    pos, ori = cube(M, 40.0, min_dist=0.4, seed=2020)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-8
    integration = "analytical" if analytical else "numerical"
    out = {}
    for solver in ("gmres", "direct"):
        with _sim(params, pos, ori, solver=solver, integration=integration, refinement_tolerance=1e-10) as sim:
            for t in range(20):
                sim.step(t)
                assert sim.solver_stats()["converged"]
            out[solver] = sim.state()
    ep, eo = rel_err(out["gmres"][0], out["direct"][0]), rel_err(out["gmres"][1], out["direct"][1])
    print(f"cube M=2000 {integration}, 20 steps, GMRES vs LU: pos {ep:.1e} ori {eo:.1e}")
    assert ep < 1e-7 and eo < 1e-6


@pytest.mark.parametrize("solver", ["direct", "mixed-bf16"])
@pytest.mark.parametrize("integration", ["numerical", "analytical"])
def test_direct_solvers_on_theta_and_dense_forms(solver, integration):
    """The direct solvers factor the expanded theta form (the default) or a
    copy of the dense matrix: the two systems differ by the FP32 rounding of
    the two assembly kernels, the dumps agree to that, each solve meets its
    own system's refinement tolerance (seen through the FP32 dump), and the
    states after 3 steps agree."""
    M = 150
    pos, ori = random_scene(M, side=7.0, seed=31, min_dist=0.4)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    out = {}
    for fmt in ("auto", "dense"):
        with _sim(params, pos, ori, solver=solver, integration=integration, matrix_format=fmt,
                  refinement_tolerance=1e-10) as sim:
            sim.assemble_system()
            sim.solve_system()
            st = sim.solver_stats()
            A, b, x = sim.system(with_matrix=True)
            res = np.linalg.norm(b - A.astype(np.float64) @ x) / np.linalg.norm(b)
            # the dense dump is the system itself; the theta form's dump rounds
            # the blocks its matrix-vector product forms in FP64 to FP32
            assert st["converged"] and res <= (1e-9 if fmt == "dense" else 1e-7), (fmt, st, res)
            sim.update_velocities()
            sim.update_fibers(True)
            sim.advance()
            for t in range(1, 3):
                sim.step(t)
            out[fmt] = (A, x, sim.state())
    eA = rel_err(out["auto"][0], out["dense"][0])
    ex = rel_err(out["auto"][1], out["dense"][1])
    ep, eo = rel_err(out["auto"][2][0], out["dense"][2][0]), rel_err(out["auto"][2][1], out["dense"][2][1])
    print(f"{solver} {integration} theta vs dense: A {eA:.1e} x {ex:.1e} pos {ep:.1e} ori {eo:.1e}")
    assert eA < 1e-6 and ex < 1e-5 and ep < 1e-7 and eo < 1e-6


@pytest.mark.parametrize("solver", ["gmres", "bicgstab", "gmres-ir"])
def test_extrapolated_guess(oracle64, solver):
    """Starting each solve from 2 x_n - x_{n-1}: every step converges, and
    after 10 steps the state is the one the oracle's direct solves give."""
    from oracle.pyoracle import make_params

    M, pos, ori = ref100()
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-8
    params.BICGSTAB_TOLERANCE = 1e-8
    with _sim(params, pos, ori, solver=solver, initial_guess="extrapolated") as sim:
        for t in range(10):
            sim.step(t)
            assert sim.solver_stats()["converged"]
        p, o = sim.state()
    p_ref, o_ref = oracle64.simulate(make_params(M), pos, ori, 10, solver="direct")
    ep, eo = rel_err(p, p_ref), rel_err(o, o_ref)
    print(f"{solver} extrapolated, 10 steps: pos {ep:.1e} ori {eo:.1e}")
    assert ep < 1e-5 and eo < 1e-5


@pytest.mark.parametrize("integration,solver,N,nqi,M", [("numerical", "gmres", 5, 8, 120),
                                                         ("analytical", "gmres", 5, 8, 120),
                                                         ("numerical", "bicgstab", 5, 8, 120),
                                                         ("analytical", "gmres-ir", 5, 8, 120),
                                                         ("numerical", "gmres", 6, 4, 120),
                                                         ("analytical", "gmres", 5, 8, 4500)])
def test_residual_from_assembly(integration, solver, N, nqi, M):
    """The first residual b - A x0 of a warm-started solve, formed by the
    theta assembly kernel (each workgroup reads back the rows it stored and
    multiplies them with x0), gives the solve that the separate matvec gives.
    The iterations are capped, so a wrong residual could not be repaired by a
    later restart: one warm step must agree to rounding in solution,
    velocities and reported residual, with one matvec fewer.  M = 120 splits
    the sources over 15 workgroups per target (partial products summed),
    M = 4500 gives one workgroup per target as at the large sizes."""
    from rigidfibers_amd.scenes import cube

    pos, ori = cube(M, 6.0 * (M / 120) ** (1 / 3), min_dist=0.4, seed=5)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.NUMBER_OF_TERMS_IN_FORCE_EXPANSION = N
    params.NUMBER_OF_QUADRATURE_INTERVALS = nqi
    params.GMRES_RESTART = 10
    params.GMRES_MAX_ITERATIONS = 3
    params.BICGSTAB_MAX_ITERATIONS = 2
    out = {}
    for separate in (False, True):
        with _sim(params, pos, ori, solver=solver, integration=integration, separate_residual=separate) as sim:
            sim.step(0)  # cold start: no residual product
            sim.step(1)
            _, _, x = sim.system()
            tv, rv = sim.velocities()
            out[separate] = (x.copy(), tv.copy(), rv.copy(), sim.solver_stats(), sim.timings()["gemv_launches"])
    (x0, t0, r0, s0, g0), (x1, t1, r1, s1, g1) = out[False], out[True]
    ex, ev, ew = rel_err(x0, x1), rel_err(t0, t1), rel_err(r0, r1)
    print(f"{integration} {solver} N={N} nqi={nqi} M={M}: x {ex:.1e} v {ev:.1e} w {ew:.1e}, residuals "
          f"{s0['relative_residual']:.3e} / {s1['relative_residual']:.3e}, matvecs {g0} / {g1}")
    assert ex < 1e-9 and ev < 1e-9 and ew < 1e-9
    assert s0["iterations"] == s1["iterations"]
    assert abs(s0["relative_residual"] - s1["relative_residual"]) <= 1e-6 * s1["relative_residual"]
    assert g0 == g1 - 1


@pytest.mark.parametrize("analytical", [False, True])
def test_theta_form_matches_dense(analytical):
    """The theta-form system (the iterative solvers' storage) expands to the
    dense assembly and gives the same GMRES solution and state, to FP32
    rounding (the theta form projects G onto the Legendre weights on the
    matrix cores, the dense form with vector FMAs in another order)."""
    M = 50
    pos, ori = random_scene(M, side=4.5, seed=21, min_dist=0.4)
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    params.GMRES_TOLERANCE = 1e-10
    out = {}
    for fmt in ("dense", "theta"):
        with _sim(params, pos, ori, integration="analytical" if analytical else "numerical",
                  matrix_format=fmt) as sim:
            sim.assemble_system()
            A, b, _ = sim.system(with_matrix=True)
            sim.solve_system()
            _, _, x = sim.system()
            sim.update_velocities()
            sim.update_fibers(True)
            p, o = sim.next_state()
        out[fmt] = (A, b, x, p, o)
    Ad, bd, xd, pd, od = out["dense"]
    At, bt, xt, pt, ot = out["theta"]
    print("A", rel_err(At, Ad), "b", rel_err(bt, bd), "x", rel_err(xt, xd), "pos", rel_err(pt, pd))
    assert rel_err(At, Ad) < 1e-6 and rel_err(bt, bd) < 1e-6
    assert rel_err(xt, xd) < 1e-6 and rel_err(pt, pd) < 1e-8 and rel_err(ot, od) < 1e-7


@pytest.mark.parametrize("solver,integration", [("gmres", "analytical"), ("gmres", "numerical"),
                                                ("bicgstab", "numerical"), ("direct", "numerical"),
                                                ("mixed-bf16", "analytical"), ("gmres-ir", "numerical")])
def test_coincident_fibers_fail_cleanly(solver, integration):
    """Two coincident fibers make the kernel integrals infinite: the step
    fails with RF_ERR_NONFINITE (a singular LU may report RF_ERR_SINGULAR)
    instead of returning NaN, and the same library then steps a valid
    scene."""
    from rigidfibers_amd import _lib

    M = 20
    pos, ori = random_scene(M, side=4.0, seed=41, min_dist=0.4)
    pos, ori = pos.copy(), ori.copy()
    pos[3:6], ori[3:6] = pos[0:3], ori[0:3]  # fiber 1 on top of fiber 0
    params = ref_config()
    params.NUMBER_OF_FIBERS = M
    allowed = {_lib.RF_ERR_NONFINITE} | ({_lib.RF_ERR_SINGULAR} if solver in ("direct", "mixed-bf16") else set())
    with _sim(params, pos, ori, solver=solver, integration=integration) as sim:
        with pytest.raises(_lib.RfError) as err:
            sim.step(0)
    print(solver, integration, err.value)
    assert err.value.code in allowed
    pos2, ori2 = random_scene(M, side=4.0, seed=41, min_dist=0.4)
    with _sim(params, pos2, ori2, solver=solver, integration=integration) as sim:
        sim.step(0)
        p, o = sim.state()
    assert np.isfinite(p).all() and np.isfinite(o).all()


@pytest.mark.parametrize("restart", [1, 1024])
def test_gmres_restart_extremes(restart):
    """GMRES(1), a cycle of one step, and GMRES(1024), the largest restart,
    longer than the n = 300 unknowns of the scene: both reach a tight true
    residual and the FP64-refined LU solution of the same dense system."""
    M = 20
    pos, ori = random_scene(M, side=4.0, seed=43, min_dist=0.4)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_RESTART = restart
    params.GMRES_TOLERANCE = 1e-10
    with _sim(params, pos, ori, solver="mixed-bf16", matrix_format="dense", refinement_tolerance=1e-12) as sim:
        sim.step(0)
        _, _, x_ref = sim.system()
    with _sim(params, pos, ori, solver="gmres", matrix_format="dense", confirm_residual=True) as sim:
        sim.step(0)
        _, _, x = sim.system()
        st, mv = sim.solver_stats(), sim.timings()["gemv_launches"]
    its = st["iterations"]
    print(f"restart {restart}: {its} steps, {mv} products, residual {st['relative_residual']:.2e}, "
          f"x vs LU {rel_err(x, x_ref):.1e}")
    assert st["converged"] and st["relative_residual"] <= 1e-10
    assert mv == its + (its - 1) // restart + 1  # steps, restarts, confirmation
    assert rel_err(x, x_ref) < 1e-8


def _near_coincident_scene(d):
    """64 fibers (n = 960, fifteen 64-row pivot blocks) with fiber 40 moved to
    a distance d of fiber 0, parallel to it: rows 0-14 and 600-614 of A,
    which lie in different pivot blocks, couple with entries ~ 1/d.  Partial
    pivoting over the whole column (sgesv, fortran/solve_system.incl:3) would
    exchange rows across blocks; the pivots here stay within each block, so
    the factors carry an element growth ~ 1/d that the refinement, or the
    GMRES fallback, has to overcome."""
    M = 64
    pos, ori = random_scene(M, side=5.0, seed=3, min_dist=0.4)
    pos, ori = pos.reshape(M, 3).copy(), ori.reshape(M, 3).copy()
    perp = np.cross(ori[0], [1.0, 0.0, 0.0])
    perp /= np.linalg.norm(perp)
    pos[40] = pos[0] + d * perp
    ori[40] = ori[0]
    return M, pos.ravel(), ori.ravel()


@pytest.mark.parametrize("solver", ["direct", "mixed-bf16", "mixed-fp16"])
@pytest.mark.parametrize("d,max_sweeps", [(1e-3, 50), (1e-4, 50), (1e-4, 1)])
def test_direct_solve_without_cross_block_pivoting(oracle64, solver, d, max_sweeps):
    """Systems that need cross-block row exchanges.  d = 1e-3 (entries up to
    5e3, condition number 1.5e4): the in-block LU's element growth is
    overcome by the FP64 refinement: x is the solution of the FP32 system it
    solved (against an LU with full partial pivoting of that system, in
    FP64) and matches the oracle's to what the FP32 rounding of the assembly
    leaves at that condition number.  d = 1e-4 (entries up to 5e6, condition
    number 1.5e7, past what an FP32-stored system resolves): the step either
    converges or reports RF_WARN_NOT_CONVERGED, and the residual it reports is
    the true one of the returned x in the system it solved -- never a
    silently wrong x; capped at one refinement sweep the solve hands over to
    GMRES on the dense system (rf_timings.direct_fallback)."""
    from oracle.pyoracle import make_params
    from rigidfibers_amd import _lib

    M, pos, ori = _near_coincident_scene(d)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.GMRES_MAX_ITERATIONS = 2000
    A_ref, b_ref = oracle64.assemble(make_params(M), pos, ori, False)
    x_ref = oracle64.solve_direct(A_ref, b_ref)
    with _sim(params, pos, ori, solver=solver, matrix_format="dense", refinement_tolerance=1e-10,
              refinement_max_iterations=max_sweeps) as sim:
        code = sim.step(0)
        A, b, x = sim.system(with_matrix=True)
        st, tm = sim.solver_stats(), sim.timings()
    e = rel_err(x, x_ref)
    A64 = A.astype(np.float64)
    res = np.linalg.norm(b - A64 @ x) / np.linalg.norm(b)
    e_sys = rel_err(x, np.linalg.solve(A64, b))
    print(f"{solver} d={d:g} max sweeps {max_sweeps}: status {code}, {st}, fallback {tm['direct_fallback']}, "
          f"residual of x in the solved system {res:.1e}, x against its solution {e_sys:.1e}, "
          f"against the oracle {e:.1e}")
    if max_sweeps == 1:
        assert tm["direct_fallback"] == 1
    if d == 1e-3:
        assert code == _lib.RF_OK and st["converged"] and tm["direct_fallback"] == 0
        assert e_sys <= 1e-8 and e <= 1e-4
    if st["converged"]:
        assert code == _lib.RF_OK and st["relative_residual"] <= 1e-10 and res <= 1e-9
    else:
        assert code == _lib.RF_WARN_NOT_CONVERGED and abs(res - st["relative_residual"]) <= 1e-3 * res


def test_bicgstab_restarts_from_the_true_residual():
    """BiCGStab restarts from the true residual every 200 iterations and on
    a degenerate search direction, as ViennaCL's solver does
    (lib/viennacl/linalg/bicgstab.hpp:119-131,150-152).  On the
    near-coincident system of condition number 1.5e7 it needs more than 200
    iterations: the solve restarts, and either converges -- its true
    residual then meets the tolerance -- or reports non-convergence with the
    true residual of the returned x."""
    from rigidfibers_amd import _lib

    M, pos, ori = _near_coincident_scene(1e-4)
    params = ref_config(validate=False)
    params.NUMBER_OF_FIBERS = M
    params.BICGSTAB_TOLERANCE = 1e-9
    params.BICGSTAB_MAX_ITERATIONS = 1000
    with _sim(params, pos, ori, solver="bicgstab", matrix_format="dense", confirm_residual=True) as sim:
        code = sim.step(0)
        A, b, x = sim.system(with_matrix=True)
        st, tm = sim.solver_stats(), sim.timings()
    res = np.linalg.norm(b - A.astype(np.float64) @ x) / np.linalg.norm(b)
    print(f"bicgstab: status {code}, {st}, restarts {tm['solver_restarts']}, true residual {res:.1e}")
    # periodic restarts before iterations 202, 404, ... (0-based), and any degenerate ones
    assert st["iterations"] > 202 and tm["solver_restarts"] >= (st["iterations"] - 1) // 202
    assert abs(res - st["relative_residual"]) <= 1e-3 * res
    if st["converged"]:
        assert code == _lib.RF_OK and res <= 1e-8
    else:
        assert code == _lib.RF_WARN_NOT_CONVERGED
