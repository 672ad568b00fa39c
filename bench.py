#!/usr/bin/env python3
"""Throughput of the rigid-fiber timestep on one MI355X, in fiber-steps/s.

Workload (the configuration the headline metric is quoted on): M = 5000 fibers
of length 2 with centres uniform in a ball of radius 35 (minimum centre
distance 0.4, seed 2026), orientations uniform on the sphere, slenderness 0.01,
N = 5 force terms, 8 x 3 quadrature, unit gravity, dt = 0.1, analytical inner
integrals (FP64 recurrences), GMRES(20) to a relative residual of 1e-6 starting
from the previous solution.  n = 75 000 unknowns.

One "step" is one call of the library's rf_step (assemble -> solve -> update
velocities -> update fibers) on device-resident state.  W warm-up steps are
run untimed, then exactly K steps are timed between two device
synchronisations.  value = M * K / elapsed.

Also reported:
  roofline      the step's dominant kernel, the assembly (VALU-bound): its
                algorithmic flops per launch / the mean duration of the
                assembly kernel (HIP events around it on the library's
                stream), against the peak for its FP64/FP32 mix at MI355X's
                published 78.6 / 157.3 TFLOP/s (the time the flops take at
                those rates).  Flops: numerical, 63 216 FP32 per ordered fiber
                pair; analytical, per (pair, target node): 624 FP64 (moment
                recurrences) + 400 FP32 (recombination) where the node lies
                within the far-field radius of the source's centre, else 576
                FP32 (8-point far-field rule: 8 kernel evaluations of 30 flop,
                6 mirror sums and 36 FMAs per node pair), plus 7 920 FP32 per
                pair for the projections; near and far nodes are counted on
                the initial scene.  reference_algorithm_count
                prices the reference's algorithm instead (FP64 recurrences for
                every pair: 32 500 flop per ordered pair, 15 000 of them FP64)
                against the same kernel time;
  roofline_solver  iterative solvers: the solver's matvec (HBM-bound),
                algorithmic bytes per launch (theta form: M * M * N * (3 +
                6(N-1)) * 4, the FP32 numbers that define A; dense: n^2 * 4) /
                mean launch duration, against 8 TB/s.  Direct solvers: the
                LU's trailing-update GEMMs (MFMA-bound), their 2 m n K flops /
                summed duration, against the dense MFMA peak of the mode (FP32
                157.3, BF16/FP16 2500 TFLOP/s);
  precision     plain FP32 LU against the BF16- and FP16-MFMA LUs with FP64
                refinement (to 1e-10) on the same scene: fiber-steps/s, the
                speed-up, the relative position error against the FP32-LU run
                after the same steps, and the error against the FP64 oracle:
                the residual of the last step's solution in the oracle's FP64
                rows of three groups of target fibers, and the position and
                orientation error of 10 steps on a 250-fiber ball at the same
                number density against the oracle's trajectory;
  serving       the serving stream with continuous batching against the same
                stream one request at a time: p50/p99 step and request latency;
  large         M = 10 000: GMRES-IR on the FP16 theta copy against plain
                GMRES on the FP32 theta system (speed-up, position difference);
  cpu_baseline  the FP32 OpenMP port of the reference Fortran (oracle build
                librf_oracle32.so) on a bounded sample: the rows of S target
                fibers assembled against all 5000 sources, their velocities,
                and the GMRES matvecs of those rows; scaled by M/S to a step.
                Threads: OMP_NUM_THREADS, else every CPU the process may run
                on, bound close (OMP_PROC_BIND); the CPU model is recorded.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# CPUs this process may run on, read before the oracle's OpenMP runtime binds
# the main thread to one place
CPUS_AVAILABLE = len(os.sched_getaffinity(0))
MFMA_PEAK_TFLOPS = {"direct": 157.3, "mixed-bf16": 2500.0, "mixed-fp16": 2500.0}
GEMM_KERNEL = {"direct": "gemm_packed_sub_kernel<float>", "mixed-bf16": "gemm_packed_sub_kernel<__bf16>",
               "mixed-fp16": "gemm_packed_sub_kernel<_Float16>"}
ITERATIVE = ("gmres", "bicgstab", "gmres-ir")


def build_scene(M, radius, seed):
    from rigidfibers_amd.scenes import ball

    return ball(M, radius, min_dist=0.4, seed=seed)


def run_gpu(args, pos, ori, solver, steps, warmup, tag="", capture_last=False):
    """Runs warmup + steps timesteps; with capture_last also returns the state
    the last step assembled its system from and that step's solution."""
    from rigidfibers_amd.config import Parameters
    from rigidfibers_amd.simulation import Simulation

    M = args.fibers
    params = Parameters(NUMBER_OF_FIBERS=M, TIMESTEP=0.1, SLENDERNESS=0.01, NUMBER_OF_TERMS_IN_FORCE_EXPANSION=5,
                        NUMBER_OF_QUADRATURE_INTERVALS=8, GMRES_RESTART=args.restart,
                        GMRES_MAX_ITERATIONS=1000, GMRES_TOLERANCE=args.tol, BICGSTAB_TOLERANCE=args.tol)
    matrix_format = args.matrix_format if solver in ITERATIVE else "auto"
    sim = Simulation(params, pos, ori, solver=solver, integration=args.integration, matrix_format=matrix_format,
                     velocity_operator=args.velocity_operator, refinement_tolerance=1e-10,
                     initial_guess=args.initial_guess,
                     separate_residual=args.separate_residual, confirm_residual=args.confirm_residual)
    try:
        for t in range(warmup):
            sim.step(t)
            print(f"# {tag}{solver} warmup step {t}: {sim.timings()['total']:.1f} ms, solver {sim.solver_stats()}",
                  flush=True)
        phases = {"assemble_system": 0.0, "solve_system": 0.0, "update_velocities": 0.0, "update_fibers": 0.0}
        acc = {"gemv": 0.0, "gemv_launches": 0, "gemm": 0.0, "gemm_flop": 0.0, "gemm_launches": 0,
               "assemble_kernel": 0.0}
        iters = []
        t0 = time.perf_counter()
        last = None
        for t in range(warmup, warmup + steps):
            if capture_last and t == warmup + steps - 1:
                last = sim.state()
            sim.step(t)
            tm = sim.timings()
            for k in phases:
                phases[k] += tm[k]
            for k in acc:
                acc[k] += tm[k]
            iters.append(sim.solver_stats()["iterations"])
        elapsed = time.perf_counter() - t0
        stats = sim.solver_stats()
        pos_end, ori_end = sim.state()
        if last is not None:
            last = {"pos": last[0], "ori": last[1], "x": sim.system()[2]}
        device_bytes = sim.device_bytes()
        n = sim.n
    finally:
        sim.close()
    theta = solver in ITERATIVE and matrix_format in ("auto", "theta")
    return {"elapsed": elapsed, "phases": phases, "acc": acc, "iters": iters, "stats": stats, "n": n, "theta": theta,
            "device_bytes": device_bytes, "pos": pos_end, "ori": ori_end, "steps": steps,
            "last": last}


def far_field_radius(N=5):
    """kFarField[N] of the library's device header (rf_device.h)."""
    import re

    with open(os.path.join(ROOT, "rigidfibers_amd", "csrc", "rf_device.h")) as fh:
        src = fh.read()
    vals = re.search(r"kFarField\[kMaxTerms \+ 1\] = \{([^}]*)\}", src).group(1).split(",")
    return float(vals[N])


def assembly_flops(integration, pos, ori, N=5, nqi=8):
    """(FP64, FP32) algorithmic flops of one assembly of this scene and the
    number of near (pair, node) combinations (see the module docstring)."""
    import numpy as np
    from scipy.spatial import cKDTree

    M = pos.size // 3
    pairs = float(M) * (M - 1)
    if integration == "numerical":
        return 0.0, pairs * 63_216.0, 0
    g = np.sqrt(15.0) / 5.0
    h = 2.0 / nqi
    s = np.array([(2.0 * (-1.0 + iv * h) + h + gp * h) / 2.0 for iv in range(nqi) for gp in (-g, 0.0, g)])
    P, T = pos.reshape(M, 3), ori.reshape(M, 3)
    X = (P[:, None, :] + s[None, :, None] * T[:, None, :]).reshape(-1, 3)
    # source centres within the radius of each target node, less the node's own fiber
    near = int(cKDTree(P).query_ball_point(X, r=far_field_radius(N), return_length=True).sum()) - M * s.size
    far = pairs * s.size - near
    return 624.0 * near, 400.0 * near + 576.0 * far + 7_920.0 * pairs, near


def step_bound(args, g):
    """The step at its bounds: the assembly's algorithmic flops at the
    published FP64 / FP32 peaks, every matvec of the step and the stored
    velocity pass at 8 TB/s."""
    if not (args.solver == "gmres" and g["theta"]):
        return None
    M, N, K = args.fibers, 5, g["steps"]
    f64, f32, _ = g["asm_flops"]
    asm_ms = 1e3 * (f64 / (FP64_PEAK_TFLOPS * 1e12) + f32 / (FP32_PEAK_TFLOPS * 1e12))
    mv_ms = 1e3 * M * (3 + 6 * (N - 1)) * M * N * 4.0 / (HBM_PEAK_GBS * 1e9)
    vel_ms = 1e3 * M * 12 * M * N * 4.0 / (HBM_PEAK_GBS * 1e9)
    matvecs = g["acc"]["gemv_launches"] / K
    total = asm_ms + matvecs * mv_ms + vel_ms
    return {"assembly_ms": round(asm_ms, 2), "matvec_ms": round(mv_ms, 3), "matvecs_per_step": round(matvecs, 2),
            "velocity_ms": round(vel_ms, 3), "ms_per_step": round(total, 2),
            "fiber_steps_per_s": round(M / (total * 1e-3), 1),
            "achieved_frac": round((M * K / g["elapsed"]) / (M / (total * 1e-3)), 3)}


def measured_traffic(kernel, M, N):
    """HBM bytes per launch from the committed FETCH_SIZE/WRITE_SIZE counter
    passes (scaled by the copy-kernel calibration), or None for other shapes."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh).get(f"{kernel} M={M} N={N}")
    except (OSError, ValueError):
        return None
    return None if rec is None else float(rec["fetch_bytes"] + rec["write_bytes"])


# MI355X's published FP64 and FP32 rates (the same for vector and matrix
# instructions).
FP64_PEAK_TFLOPS, FP32_PEAK_TFLOPS = 78.6, 157.3


def roofline_assembly(args, g):
    """The assembly kernel against the time its algorithmic flops take at the
    published FP64 / FP32 peaks."""
    M, K = args.fibers, g["steps"]
    f64, f32, near = g["asm_flops"]
    pairs = float(M) * (M - 1)
    flop = f64 + f32
    peak_s = f64 / (FP64_PEAK_TFLOPS * 1e12) + f32 / (FP32_PEAK_TFLOPS * 1e12)
    avg_ms = g["acc"]["assemble_kernel"] / K
    achieved = flop / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    peak = flop / peak_s / 1e12
    theta = g["theta"]
    kernel = f"assemble_{'theta_' if theta else ''}kernel<5, 8, {'true' if args.integration == 'analytical' else 'false'}>"
    # the reference algorithm's count (FP64 moment recurrences for every pair):
    # 32 500 flop per ordered pair, 15 000 of them FP64
    ref64, ref32 = 15_000.0 * pairs, 17_500.0 * pairs
    ref_peak_s = ref64 / (FP64_PEAK_TFLOPS * 1e12) + ref32 / (FP32_PEAK_TFLOPS * 1e12)
    ref = None
    if args.integration == "analytical" and avg_ms > 0:
        ref = {"flop_per_launch": ref64 + ref32, "fp64_flop_per_launch": ref64,
               "time_at_peak_ms": round(1e3 * ref_peak_s, 3), "frac": round(ref_peak_s / (avg_ms * 1e-3), 4)}
    return {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 2), "peak": round(peak, 2),
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4) if peak > 0 else 0.0,
            "reference_algorithm_count": ref,
            "traffic": measured_traffic(kernel.split("<")[0], M, 5) if theta else None,
            "flop_per_launch": flop, "fp64_flop_per_launch": f64, "fp32_flop_per_launch": f32,
            "near_pair_nodes": near, "pair_nodes": pairs * 24,
            "time_at_peak_ms": round(1e3 * peak_s, 3), "avg_launch_ms": round(avg_ms, 4)}


def roofline(args, solver, g):
    M, N = args.fibers, 5
    acc = g["acc"]
    if solver in ITERATIVE:
        theta = args.matrix_format == "theta" or args.matrix_format == "auto"
        if theta:  # 3 + 6(N-1) FP32 numbers per (target, source, source mode)
            per_launch, kernel = float(M) * (3 + 6 * (N - 1)) * M * N * 4.0, "gemv_theta_kernel"
        else:      # n^2 FP32 entries
            per_launch, kernel = float(g["n"]) ** 2 * 4.0, "gemv_f32_kernel"
        avg_ms = acc["gemv"] / max(acc["gemv_launches"], 1)
        achieved = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel, M, N),
                "bytes_per_launch": per_launch, "avg_launch_ms": round(avg_ms, 4)}
    peak = MFMA_PEAK_TFLOPS[solver]
    achieved = acc["gemm_flop"] / (acc["gemm"] * 1e-3) / 1e12 if acc["gemm"] > 0 else 0.0
    kernel = GEMM_KERNEL[solver]
    return {"bound": "mfma", "kernel": kernel, "achieved": round(achieved, 1), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": None,
            "flop_per_step": acc["gemm_flop"] / g["steps"], "launches_per_step": acc["gemm_launches"] / g["steps"],
            "avg_launch_ms": round(acc["gemm"] / max(acc["gemm_launches"], 1), 4)}


def rel_err(a, b):
    import numpy as np

    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b))))


def precision_study(args, pos, ori, plain=None):
    """Plain FP32 LU vs the BF16/FP16-MFMA LUs with FP64 refinement, same scene
    and steps; and, beside the headline's plain FP32 GMRES run (`plain`),
    GMRES-IR with its inner solves on the FP16 copy of the theta-form system
    (same steps and tolerance)."""
    out, ref = {}, None
    check = oracle_check(args) if args.oracle_check else {}
    for solver in ("direct", "mixed-bf16", "mixed-fp16"):
        g = run_gpu(args, pos, ori, solver, args.lu_steps, 1, tag="precision ", capture_last=args.oracle_check)
        rate = args.fibers * g["steps"] / g["elapsed"]
        entry = {"fiber_steps_per_s": round(rate, 2), "ms_per_step": round(1e3 * g["elapsed"] / g["steps"], 2),
                 "solve_ms_per_step": round(g["phases"]["solve_system"] / g["steps"], 2),
                 "refinement_sweeps": g["iters"], "relative_residual": g["stats"]["relative_residual"],
                 "roofline": roofline(args, solver, g)}
        if ref is None:
            ref = (g["pos"], g["ori"], rate)
        else:
            entry["speedup_vs_fp32_lu"] = round(rate / ref[2], 3)
            entry["pos_rel_err_vs_fp32_lu"] = rel_err(g["pos"], ref[0])
            entry["ori_rel_err_vs_fp32_lu"] = rel_err(g["ori"], ref[1])
        if args.oracle_check:
            entry["fp64_oracle"] = {"rows_m5000": oracle_rows(args, g["last"]), "ten_steps_m250": check[solver]}
        out["plain_fp32_lu" if solver == "direct" else solver.replace("-", "_")] = entry
    out["steps"] = args.lu_steps
    out["warmup"] = 1
    if plain is not None:
        g = run_gpu(args, pos, ori, "gmres-ir", args.steps, args.warmup, tag="precision ")
        rate = args.fibers * g["steps"] / g["elapsed"]
        plain_rate = args.fibers * plain["steps"] / plain["elapsed"]
        out["gmres_ir_fp16"] = {
            "fiber_steps_per_s": round(rate, 2), "ms_per_step": round(1e3 * g["elapsed"] / g["steps"], 2),
            "solve_ms_per_step": round(g["phases"]["solve_system"] / g["steps"], 2),
            "plain_solve_ms_per_step": round(plain["phases"]["solve_system"] / plain["steps"], 2),
            "iterations_per_step": g["iters"], "steps": g["steps"],
            "speedup_vs_plain_fp32_gmres": round(rate / plain_rate, 3),
            "pos_rel_err_vs_plain_fp32_gmres": rel_err(g["pos"], plain["pos"]),
            "ori_rel_err_vs_plain_fp32_gmres": rel_err(g["ori"], plain["ori"])}
    return out


def oracle_rows(args, last):
    """The last step's solution in the FP64 oracle's rows of three groups of
    three target fibers (fortran/assemble_matrix.incl restated in FP64, at the
    state that step assembled from): ||b - A x|| / ||b|| over those rows."""
    import numpy as np

    from oracle.pyoracle import Oracle, make_params

    o = Oracle(64)
    M = args.fibers
    p = make_params(M, threads=cpu_threads())
    res = []
    for i0 in (0, M // 2 - 1, M - 3):
        A, b = o.assemble_rows(p, last["pos"], last["ori"], i0, i0 + 3, args.integration == "analytical")
        res.append(float(np.linalg.norm(b - A @ last["x"]) / np.linalg.norm(b)))
    return {"targets": [0, M // 2 - 1, M - 3], "fibers_per_group": 3, "relative_residual": res}


def oracle_check(args, M=250, steps=10):
    """Each LU form, 10 steps on a ball of M fibers at the number density of
    the headline ball (same generator, integrals and tolerances), against the
    FP64 oracle's trajectory (its GMRES at 1e-13, which is its LU solution):
    the largest relative position and orientation error over the steps."""
    import numpy as np

    from oracle.pyoracle import Oracle, make_params
    from rigidfibers_amd.config import Parameters
    from rigidfibers_amd.simulation import Simulation

    radius = args.radius * (M / args.fibers) ** (1.0 / 3.0)
    pos, ori = build_scene(M, radius, args.seed)
    rec = []
    Oracle(64).simulate(make_params(M, threads=cpu_threads()), pos, ori, steps, solver="gmres",
                        analytical=args.integration == "analytical", gmres=(30, 2000, 1e-13), record=rec)
    params = Parameters(NUMBER_OF_FIBERS=M, TIMESTEP=0.1, SLENDERNESS=0.01, NUMBER_OF_TERMS_IN_FORCE_EXPANSION=5,
                        NUMBER_OF_QUADRATURE_INTERVALS=8)
    out = {}
    for solver in ("direct", "mixed-bf16", "mixed-fp16"):
        ep = eo = 0.0
        with Simulation(params, pos, ori, solver=solver, integration=args.integration,
                        refinement_tolerance=1e-10) as sim:
            for t in range(steps):
                sim.step(t)
                p, o = sim.state()
                ep, eo = max(ep, rel_err(p, rec[t]["pos"])), max(eo, rel_err(o, rec[t]["ori"]))
        out[solver] = {"workload": f"M={M} ball of radius {radius:.2f}, {steps} steps",
                       "pos_rel_err": ep, "ori_rel_err": eo}
        print(f"# oracle check {solver}: M={M}, {steps} steps: positions {ep:.2e}, orientations {eo:.2e}",
              flush=True)
    return out


def serving(args):
    """The serving stream: suspensions with M ~ U[16, 400] from the cube
    generator at number density 0.1 (own seed each), 20 steps each, numerical
    integrals, GMRES tol 1e-6, Poisson arrivals; served with continuous
    batching, then the same stream one request at a time."""
    from rigidfibers_amd.config import Parameters
    from rigidfibers_amd.serve import Server, request_stream, serial_stream, serve_stream, summarize

    params = Parameters(NUMBER_OF_FIBERS=1, TIMESTEP=0.1, SLENDERNESS=0.01, NUMBER_OF_TERMS_IN_FORCE_EXPANSION=5,
                        NUMBER_OF_QUADRATURE_INTERVALS=8, GMRES_RESTART=10, GMRES_MAX_ITERATIONS=1000,
                        GMRES_TOLERANCE=1e-6)
    reqs = request_stream(count=args.serve_requests, m_range=(16, 400), density=0.1, rate=args.serve_rate,
                          steps=20, seed=args.seed)
    with Server(params, integration="numerical", memory_budget=0, max_active=args.serve_slots,
                eager=args.serve_eager) as srv:
        res, wall = serve_stream(srv, reqs)
        st = srv.stats()
    out = {"workload": f"{len(reqs)} suspensions, M ~ U[16,400], cube at density 0.1, 20 steps, numerical, "
                       f"GMRES tol 1e-6, Poisson {args.serve_rate:g}/s",
           "batched": summarize(res, wall),
           "server": {"ticks": st["ticks"], "busy_s": round(st["busy_seconds"], 3), "slots": args.serve_slots,
                      "launches": "eager" if args.serve_eager else "graph",
                      "peak_bytes": st["peak_bytes"], "memory_budget": st["memory_budget"],
                      "failed": sum(1 for r in res if r["status"] not in (0, 100))}}
    if args.serve_serial > 0:
        sub = reqs[:args.serve_serial]
        sres, swall = serial_stream(params, sub, integration="numerical")
        out["one_at_a_time"] = summarize(sres, swall)
        out["one_at_a_time"]["requests_of_stream"] = len(sub)
    return out


def large_suspension(args):
    """M = 10 000 fibers in a cube of side 60 (n = 150 000), numerical
    integrals, GMRES restart 30 to 1e-8, 5 steps: GMRES-IR on the FP16 theta
    copy (FP32 residuals) against plain GMRES on the FP32 theta system."""
    from rigidfibers_amd.config import Parameters
    from rigidfibers_amd.scenes import cube
    from rigidfibers_amd.simulation import Simulation

    M = args.large_fibers
    pos, ori = cube(M, 60.0 * (M / 10000.0) ** (1.0 / 3.0), min_dist=0.4, seed=args.seed)
    params = Parameters(NUMBER_OF_FIBERS=M, TIMESTEP=0.1, SLENDERNESS=0.01, NUMBER_OF_TERMS_IN_FORCE_EXPANSION=5,
                        NUMBER_OF_QUADRATURE_INTERVALS=8, GMRES_RESTART=30, GMRES_MAX_ITERATIONS=2000,
                        GMRES_TOLERANCE=1e-8)
    out, ref = {"workload": f"M={M} cube of side {60.0 * (M / 10000.0) ** (1.0 / 3.0):g}, numerical, GMRES(30) "
                            f"tol 1e-8, 1 warm-up + {args.large_steps} timed steps"}, None
    for solver in ("gmres", "gmres-ir"):
        with Simulation(params, pos, ori, solver=solver, integration="numerical") as sim:
            iters, gemv, gemv_n, half, half_n, solve = [], 0.0, 0, 0.0, 0, 0.0
            sim.step(0)  # untimed warm-up (cold GMRES start)
            t0 = time.perf_counter()
            for t in range(1, args.large_steps + 1):
                sim.step(t)
                tm = sim.timings()
                iters.append(sim.solver_stats()["iterations"])
                gemv += tm["gemv"]
                gemv_n += tm["gemv_launches"]
                half += tm["gemv_half"]
                half_n += tm["gemv_half_launches"]
                solve += tm["solve_system"]
            elapsed = time.perf_counter() - t0
            p_end, o_end = sim.state()
            dev = sim.device_bytes()
        per32 = float(M) * 27 * M * 5 * 4
        entry = {"fiber_steps_per_s": round(M * args.large_steps / elapsed, 1),
                 "ms_per_step": round(1e3 * elapsed / args.large_steps, 2),
                 "solve_ms_per_step": round(solve / args.large_steps, 2), "iterations_per_step": iters,
                 "fp32_matvecs": gemv_n, "fp32_matvec_ms": round(gemv / max(gemv_n, 1), 4),
                 "fp32_matvec_GBps": round(per32 / (gemv / max(gemv_n, 1) * 1e-3) / 1e9, 1) if gemv_n else None,
                 "device_bytes": dev}
        if half_n:
            per16 = per32 / 2 + float(M) * M * 4
            entry.update({"fp16_matvecs": half_n, "fp16_matvec_ms": round(half / half_n, 4),
                          "fp16_matvec_GBps": round(per16 / (half / half_n * 1e-3) / 1e9, 1)})
        if ref is None:
            ref = (p_end, o_end, elapsed)
        else:
            entry["speedup_vs_fp32_gmres"] = round(ref[2] / elapsed, 3)
            entry["pos_rel_err_vs_fp32_gmres"] = rel_err(p_end, ref[0])
            entry["ori_rel_err_vs_fp32_gmres"] = rel_err(o_end, ref[1])
        out[solver.replace("-", "_")] = entry
    return out


def cpu_threads():
    """OMP_NUM_THREADS, else every CPU this process may run on."""
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or CPUS_AVAILABLE


def cpu_quota():
    """CPUs of the cgroup's quota (cpu.max), or None."""
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, p = fh.read().split()[:2]
        return None if q == "max" else float(q) / float(p)
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, pos, ori, iters_per_step):
    """FP32 OpenMP port of the reference Fortran on a bounded sample of rows."""
    import numpy as np

    from oracle.pyoracle import Oracle, make_params

    threads = cpu_threads()
    o = Oracle(32)
    M = args.fibers
    p = make_params(M, threads=threads)
    analytical = args.integration == "analytical"
    pos32, ori32 = pos.astype(np.float32), ori.astype(np.float32)
    # size the sample from a small probe so that it takes ~budget seconds
    t = time.perf_counter()
    o.assemble_rows(p, pos32, ori32, 0, threads, analytical)
    probe = (time.perf_counter() - t) / threads
    S = int(max(8, min(M, args.cpu_budget / 2.0 / max(probe, 1e-6))))
    t = time.perf_counter()
    A_rows, _ = o.assemble_rows(p, pos32, ori32, 0, S, analytical)
    t_asm = time.perf_counter() - t
    x = np.random.default_rng(0).normal(size=3 * 5 * M).astype(np.float32)
    t = time.perf_counter()
    reps = 3
    for _ in range(reps):
        o.matvec_rows(A_rows, x)
    t_mv = (time.perf_counter() - t) / reps
    del A_rows
    t = time.perf_counter()
    o.velocities_rows(p, pos32, ori32, x, 0, S, analytical)
    t_vel = time.perf_counter() - t
    step_s = (t_asm + t_vel + iters_per_step * t_mv) * M / S
    return {
        "value": M / step_s,
        "unit": "fiber-steps/s",
        "cores": min([threads, CPUS_AVAILABLE] + ([cpu_quota()] if cpu_quota() else [])),
        "threads": threads,
        "cpu_model": cpu_model(),
        "cpus_available": CPUS_AVAILABLE,
        "cpu_quota": cpu_quota(),
        "kind": "port",
        "sample": (f"FP32 OpenMP port of the reference Fortran (oracle/librf_oracle32.so), {threads} threads: "
                   f"assembly of the rows of {S} of {M} target fibers ({t_asm:.2f} s), their velocities "
                   f"({t_vel:.2f} s) and {iters_per_step:.1f} GMRES matvecs of those rows ({t_mv:.3f} s each), "
                   f"scaled by M/{S} to one step ({step_s:.1f} s)"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fibers", type=int, default=5000)
    ap.add_argument("--radius", type=float, default=35.0)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--solver", default="gmres", choices=["gmres", "bicgstab", "direct", "mixed-bf16", "mixed-fp16"])
    ap.add_argument("--integration", default="analytical", choices=["analytical", "numerical"])
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--restart", type=int, default=20)
    ap.add_argument("--matrix-format", default="auto", choices=["auto", "dense", "theta"])
    ap.add_argument("--velocity-operator", default="stored", choices=["stored", "pairwise"])
    ap.add_argument("--initial-guess", default="previous", choices=["previous", "zero", "extrapolated"],
                    help="iterative solvers' start: the previous solution or 2 x_n - x_{n-1}")
    ap.add_argument("--lu-steps", type=int, default=2, help="timed steps per solver of the precision study")
    ap.add_argument("--separate-residual", action="store_true",
                    help="form the first residual of a warm start with a matvec, not in the assembly kernel")
    ap.add_argument("--confirm-residual", action="store_true",
                    help="GMRES and BiCGStab: confirm a converged solve with the true residual (one matvec more)")
    ap.add_argument("--no-precision-study", action="store_true")
    ap.add_argument("--no-oracle-check", dest="oracle_check", action="store_false",
                    help="precision study without its comparisons against the FP64 oracle")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=50.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-slots", type=int, default=64, help="suspensions stepped together")
    ap.add_argument("--serve-serial", type=int, default=512, help="requests of the one-at-a-time comparison")
    ap.add_argument("--serve-eager", action="store_true",
                    help="launch each GMRES iteration's kernels one by one instead of replaying a graph")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--large-fibers", type=int, default=10000)
    ap.add_argument("--large-steps", type=int, default=5)
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline sampling")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the timestep on one GPU; multi-GPU runs are not part of it "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    # the oracle's OpenMP threads stay on neighbouring cores (read when the
    # oracle library first starts its runtime)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    pos, ori = build_scene(args.fibers, args.radius, args.seed)
    g = run_gpu(args, pos, ori, args.solver, args.steps, args.warmup)
    g["asm_flops"] = assembly_flops(args.integration, pos, ori)
    M, K = args.fibers, args.steps
    value = M * K / g["elapsed"]
    fmt = ("theta" if args.matrix_format in ("auto", "theta") else "dense") if args.solver in ITERATIVE else "dense"
    line = {
        "metric": "fiber-steps/s at M=5000 (plain FP32 vs mixed-prec, pos err<=1e-6); p99 serve latency",
        "value": round(value, 2),
        "unit": "fiber-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * g["elapsed"] / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 matrix, f64 recurrences/Krylov",
        "data": "synthetic (seeded ball scene)",
        "config": {"workload": f"rigid-fiber timestep, M={M} fibers in a ball of radius {args.radius:g}, "
                               f"min centre distance 0.4, N=5, 8x3 quadrature, {args.integration} integrals, "
                               f"{args.solver} tol {args.tol:g} (restart {args.restart}), dt=0.1",
                   "fibers": M, "unknowns": g["n"], "solver": args.solver, "integration": args.integration,
                   "matrix_format": fmt},
        "roofline": None,
        "roofline_solver": roofline(args, args.solver, g),
        "kernels": {k: round(v / K, 3) for k, v in g["phases"].items()},
        "solver": {"iterations_per_step": g["iters"], "last": g["stats"]},
        "device_bytes": g["device_bytes"],
    }
    # the roofline of the step's dominant kernel: the assembly for the
    # iterative solvers, the LU's trailing updates for the direct ones
    asm = roofline_assembly(args, g)
    line["roofline_assembly"] = asm
    line["roofline"] = asm if args.solver in ITERATIVE else line["roofline_solver"]
    bound = step_bound(args, g)
    if bound is not None:
        line["step_bound"] = bound
    line["kernels"]["gemv_ms_per_step"] = round(g["acc"]["gemv"] / K, 3)
    if not args.no_precision_study:
        line["precision"] = precision_study(args, pos, ori, plain=g if args.solver == "gmres" else None)
    if not args.no_serving:
        line["serving"] = serving(args)
    if not args.no_large:
        line["large"] = large_suspension(args)
    if not args.no_cpu_baseline:
        # the matvecs the GPU's solves did per step (Arnoldi steps and true residuals)
        iters = g["acc"]["gemv_launches"] / K if args.solver in ITERATIVE else 1.0
        line["cpu_baseline"] = cpu_baseline(args, pos, ori, iters)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
