#!/usr/bin/env python3
"""Throughput of the rigid-fiber timestep on one MI355X, in fiber-steps/s.

Workload (the configuration the headline metric is quoted on): M = 5000 fibers
of length 2 with centres uniform in a ball of radius 35 (minimum centre
distance 0.4, seed 2026), orientations uniform on the sphere, slenderness 0.01,
N = 5 force terms, 8 x 3 quadrature, unit gravity, dt = 0.1, analytical inner
integrals (FP64 recurrences), GMRES(10) to a relative residual of 1e-6 starting
from the previous solution.  n = 75 000 unknowns, A = 22.5 GB FP32 resident in HBM.

One "step" is one call of the library's rf_step (assemble -> solve -> update
velocities -> update fibers) on device-resident state.  W warm-up steps are
run untimed, then exactly K steps are timed between two device
synchronisations.  value = M * K / elapsed.

Also reported:
  roofline      the matvec of the solver (HBM-bound): its algorithmic bytes
                per launch (theta form: M * M * N * (3 + 6(N-1)) * 4, the FP32
                numbers that define A; dense: n^2 * 4) / mean launch duration
                (HIP events on the library's stream), against 8 TB/s;
  kernels       per-phase times and the assembly kernel's rate;
  cpu_baseline  the FP32 OpenMP port of the reference Fortran (oracle build
                librf_oracle32.so) on a bounded sample: the rows of S target
                fibers assembled against all 5000 sources, their velocities,
                and the GMRES matvecs of those rows; scaled by M/S to a step.
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
FP32_VALU_PEAK_TFLOPS = 157.3


def build_scene(M, radius, seed):
    from rigidfibers_amd.scenes import ball

    return ball(M, radius, min_dist=0.4, seed=seed)


def run_gpu(args, pos, ori):
    from rigidfibers_amd.config import Parameters
    from rigidfibers_amd.simulation import Simulation

    M = args.fibers
    params = Parameters(NUMBER_OF_FIBERS=M, TIMESTEP=0.1, SLENDERNESS=0.01, NUMBER_OF_TERMS_IN_FORCE_EXPANSION=5,
                        NUMBER_OF_QUADRATURE_INTERVALS=8, GMRES_RESTART=args.restart,
                        GMRES_MAX_ITERATIONS=1000, GMRES_TOLERANCE=args.tol, BICGSTAB_TOLERANCE=args.tol)
    sim = Simulation(params, pos, ori, solver=args.solver, integration=args.integration,
                     matrix_format=args.matrix_format, velocity_operator=args.velocity_operator)
    try:
        for t in range(args.warmup):
            sim.step(t)
            print(f"# warmup step {t}: {sim.timings()['total']:.1f} ms, solver {sim.solver_stats()}", flush=True)
        phases = {"assemble_system": 0.0, "solve_system": 0.0, "update_velocities": 0.0, "update_fibers": 0.0}
        gemv_ms, gemv_n, iters = 0.0, 0, []
        t0 = time.perf_counter()
        for t in range(args.warmup, args.warmup + args.steps):
            sim.step(t)
            tm = sim.timings()
            for k in phases:
                phases[k] += tm[k]
            gemv_ms += tm["gemv"]
            gemv_n += tm["gemv_launches"]
            iters.append(sim.solver_stats()["iterations"])
        elapsed = time.perf_counter() - t0
        stats = sim.solver_stats()
        pos_end, ori_end = sim.state()
        device_bytes = sim.device_bytes()
        n = sim.n
    finally:
        sim.close()
    return {"elapsed": elapsed, "phases": phases, "gemv_ms": gemv_ms, "gemv_n": gemv_n, "iters": iters,
            "stats": stats, "n": n, "device_bytes": device_bytes, "pos": pos_end, "ori": ori_end}


def cpu_baseline(args, pos, ori, iters_per_step):
    """FP32 OpenMP port of the reference Fortran on a bounded sample of rows."""
    import numpy as np

    from oracle.pyoracle import Oracle, make_params

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
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
        "cores": threads,
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
    ap.add_argument("--restart", type=int, default=10)
    ap.add_argument("--matrix-format", default="auto", choices=["auto", "dense", "theta"])
    ap.add_argument("--velocity-operator", default="stored", choices=["stored", "pairwise"])
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline sampling")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the timestep on one GPU; multi-GPU runs are not part of it "
              "(use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    pos, ori = build_scene(args.fibers, args.radius, args.seed)
    g = run_gpu(args, pos, ori)
    M, K, n = args.fibers, args.steps, g["n"]
    value = M * K / g["elapsed"]
    gemv_avg_ms = g["gemv_ms"] / max(g["gemv_n"], 1)
    iterative = args.solver in ("gmres", "bicgstab")
    theta = args.matrix_format == "theta" or (args.matrix_format == "auto" and iterative)
    N = 5
    if theta:  # 3 + 6(N-1) FP32 numbers per (target, source, source mode)
        gemv_bytes = float(M) * (3 + 6 * (N - 1)) * M * N * 4.0
        gemv_kernel = "gemv_theta_kernel"
    else:      # n^2 FP32 entries
        gemv_bytes = float(n) * n * 4.0
        gemv_kernel = "gemv_f32_kernel"
    achieved = gemv_bytes / (gemv_avg_ms * 1e-3) / 1e9
    asm_ms = g["phases"]["assemble_system"] / K
    # algorithmic count of the analytical assembly: per ordered pair and target
    # node, FP64 moments (~2 log + 2 sqrt + 26 x 11 recurrence ops) and the
    # G of 6 modes (~6 x 110) plus the projection (~(3N+1) x 3N x Q / Q per node)
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
                   "fibers": M, "unknowns": n, "solver": args.solver, "integration": args.integration,
                   "matrix_format": "theta" if theta else "dense"},
        "roofline": {"bound": "hbm", "kernel": gemv_kernel, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_launch": gemv_bytes, "avg_launch_ms": round(gemv_avg_ms, 4)},
        "kernels": {k: round(v / K, 3) for k, v in g["phases"].items()},
        "solver": {"iterations_per_step": g["iters"], "last": g["stats"]},
        "device_bytes": g["device_bytes"],
    }
    line["kernels"]["gemv_ms_per_step"] = round(g["gemv_ms"] / K, 3)
    line["kernels"]["assemble_ms_per_step"] = round(asm_ms, 3)
    if not args.no_cpu_baseline:
        iters = sum(g["iters"]) / len(g["iters"]) + 1.0
        line["cpu_baseline"] = cpu_baseline(args, pos, ori, iters)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
