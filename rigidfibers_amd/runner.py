"""Command line of the reference (fibers_runner.py:17-171), run on the HIP path.

    fibers_runner.py [--verbose] {cuda,fortran} {run,validate,benchmark} ...

cuda:     (--magma | --gmres | --bicgstab) (--numerical | --analytical)
          (--D1 | --D2 | --D3) configuration_file [fibers_file]
fortran:  (--direct | --gmres) (--numerical | --analytical) [--threads T]
          configuration_file [fibers_file]

Both platforms run the same library (include/rigidfibers.h); the platform
picks the reference's file names and semantics for outputs.  The three grid
decompositions of the CUDA code (--D1/--D2/--D3, src/simulation.cu:338-396)
are accepted and run the one assembly kernel.  Instead of generating a
constants header and rebuilding per run (runner/cuda_runner.py:49-152), the
configuration is read at run time.

Additions: --mixed-bf16/--mixed-fp16 (cuda solvers), --velocity-integration,
--initial-guess, --output-dir, --reference-dir and tolerances for validate,
--sizes/--steps for benchmark.
"""
from __future__ import annotations

import argparse
import csv
import math
import os
import sys
import time

import numpy as np

from .config import ConfigError, Parameters, read_configuration, read_fibers

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _solver(args):
    for flag, name in (("magma", "direct"), ("direct", "direct"), ("gmres", "gmres"), ("bicgstab", "bicgstab"),
                       ("mixed_bf16", "mixed-bf16"), ("mixed_fp16", "mixed-fp16")):
        if getattr(args, flag, False):
            return name
    raise SystemExit("no solver selected")


def _integration(args):
    return "analytical" if args.analytical else "numerical"


def _add_common(p, platform, command):
    p.add_argument("configuration_file", help="the configuration file")
    if command != "benchmark":
        p.add_argument("fibers_file", help="the file containing the initial positions and orientations")
    solvers = p.add_argument_group("supported solvers").add_mutually_exclusive_group(required=True)
    if platform == "cuda":
        solvers.add_argument("--magma", action="store_true", help="direct LU solve (FP32)")
        solvers.add_argument("--gmres", action="store_true")
        solvers.add_argument("--bicgstab", action="store_true")
        solvers.add_argument("--mixed-bf16", dest="mixed_bf16", action="store_true",
                             help="LU with BF16 MFMA trailing updates + FP64 iterative refinement")
        solvers.add_argument("--mixed-fp16", dest="mixed_fp16", action="store_true",
                             help="LU with FP16 MFMA trailing updates + FP64 iterative refinement")
    else:
        solvers.add_argument("--direct", action="store_true")
        solvers.add_argument("--gmres", action="store_true")
        p.add_argument("--threads", type=int, help="accepted for compatibility (the timestep runs on the GPU)")
    algos = p.add_argument_group("supported algorithms").add_mutually_exclusive_group(required=True)
    algos.add_argument("--numerical", action="store_true")
    algos.add_argument("--analytical", action="store_true")
    if platform == "cuda":
        dims = p.add_argument_group("supported parallization dimensions").add_mutually_exclusive_group(required=True)
        dims.add_argument("--D1", action="store_true")
        dims.add_argument("--D2", action="store_true")
        dims.add_argument("--D3", action="store_true")
    p.add_argument("--velocity-integration", choices=["follow", "numerical", "analytical"], default="follow")
    p.add_argument("--initial-guess", choices=["previous", "zero", "extrapolated"], default="previous")
    p.add_argument("--confirm-residual", action="store_true",
                   help="GMRES and BiCGStab: confirm a converged solve with the true residual, as the Fortran "
                        "solver does")
    p.add_argument("--output-dir", default=".")
    if command == "validate":
        p.add_argument("--reference-dir", default=None,
                       help="directory of <step>_<FIELD>.ref vectors (default: the golden vectors matching the flags)")
        p.add_argument("--atol", type=float, default=1e-5)
        p.add_argument("--rtol", type=float, default=1e-5)
    if command == "benchmark":
        p.add_argument("--max_rse", type=float, default=0.2, help="the maximum relative standard error (default: 0.2)")
        p.add_argument("--sizes", default="2000:100:-100", help="fiber counts, start:stop:step or a comma list")
        p.add_argument("--density", type=float, default=0.031, help="fibers per unit volume of the generated cube")
        p.add_argument("--seed", type=int, default=1)


def build_parser():
    parser = argparse.ArgumentParser(
        description="Helper program for running, validating and benchmarking the rigid fibers simulation")
    parser.add_argument("--verbose", action="store_true", help="be verbose")
    platforms = parser.add_subparsers(title="supported platforms", dest="platform")
    platforms.required = True
    for platform, help_ in (("fortran", "the reference's OpenMP Fortran surface (runs on the GPU here)"),
                            ("cuda", "the reference's CUDA surface (runs on the GPU here)")):
        pp = platforms.add_parser(platform, help=help_)
        cmds = pp.add_subparsers(title="supported commands", dest="command")
        cmds.required = True
        for command, func in (("run", run), ("validate", validate), ("benchmark", benchmark)):
            c = cmds.add_parser(command, help=f"{command} the simulation")
            _add_common(c, platform, command)
            c.set_defaults(func=func, run=command == "run", validate=command == "validate",
                           benchmark=command == "benchmark")
    return parser


def _make_sim(args, params, pos, ori):
    from .simulation import Simulation

    return Simulation(params, pos, ori, solver=_solver(args), integration=_integration(args),
                      velocity_integration=args.velocity_integration, initial_guess=args.initial_guess,
                      confirm_residual=args.confirm_residual)


def _report_solver(sim, verbose):
    st = sim.solver_stats()
    if verbose or not st["converged"]:
        print(f"     [GPU]      : No. of iters : {st['iterations']}")
        print(f"     [GPU]      : Est. error   : {st['relative_residual']:.3e}"
              + ("" if st["converged"] else "  (NOT CONVERGED)"))


def run(args):
    from . import output

    params = read_configuration(args.configuration_file)
    M, pos, ori = read_fibers(args.fibers_file)
    params.NUMBER_OF_FIBERS = M
    os.makedirs(args.output_dir, exist_ok=True)
    with _make_sim(args, params, pos, ori) as sim:
        for t in range(params.NUMBER_OF_TIMESTEPS):
            if args.platform == "fortran":
                # the Fortran saves the state before the update (fortran/fibers.f95:305-322)
                p, o = sim.state()
            code = sim.step(t)
            tm = sim.timings()
            print(f"step {t}: {tm['total']:.2f} ms (assemble {tm['assemble_system']:.2f}, solve "
                  f"{tm['solve_system']:.2f}, velocities {tm['update_velocities']:.2f}, update "
                  f"{tm['update_fibers']:.2f})")
            _report_solver(sim, args.verbose or code != 0)
            if args.platform == "fortran":
                if params.STATE_SAVE_INTERVAL > 0 and t % params.STATE_SAVE_INTERVAL == 0:
                    output.write_fortran_state(os.path.join(args.output_dir, f"{t:05d}.state"), p, o)
                if params.VELOCITY_SAVE_INTERVAL > 0 and t % params.VELOCITY_SAVE_INTERVAL == 0:
                    tv, _ = sim.velocities()
                    output.write_fortran_velocity(os.path.join(args.output_dir, f"{t:05d}.velocity"), tv)
            else:
                if params.STATE_SAVE_INTERVAL > 0 and t % params.STATE_SAVE_INTERVAL == 0:
                    p, o = sim.state()
                    output.write_state(os.path.join(args.output_dir, f"XcT_res_{t + 1:05d}.out"), p, o)
                if params.VELOCITY_SAVE_INTERVAL > 0 and t % params.VELOCITY_SAVE_INTERVAL == 0:
                    tv, rv = sim.velocities()
                    output.write_state(os.path.join(args.output_dir, f"TcR_res_{t + 1:05d}.out"), tv, rv)
    print("**************************************************")
    return 0


def _default_reference_dir(args):
    variant = ("analytic" if args.analytical else "numeric") + "_" + (
        "direct" if _solver(args) in ("direct", "mixed-bf16", "mixed-fp16") else "gmres")
    return os.path.join(_ROOT, "tests", "golden", "reference", f"100_{variant}")


def validate(args):
    """Two timesteps with every VALIDATE dump, compared against the reference
    vectors (runner/cuda_runner.py:165-215).  Exits non-zero on a mismatch."""
    from . import output
    from .validate import compare

    params = read_configuration(args.configuration_file, validate=True)
    M, pos, ori = read_fibers(args.fibers_file)
    params.NUMBER_OF_FIBERS = M
    ref_dir = args.reference_dir or _default_reference_dir(args)
    os.makedirs(args.output_dir, exist_ok=True)
    failures = 0
    with _make_sim(args, params, pos, ori) as sim:
        for t in range(params.NUMBER_OF_TIMESTEPS):
            sim.assemble_system()
            A, b, _ = sim.system(with_matrix=args.verbose)
            sim.solve_system()
            _, _, x = sim.system()
            sim.update_velocities()
            tv, rv = sim.velocities()
            sim.update_fibers(t == 0)
            p, o = sim.next_state()
            sim.advance()
            names = output.dump_step(args.output_dir, args.platform, t, b, x, tv, rv, p, o, A)
            if t == 0 and args.platform == "cuda":  # written once: it does not change
                output.write_mapping(os.path.join(args.output_dir, "current.map"), M,
                                     params.NUMBER_OF_TERMS_IN_FORCE_EXPANSION)
            for key, field in (("b", "BVec"), ("x", "XVec"), ("t", "TRANSVel"), ("r", "ROTVel"), ("p", "POS"),
                               ("o", "ORIENT")):
                ref = os.path.join(ref_dir, f"{t}_{field}.ref")
                if not os.path.exists(ref):
                    print(f"  (no reference vector {ref})")
                    continue
                rep = compare(os.path.join(args.output_dir, names[key]), ref, args.atol, args.rtol)
                print(rep.text())
                failures += 0 if rep.passed else 1
    print("**************************************************")
    print("Validation " + ("passed" if failures == 0 else f"FAILED ({failures} comparisons)"))
    return 0 if failures == 0 else 1


def _sizes(spec):
    if ":" in spec:
        a, b, c = (int(v) for v in spec.split(":"))
        return list(range(a, b + (1 if c > 0 else -1), c))
    return [int(v) for v in spec.split(",") if v]


def benchmark(args):
    """Per-phase step times over fiber counts, repeated until the relative
    standard error of the total is below --max_rse (runner/cuda_runner.py:217-339),
    first step excluded as warm-up; writes results.csv (tab-separated)."""
    from . import output
    from .scenes import cube

    params = read_configuration(args.configuration_file)
    names = ("assemble_system", "solve_system", "update_velocities", "update_fibers")
    results = {}
    for idx, M in enumerate(_sizes(args.sizes)):
        print(f"  [BENCHMARK]   : {M} fibers ({idx + 1})")
        side = (M / args.density) ** (1.0 / 3.0)
        runs = []
        rse = math.inf
        rounds = 0
        while rse > args.max_rse and rounds < 8:
            rounds += 1
            for _ in range(4):
                pos, ori = cube(M, side, min_dist=0.4, seed=args.seed + len(runs))
                p = Parameters(**{**params.__dict__, "NUMBER_OF_FIBERS": M})
                with _make_sim(args, p, pos, ori) as sim:
                    acc = {k: [] for k in names}
                    for t in range(max(2, p.NUMBER_OF_TIMESTEPS)):
                        sim.step(t)
                        tm = sim.timings()
                        if t > 0:  # the first step is the warm-up (src/performance.cc:68-78)
                            for k in names:
                                acc[k].append(tm[k] / 1e3)
                runs.append({k: float(np.mean(v)) for k, v in acc.items()})
            totals = [sum(r.values()) for r in runs]
            mean = float(np.mean(totals))
            std = float(np.std(totals, ddof=1)) if len(totals) > 1 else 0.0
            rse = (std / math.sqrt(len(totals))) / mean if mean > 0 else 0.0
        row = {}
        for k in names:
            vals = [r[k] for r in runs]
            row[k] = float(np.mean(vals))
            row[k + "_std"] = float(np.std(vals, ddof=1)) if len(vals) > 1 else 0.0
        row["total"] = mean
        row["total_std"] = std
        results[M] = row
        output.write_performance(os.path.join(args.output_dir, "performance.out"), [(k, row[k]) for k in names])
    keys = sorted(next(iter(results.values())).keys())
    with open(os.path.join(args.output_dir, "results.csv"), "w", newline="") as fh:
        w = csv.writer(fh, dialect="excel-tab")
        w.writerow(["X"] + keys)
        print("**************************************************")
        print("Benchmark:")
        print("  " + "\t".join(["X"] + keys))
        for M in sorted(results):
            w.writerow([M] + [results[M][k] for k in keys])
            print("  " + "\t".join(str(v) for v in [M] + [results[M][k] for k in keys]))
    print("**************************************************")
    return 0


def main(argv=None):
    parser = build_parser()
    args = parser.parse_args(argv)
    try:
        return args.func(args)
    except (ConfigError, FileNotFoundError) as exc:
        print(f"error: {exc}", file=sys.stderr)
        return 1
