"""Output files of the reference, written from the host.

  XcT_res_%05d.out / TcR_res_%05d.out   saveFibers / saveVelocities
                                        (src/simulation.cu:530-604): M, then
                                        "x y z" and "tx ty tz" lines, 8 decimals
  %05d.state / %05d.velocity            the Fortran saves (fortran/fibers.f95:305-322)
  VALIDATE dumps                        a_matrix_t / b_vector_t / x_vector_t /
                                        t_vel_t / r_vel_t / positions_t /
                                        orientations_t .out (src/simulation.cu:606-822)
                                        and the Fortran t_AMat / t_BVec / ... names
                                        (fortran/fibers.f95:205-303)
  current.map                           per entry of A the simulation variables it
                                        couples (src/simulation.cu:711-740,
                                        src/kernels/assemble_system_2D.cu:71-92,162-225)
  performance.out                       name,time rows (src/performance.cc:110-134)
"""
from __future__ import annotations

import os

import numpy as np


def _fmt_signed(v):
    """%.8f padded as the reference's dumps: 5 spaces before negative values, 6 before others."""
    s = f"{v:.8f}"
    return ("     " if v < 0 else "      ") + s


def write_column(path, values):
    with open(path, "w") as fh:
        for v in np.asarray(values, dtype=np.float64).ravel():
            fh.write(_fmt_signed(v) + "\n")


def write_matrix(path, A):
    """Row per line, values as in dumpLinearSystem (src/simulation.cu:683-708)."""
    A = np.asarray(A)
    with open(path, "w") as fh:
        for r in range(A.shape[0]):
            fh.write("".join(_fmt_signed(float(v)) for v in A[r]) + "\n")


def write_state(path, positions, orientations):
    p = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    o = np.asarray(orientations, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        fh.write(f"{p.shape[0]}\n")
        for a, b in zip(p, o):
            fh.write(f"{a[0]:.8f} {a[1]:.8f} {a[2]:.8f}\n{b[0]:.8f} {b[1]:.8f} {b[2]:.8f}\n")


def write_fortran_state(path, positions, orientations):
    """%05d.state: M then 2 lines per fiber in F16.8 columns (fortran/fibers.f95:305-315)."""
    p = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    o = np.asarray(orientations, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        fh.write(f"{p.shape[0]:12d}\n")
        for a, b in zip(p, o):
            fh.write("".join(f"{v:16.8f}" for v in a) + "\n" + "".join(f"{v:16.8f}" for v in b) + "\n")


def write_fortran_velocity(path, translational):
    t = np.asarray(translational, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        for a in t:
            fh.write("".join(f"{v:16.8f}" for v in a) + "\n")


def write_performance(path, rows):
    """rows: iterable of (name, seconds)."""
    with open(path, "w") as fh:
        fh.write("name,time\n")
        for name, t in rows:
            fh.write(f"{name},{t:.9f}\n")


def mapping_fields(rows, cols, N):
    """(i, j, k_j, k_i, a, b) of the entries (rows, cols) of A: row 3N i + 3 k_i + a,
    column 3N j + 3 k_j + b.  Entries of a fiber's own block other than its unit
    diagonal are never assembled and carry -1 in every field."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    w = 3 * N
    i, ri = np.divmod(rows, w)
    j, rj = np.divmod(cols, w)
    ki, a = np.divmod(ri, 3)
    kj, b = np.divmod(rj, 3)
    f = np.stack([i, j, kj, ki, a, b])
    unset = (i == j) & ~((ki == kj) & (a == b))
    f[:, unset] = -1
    return f


def write_mapping(path, M, N, rows_per_chunk=256):
    """current.map: one line "[V]|i|j|k_j|k_i|a|b|row|col" per entry of the
    n x n matrix, rows outer, columns inner."""
    n = 3 * N * M
    cols = np.arange(n, dtype=np.int64)
    with open(path, "w") as fh:
        for r0 in range(0, n, rows_per_chunk):
            r1 = min(n, r0 + rows_per_chunk)
            rows = np.repeat(np.arange(r0, r1, dtype=np.int64), n)
            cc = np.tile(cols, r1 - r0)
            f = mapping_fields(rows, cc, N)
            fh.write("".join(f"[V]|{a}|{b}|{c}|{d}|{e}|{g}|{r}|{q}\n"
                             for a, b, c, d, e, g, r, q in zip(*f.tolist(), rows.tolist(), cc.tolist())))


def dump_step(out_dir, platform, step, b, x, tvel, rvel, positions, orientations, A=None):
    """The VALIDATE dumps of one step under the reference's names for the platform."""
    os.makedirs(out_dir, exist_ok=True)
    if platform == "fortran":
        names = {"A": f"{step}_AMat.out", "b": f"{step}_BVec.out", "x": f"{step}_XVec.out",
                 "t": f"{step}_TRANSVel.out", "r": f"{step}_ROTVel.out", "p": f"{step}_POS.out",
                 "o": f"{step}_ORIENT.out"}
    else:
        names = {"A": f"a_matrix_{step}.out", "b": f"b_vector_{step}.out", "x": f"x_vector_{step}.out",
                 "t": f"t_vel_{step}.out", "r": f"r_vel_{step}.out", "p": f"positions_{step}.out",
                 "o": f"orientations_{step}.out"}
    if A is not None:
        write_matrix(os.path.join(out_dir, names["A"]), A)
    write_column(os.path.join(out_dir, names["b"]), b)
    write_column(os.path.join(out_dir, names["x"]), x)
    write_column(os.path.join(out_dir, names["t"]), tvel)
    write_column(os.path.join(out_dir, names["r"]), rvel)
    write_column(os.path.join(out_dir, names["p"]), positions)
    write_column(os.path.join(out_dir, names["o"]), orientations)
    return names
