"""Seeded fiber configurations.

The reference's generators (tools/gen2.f95:29-102, tools/gen3.py:162-251) are
seeded from /dev/urandom and enforce a minimum segment distance; the scenes
here are reproducible:

  cube, ball     centres uniform in a cube or a ball, rejected while closer
                 than a minimum centre distance to an accepted one (cell-list
                 check), orientations uniform on the unit sphere;
  grown_sphere   the gen2 construction (tools/gen2.f95:51-89): fibers drawn in
                 the shell max(0, R - 4d) < |x| < R with orientations from a
                 normalised uniform sample of [-1, 1]^3, rejected while the
                 segment x +- t comes within d of an accepted fiber's segment;
                 after 1000 rejections in a row R grows by d.

`python -m rigidfibers_amd.scenes {cube,ball,gen2} M ...` writes a fibers file.
"""
from __future__ import annotations

import math

import numpy as np


def _sample(M, draw, min_dist, rng, max_tries):
    cell = max(min_dist, 1e-9)
    grid = {}
    pts = []
    tries = 0
    while len(pts) < M:
        if tries >= max_tries:
            raise RuntimeError(f"could only place {len(pts)} of {M} fibers with min distance {min_dist}")
        tries += 1
        c = draw(rng)
        key = tuple(int(math.floor(v / cell)) for v in c)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        if (c[0] - q[0]) ** 2 + (c[1] - q[1]) ** 2 + (c[2] - q[2]) ** 2 < min_dist * min_dist:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid.setdefault(key, []).append(c)
            pts.append(c)
    return np.array(pts, dtype=np.float64)


def _orientations(M, rng):
    o = rng.normal(size=(M, 3))
    return o / np.linalg.norm(o, axis=1, keepdims=True)


def cube(M, side, min_dist=0.4, seed=0):
    """(positions[3M], orientations[3M]) with centres uniform in [-side/2, side/2]^3."""
    rng = np.random.default_rng(seed)
    h = side / 2.0
    pts = _sample(M, lambda r: r.uniform(-h, h, size=3), min_dist, rng, 200 * M + 1000)
    return pts.ravel(), _orientations(M, rng).ravel()


def ball(M, radius, min_dist=0.4, seed=0):
    """(positions[3M], orientations[3M]) with centres uniform in a ball."""
    rng = np.random.default_rng(seed)

    def draw(r):
        while True:
            c = r.uniform(-radius, radius, size=3)
            if c @ c <= radius * radius:
                return c

    pts = _sample(M, draw, min_dist, rng, 200 * M + 1000)
    return pts.ravel(), _orientations(M, rng).ravel()


def segment_distance2(p, o, P, O):
    """Squared distances between the segment p +- o and each segment P[k] +- O[k]
    (p, o: 3-vectors; P, O: (K, 3)), minimising |X(s) - Y(t)|^2 over s, t in
    [0, 1] for X(s) = p - o + 2 s o, Y(t) = P - O + 2 t O."""
    p = np.asarray(p, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    P = np.atleast_2d(np.asarray(P, dtype=np.float64))
    O = np.atleast_2d(np.asarray(O, dtype=np.float64))
    u = 2.0 * o
    v = 2.0 * O
    w = (p - o) - (P - O)
    a = float(u @ u)
    b = v @ u
    c = np.einsum("ij,ij->i", v, v)
    d = w @ u
    e = np.einsum("ij,ij->i", v, w)
    D = a * c - b * b
    # s on the line through the first segment, clamped; then t for that s,
    # clamped; then s again for the clamped t
    par = D <= 1e-12 * a * c
    s = np.where(par, 0.0, (b * e - c * d) / np.where(par, 1.0, D))
    s = np.clip(s, 0.0, 1.0)
    t = np.clip((b * s + e) / c, 0.0, 1.0)
    s = np.clip((b * t - d) / a, 0.0, 1.0)
    r = w + s[:, None] * u[None, :] - t[:, None] * v
    return np.einsum("ij,ij->i", r, r)


def grown_sphere(M, min_dist, seed=0, retries=1000):
    """(positions[3M], orientations[3M]) built as tools/gen2.f95:51-89 builds
    them (see the module docstring), from a seeded generator.  The `retries`
    candidates of one radius are drawn at once and the first one (in draw
    order) that keeps the distance is taken; segments of half-length 1 closer
    than d have centres closer than 2 + d, so only those neighbours (KD-tree
    of the accepted centres) are measured."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    pos = np.empty((M, 3))
    ori = np.empty((M, 3))
    R = float(min_dist)
    d2 = float(min_dist) * float(min_dist)
    reach = 2.0 + float(min_dist)
    i = 0
    while i < M:
        lo2 = max(0.0, R - 4.0 * min_dist) ** 2
        X = np.empty((0, 3))
        while len(X) < retries:  # centres uniform in the shell, by rejection from the cube
            c = rng.uniform(-R, R, size=(2 * retries, 3))
            r2 = np.einsum("ij,ij->i", c, c)
            X = np.concatenate([X, c[(r2 > lo2) & (r2 < R * R)]])
        X = X[:retries]
        T = rng.uniform(-1.0, 1.0, size=(retries, 3))
        T /= np.linalg.norm(T, axis=1, keepdims=True)
        tree = cKDTree(pos[:i]) if i > 0 else None
        first = -1
        for c0 in range(0, retries, 64):  # in draw order, 64 at a time
            c1 = min(retries, c0 + 64)
            ok = np.ones(c1 - c0, dtype=bool)
            if tree is not None:
                nbr = tree.query_ball_point(X[c0:c1], r=reach)
                cand = np.repeat(np.arange(c1 - c0), [len(v) for v in nbr])
                if len(cand):
                    other = np.concatenate([np.asarray(v, dtype=np.int64) for v in nbr if len(v)])
                    dist2 = _segment_distance2_pairs(X[c0 + cand], T[c0 + cand], pos[other], ori[other])
                    bad = np.zeros(c1 - c0, dtype=bool)
                    np.logical_or.at(bad, cand, dist2 < d2)
                    ok = ~bad
            hit = np.flatnonzero(ok)
            if len(hit):
                first = c0 + int(hit[0])
                break
        if first < 0:
            R += min_dist
            continue
        pos[i], ori[i] = X[first], T[first]
        i += 1
    return pos.ravel(), ori.ravel()


def _segment_distance2_pairs(p, o, P, O):
    """Row-wise segment_distance2: segment p[k] +- o[k] against P[k] +- O[k]."""
    u = 2.0 * o
    v = 2.0 * O
    w = (p - o) - (P - O)
    a = np.einsum("ij,ij->i", u, u)
    b = np.einsum("ij,ij->i", v, u)
    c = np.einsum("ij,ij->i", v, v)
    d = np.einsum("ij,ij->i", u, w)
    e = np.einsum("ij,ij->i", v, w)
    D = a * c - b * b
    par = D <= 1e-12 * a * c
    s = np.where(par, 0.0, (b * e - c * d) / np.where(par, 1.0, D))
    s = np.clip(s, 0.0, 1.0)
    t = np.clip((b * s + e) / c, 0.0, 1.0)
    s = np.clip((b * t - d) / a, 0.0, 1.0)
    r = w + s[:, None] * u - t[:, None] * v
    return np.einsum("ij,ij->i", r, r)


def write_fibers(path, positions, orientations):
    """Fibers file: M, then per fiber a line of the centre and a line of the
    orientation (the layout src/parameters.cc:75-116 reads; 16 decimals)."""
    p = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    o = np.asarray(orientations, dtype=np.float64).reshape(-1, 3)
    with open(path, "w") as fh:
        fh.write(f"{p.shape[0]:10d}\n")
        for a, b in zip(p, o):
            fh.write("".join(f"{v:32.16f}" for v in a) + "\n")
            fh.write("".join(f"{v:32.16f}" for v in b) + "\n")


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(description="Write a seeded fibers file.")
    ap.add_argument("kind", choices=["cube", "ball", "gen2"])
    ap.add_argument("fibers", type=int)
    ap.add_argument("--size", type=float, default=None,
                    help="cube side or ball radius (default: number density 0.1)")
    ap.add_argument("--min-dist", type=float, default=0.4,
                    help="minimum centre distance (cube, ball) or segment distance (gen2)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("-o", "--output", default=None, help="default: XcT_gen<M>.in")
    args = ap.parse_args(argv)
    M = args.fibers
    if args.kind == "cube":
        pos, ori = cube(M, args.size or (M / 0.1) ** (1.0 / 3.0), args.min_dist, args.seed)
    elif args.kind == "ball":
        pos, ori = ball(M, args.size or (3.0 * M / (4.0 * math.pi * 0.1)) ** (1.0 / 3.0), args.min_dist, args.seed)
    else:
        pos, ori = grown_sphere(M, args.min_dist, args.seed)
    out = args.output or f"XcT_gen{M}.in"
    write_fibers(out, pos, ori)
    print(out)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
