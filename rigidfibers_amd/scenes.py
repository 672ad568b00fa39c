"""Seeded fiber configurations.

The reference's generators (tools/gen2.f95:29-102, tools/gen3.py:162-251) are
seeded from /dev/urandom and enforce a minimum segment distance; the scenes
here are reproducible: centres uniform in a cube or a ball, rejected while
closer than a minimum centre distance to an accepted one (cell-list check),
orientations uniform on the unit sphere.
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
