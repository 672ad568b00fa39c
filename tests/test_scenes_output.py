"""Scene generators and output files (CPU only).

  * segment_distance2 against a brute-force grid search, with parallel,
    crossing and collinear segments among the cases;
  * grown_sphere (the gen2 construction) keeps every pair of segments at
    least the minimum distance apart, is seeded and returns unit orientations;
  * current.map fields follow A's layout: the oracle's assembled A is exactly
    zero wherever the map leaves an entry unset and 1 on the unit diagonal;
  * the generator's fibers file reads back through the runner's reader.
"""
import os

import numpy as np
import pytest


def _brute_distance2(p, o, P, O, n=401):
    s = np.linspace(-1.0, 1.0, n)
    X = p[None, :] + s[:, None] * o[None, :]
    Y = P[None, :] + s[:, None] * O[None, :]
    d = X[:, None, :] - Y[None, :, :]
    return float(np.min(np.einsum("ijk,ijk->ij", d, d)))


def test_segment_distance_matches_brute_force():
    from rigidfibers_amd.scenes import segment_distance2

    rng = np.random.default_rng(3)
    cases = []
    for _ in range(40):
        p, P = rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)
        o, O = rng.normal(size=3), rng.normal(size=3)
        cases.append((p, o / np.linalg.norm(o), P, O / np.linalg.norm(O)))
    t = np.array([0.0, 0.0, 1.0])
    cases.append((np.zeros(3), t, np.array([0.5, 0.0, 0.3]), t))            # parallel, overlapping
    cases.append((np.zeros(3), t, np.array([0.5, 0.0, 5.0]), t))            # parallel, apart along t
    cases.append((np.zeros(3), t, np.zeros(3), np.array([1.0, 0.0, 0.0])))  # crossing: distance 0
    cases.append((np.zeros(3), t, np.array([0.0, 0.0, 3.0]), t))            # collinear, gap 1
    for p, o, P, O in cases:
        got = float(segment_distance2(p, o, P[None, :], O[None, :])[0])
        ref = _brute_distance2(p, o, P, O)
        # the grid (spacing 0.005 along both unit-half-length segments) can
        # only overestimate the minimum, by at most 0.01 in distance
        assert got <= ref + 1e-12
        assert np.sqrt(ref) - np.sqrt(got) <= 0.01
    assert float(segment_distance2(np.zeros(3), t, np.zeros((1, 3)), np.array([[1.0, 0.0, 0.0]]))[0]) == 0.0
    assert float(segment_distance2(np.zeros(3), t, np.array([[0.0, 0.0, 3.0]]), t[None, :])[0]) == pytest.approx(1.0)


def test_grown_sphere_keeps_segment_distance():
    from rigidfibers_amd.scenes import grown_sphere, segment_distance2

    M, d = 150, 0.4
    p, o = grown_sphere(M, d, seed=11)
    p2, o2 = grown_sphere(M, d, seed=11)
    assert np.array_equal(p, p2) and np.array_equal(o, o2)
    P, O = p.reshape(-1, 3), o.reshape(-1, 3)
    assert np.allclose(np.linalg.norm(O, axis=1), 1.0)
    for i in range(M):
        rest = np.arange(M) != i
        assert segment_distance2(P[i], O[i], P[rest], O[rest]).min() >= d * d * (1 - 1e-12)
    p3, _ = grown_sphere(M, d, seed=12)
    assert not np.array_equal(p, p3)


def test_mapping_follows_the_assembled_matrix(oracle64, tmp_path):
    from oracle.pyoracle import make_params
    from rigidfibers_amd.output import mapping_fields, write_mapping
    from rigidfibers_amd.scenes import cube

    M, N = 4, 3
    n = 3 * N * M
    pos, ori = cube(M, 6.0, min_dist=0.8, seed=2)
    A, _ = oracle64.assemble(make_params(M, N=N), pos, ori, False)
    rows, cols = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    f = mapping_fields(rows.ravel(), cols.ravel(), N)
    unset = (f[0] == -1).reshape(n, n)
    assert np.all(A[unset] == 0.0)
    assert np.all(np.diag(A) == 1.0)
    assert unset.sum() == M * ((3 * N) ** 2 - 3 * N)
    # a coupling entry: row (i=1, k_i=2, a=0), column (j=3, k_j=1, b=2)
    r, c = 3 * N * 1 + 3 * 2 + 0, 3 * N * 3 + 3 * 1 + 2
    assert tuple(f[:, r * n + c]) == (1, 3, 1, 2, 0, 2)
    assert A[r, c] != 0.0
    path = tmp_path / "current.map"
    write_mapping(str(path), M, N, rows_per_chunk=7)
    lines = path.read_text().splitlines()
    assert len(lines) == n * n
    assert lines[r * n + c] == f"[V]|1|3|1|2|0|2|{r}|{c}"
    assert lines[1] == "[V]|-1|-1|-1|-1|-1|-1|0|1"


def test_generator_file_reads_back(tmp_path):
    from rigidfibers_amd.config import read_fibers
    from rigidfibers_amd.scenes import main

    out = tmp_path / "XcT_gen20.in"
    assert main(["gen2", "20", "--min-dist", "0.3", "--seed", "4", "-o", str(out)]) == 0
    M, pos, ori = read_fibers(str(out))
    assert M == 20 and pos.shape == (60,) and ori.shape == (60,)
    assert np.allclose(np.linalg.norm(ori.reshape(-1, 3), axis=1), 1.0)
    out2 = tmp_path / "cube.in"
    assert main(["cube", "30", "--seed", "1", "-o", str(out2)]) == 0
    assert read_fibers(str(out2))[0] == 30
