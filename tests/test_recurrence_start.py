"""The far branch of the analytical moments (l >= 10) runs the backward
recurrence of fortran/compute_inner_integral_analytically.incl:91-125 from a
zero start.  The reference starts at n = 27 for every l; the device starts
later where l is large (the start values' error shrinks ~l^(-1/2) per step).
This test reads the start table from the device source
(rigidfibers_amd/csrc/rf_device.h) and checks, with the recurrence restated
in FP64 numpy, that each start changes every Legendre combination L of the
moments (the sums analytic_g forms, P_1..P_5 and GF) by less than 3e-13 of its
largest term against the reference's start at 27, over random (b, l) of each
range including |b| -> 2 sqrt(l) -- five orders below the FP32 rounding of the
G that the L form.  It also checks that the table is not looser than needed
by accident: one step less misses the bound in the ranges of the M=5000 ball.
"""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = 5
# P_F(s) = sum_p C[F][p] s^p, F = 0..5
C = np.zeros((6, 6))
C[0][0] = 1
C[1][1] = 1
C[2][[0, 2]] = [-0.5, 1.5]
C[3][[1, 3]] = [-1.5, 2.5]
C[4][[0, 2, 4]] = [3 / 8, -30 / 8, 35 / 8]
C[5][[1, 3, 5]] = [15 / 8, -70 / 8, 63 / 8]


def device_table():
    """(upper l bound, start) pairs of the device's far-branch start."""
    src = open(os.path.join(ROOT, "rigidfibers_amd", "csrc", "rf_device.h")).read()
    m = re.search(r"const int start = (.*?);", src, re.S)
    assert m, "start table not found"
    pairs = [(float(a), int(b)) for a, b in re.findall(r"l < ([0-9.]+)\s*\?\s*(\d+)", m.group(1))]
    last = int(re.findall(r":\s*(\d+)\s*$", m.group(1).strip())[0])
    return pairs, last


def backward(b, l, start):
    uu, ul = np.sqrt(1 + b + l), np.sqrt(1 - b + l)
    il, dm, dp = 1 / l, uu - ul, uu + ul
    z = np.zeros_like(b)
    i1 = [z, z]  # I_{n+1}, I_{n+2}
    i3 = [z, z]
    i5 = [z, z]
    I1 = np.zeros((N + 3,) + b.shape)
    I3, I5 = I1.copy(), I1.copy()
    for n in range(27, 1, -1):
        if n > start:
            continue
        rn = float(n)
        d = dm if n & 1 else dp
        cA, cB, cC = 1 / (rn + 1), (rn + 2) / (rn + 1) * ((1 - 2 * (rn + 2)) / (2 * (rn + 2))), (rn + 2) / (rn + 1)
        a1 = il * (cA * d + cB * b * i1[0] - cC * i1[1])
        a3 = il * (a1 - b * i3[0] - i3[1])
        a5 = il * (a3 - b * i5[0] - i5[1])
        if n < N + 3:
            I1[n], I3[n], I5[n] = a1, a3, a5
        i1, i3, i5 = [a1, i1[0]], [a3, i3[0]], [a5, i5[0]]
    return I1, I3, I5


def combinations(I1, I3, I5):
    out = []
    for F in range(6):
        c = C[F]
        out.append(np.array([sum(c[p] * I[p + o] for p in range(F + 1))
                             for I, o in ((I1, 0), (I3, 0), (I5, 0), (I3, 1), (I5, 1), (I3, 2), (I5, 2))]))
    return out


def worst(b, l, start):
    ref, got = combinations(*backward(b, l, 27)), combinations(*backward(b, l, start))
    return max(np.max(np.abs(g - r) / np.maximum(np.abs(r).max(axis=0), 1e-300)) for g, r in zip(got, ref))


def samples(lo, hi, n, rng):
    l = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    r = np.sqrt(l)
    b = rng.uniform(-2, 2, n) * r * np.sqrt(rng.uniform(0, 1, n))
    b[: n // 20] = 2 * r[: n // 20] * np.sign(rng.uniform(-1, 1, n // 20)) * (1 - 1e-6)
    return b, l


def test_start_table_is_within_bound():
    pairs, last = device_table()
    assert pairs[0] == (100.0, 27)  # the reference's start below l = 100
    rng = np.random.default_rng(11)
    bounds = [b for b, _ in pairs] + [2e4]
    starts = [s for _, s in pairs[1:]] + [last]
    for lo, hi, start in zip(bounds[:-1], bounds[1:], starts):
        b, l = samples(lo, hi, 4000, rng)
        e = worst(b, l, start)
        print(f"l in [{lo:.0f}, {hi:.0f}): start {start}, worst {e:.1e}")
        assert e < 3e-13, (lo, hi, start, e)


def test_one_step_less_misses_the_bound():
    """Where most pairs of the M=5000 ball lie (l ~ 1e3) the chosen start is
    the smallest that passes: the test above would catch a table shifted by
    one step."""
    pairs, last = device_table()
    rng = np.random.default_rng(12)
    for lo, hi in ((244.0, 745.0), (745.0, 2842.0)):
        start = [s for (ub, s) in pairs if ub > lo][0]
        b, l = samples(lo, hi, 4000, rng)
        assert worst(b, l, start) < 3e-13
        assert worst(b, l, start - 1) > 3e-13
