"""The far-field rule of the analytical inner integral (CPU).

With --analytical the inner integral G_m(q_i) = int_{-1}^{1} K(R0 - s t_j)
P_m(s) ds (fortran/compute_inner_integral_analytically.incl) is formed from
FP64 moment recurrences when the target node is near the source fiber and,
when |R0| >= kFarField[N], by the 8-point Gauss-Legendre rule of
rigidfibers_amd/csrc/rf_device.h.  These checks read the rule's thresholds,
nodes and weights from the device source and pin them:
  * nodes and weights are the 8-point Gauss-Legendre rule (numpy);
  * for every expansion order N = 1..8, at its threshold and beyond, in
    every direction sampled, the rule reproduces the integral of every mode
    m <= N (a 200-point rule, itself checked against scipy.integrate.quad)
    to 1e-8 of the largest G_0 component, far below the FP32 rounding of
    the G it forms;
  * the bound has teeth: 6 points, or N = 5's threshold for N = 8, miss it.
"""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EPS = 0.01
N = 8  # modes formed by _g: 0..N


def _device_rule():
    """(thresholds indexed by N, positive nodes, their weights)."""
    with open(os.path.join(ROOT, "rigidfibers_amd", "csrc", "rf_device.h")) as fh:
        src = fh.read()
    thr = [float(v) for v in re.search(r"kFarField\[kMaxTerms \+ 1\] = \{([^}]*)\}", src).group(1).split(",")]
    nodes = [float(v) for v in re.search(r"kGL8Node\[4\] = \{([^}]*)\}", src).group(1).split(",")]
    weights = [float(v) for v in re.search(r"kGL8Weight\[4\] = \{([^}]*)\}", src).group(1).split(",")]
    return thr, np.array(nodes), np.array(weights)


def _kernel(R):
    """K(R) = I/r + R R^T/r^3 + 2 eps^2 (I/r^3 - 3 R R^T/r^5), components xx yy zz xy xz yz."""
    r2 = (R ** 2).sum(-1)
    inv = 1.0 / np.sqrt(r2)
    inv3 = inv ** 3
    a = inv + 2 * EPS ** 2 * inv3
    c = inv3 - 6 * EPS ** 2 * inv3 * inv ** 2
    x, y, z = R[..., 0], R[..., 1], R[..., 2]
    return np.stack([a + c * x * x, a + c * y * y, a + c * z * z, c * x * y, c * x * z, c * y * z], -1)


def _g(R0, t, s, w, n=N):
    """G_m, m = 0..n, by the rule (s, w) over the source fiber."""
    k = _kernel(R0[None, :] - s[:, None] * t[None, :])
    P = np.stack([np.polynomial.legendre.legval(s, [0] * m + [1]) for m in range(n + 1)])
    return (P * w[None, :]) @ k


def _full(nodes, weights):
    return np.concatenate([-nodes[::-1], nodes]), np.concatenate([weights[::-1], weights])


def _worst(thr, s, w, n=N, samples=300, seed=0):
    rng = np.random.default_rng(seed)
    ref_s, ref_w = np.polynomial.legendre.leggauss(200)
    worst = 0.0
    for _ in range(samples):
        t = rng.normal(size=3)
        t /= np.linalg.norm(t)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        R0 = d * thr * (1.0 + 0.5 * rng.random() ** 4)  # mostly at the threshold itself
        ref = _g(R0, t, ref_s, ref_w, n)
        worst = max(worst, np.abs(_g(R0, t, s, w, n) - ref).max() / np.abs(ref[0]).max())
    return worst


def test_rule_is_gauss_legendre_8():
    _, nodes, weights = _device_rule()
    s, w = np.polynomial.legendre.leggauss(8)
    fs, fw = _full(nodes, weights)
    assert np.allclose(fs, s, rtol=0, atol=1e-15) and np.allclose(fw, w, rtol=0, atol=1e-15)


def test_reference_rule_against_quad():
    """The 200-point rule used as the exact integral agrees with adaptive
    quadrature on a few far geometries."""
    from scipy import integrate

    ref_s, ref_w = np.polynomial.legendre.leggauss(200)
    rng = np.random.default_rng(3)
    for _ in range(4):
        t = rng.normal(size=3)
        t /= np.linalg.norm(t)
        R0 = rng.normal(size=3)
        R0 *= 3.0 / np.linalg.norm(R0)
        g = _g(R0, t, ref_s, ref_w)
        for m in (0, 3, 8):
            for c in (0, 3):
                f = lambda s: _kernel(R0 - s * t)[c] * np.polynomial.legendre.legval(s, [0] * m + [1])  # noqa: E731
                v, _ = integrate.quad(f, -1, 1, epsabs=1e-15, epsrel=1e-13)
                assert abs(v - g[m, c]) <= 1e-12 * np.abs(g[0]).max()


@pytest.mark.parametrize("n", range(1, 9))
def test_far_field_accuracy_at_threshold(n):
    thr, nodes, weights = _device_rule()
    s, w = _full(nodes, weights)
    worst = _worst(thr[n], s, w, n)
    print(f"N={n}: 8-point rule at |R0| >= {thr[n]}: worst error {worst:.1e} of max |G_0|")
    assert worst <= 1e-8


@pytest.mark.parametrize("planted", ["6 points", "N=5 threshold for N=8"])
def test_planted_mistakes_fail(planted):
    thr, nodes, weights = _device_rule()
    if planted == "6 points":
        s, w = np.polynomial.legendre.leggauss(6)
        assert _worst(thr[5], s, w, 5) > 1e-8
    else:
        s, w = _full(nodes, weights)
        assert _worst(thr[5], s, w, 8) > 1e-8
