"""Pins the FP64 oracle (oracle/rf_oracle.c) before anything is compared with it.

Each check is independent of the oracle's own formulas:
  * tables        -> numpy's Gauss-Legendre rule, scipy's Legendre values,
                     lambda_k = 2 H_{k+1} (harmonic numbers);
  * K(R)          -> its eigen-decomposition (R is an eigenvector with
                     2/r - 4 eps^2/r^3, the orthogonal plane 1/r + 2 eps^2/r^3);
  * moments       -> adaptive scipy.integrate.quad in both recurrence branches
                     and the collinear branch;
  * numerical G   -> the analytical G (independent code path) for separated fibers;
  * M = 1         -> closed form of the isolated fiber;
  * everything    -> the reference's golden vectors (tests/reference/100_*).
"""
import numpy as np
import pytest
from scipy import integrate
from scipy.special import eval_legendre

from helpers import FIELDS, VARIANTS, golden, random_scene, ref100, rel_err


def test_tables_closed_forms():
    from oracle.pyoracle import Oracle

    for nqi in (1, 4, 8, 16):
        qp, qw, leg, lam, eig = Oracle.tables(nqi=nqi, N=8, eps=0.01)
        x3, w3 = np.polynomial.legendre.leggauss(3)
        h = 2.0 / nqi
        pts = np.concatenate([-1 + (k + 0.5) * h + x3 * h / 2 for k in range(nqi)])
        wts = np.concatenate([w3 * h / 2 for _ in range(nqi)])
        assert np.allclose(qp, pts, rtol=0, atol=1e-15)
        assert np.allclose(qw, wts, rtol=0, atol=1e-15)
        for k in range(8):
            assert np.allclose(leg[k], eval_legendre(k + 1, qp), rtol=0, atol=1e-13)
        # composite rule integrates degree-5 polynomials exactly
        assert abs(np.sum(qw * qp ** 4) - 2.0 / 5.0) < 1e-14
    _, _, _, lam, eig = Oracle.tables(nqi=8, N=8, eps=0.01)
    H = np.cumsum(1.0 / np.arange(1, 10))  # H[k] = H_{k+1}
    assert np.allclose(lam, 2 * H[:8], rtol=1e-15)
    d = -np.log(1e-4 * np.e)
    assert np.allclose(eig, ((d - 2 - lam) / 2) / (d - lam), rtol=1e-15)


@pytest.mark.parametrize("seed", range(5))
def test_stokes_kernel_eigenstructure(oracle64, seed):
    rng = np.random.default_rng(seed)
    R = rng.normal(size=3) * rng.uniform(0.05, 20)
    eps = 0.01
    k = oracle64.stokes_kernel(R, eps)
    K = np.array([[k[0], k[3], k[4]], [k[3], k[1], k[5]], [k[4], k[5], k[2]]])
    r = np.linalg.norm(R)
    assert np.allclose(K @ R, (2 / r - 4 * eps ** 2 / r ** 3) * R, rtol=1e-12)
    perp = np.cross(R, rng.normal(size=3))
    assert np.allclose(K @ perp, (1 / r + 2 * eps ** 2 / r ** 3) * perp, rtol=1e-12)


def _quad_moments(b, l, nmax):
    out = np.zeros((3, nmax + 1))
    for mi, m in enumerate((1, 3, 5)):
        for n in range(nmax + 1):
            out[mi, n] = integrate.quad(lambda s: s ** n * (s * s + b * s + l) ** (-m / 2), -1, 1,
                                        epsabs=1e-14, epsrel=1e-13, limit=200)[0]
    return out


@pytest.mark.parametrize("R0,t", [
    ((0.8, 0.3, -0.2), (0.0, 0.6, 0.8)),       # near branch (|R0|^2 < 10)
    ((2.5, 1.0, 0.4), (0.6, 0.0, 0.8)),        # near branch, larger
    ((3.0, 2.0, 1.5), (0.0, 0.0, 1.0)),        # far branch (backward recurrence)
    ((8.0, -5.0, 2.0), (0.48, 0.6, 0.64)),     # far branch
])
def test_moments_against_quadrature(oracle64, R0, t):
    R0, t = np.array(R0), np.array(t) / np.linalg.norm(t)
    b, l = -2 * R0 @ t, R0 @ R0
    N = 5
    I1, I3, I5 = oracle64.moments(b, l, N + 2, N)
    ref = _quad_moments(b, l, N + 2)
    for mine, r in zip((I1, I3, I5), ref):
        assert np.allclose(mine[: N + 3], r, rtol=2e-7, atol=1e-15), (mine[: N + 3], r)


def test_moments_collinear_branch(oracle64):
    """Target point on the source fiber's axis beyond its tip (l - b^2/4 = 0).
    The reference's closed forms (compute_inner_integral_analytically.incl:47-62)
    hold where 2s + b > 0 on [-1, 1]; beyond the other tip they return the
    integrals with flipped sign -- restated as the reference computes them."""
    t = np.array([0.0, 0.0, 1.0])
    R0 = np.array([0.0, 0.0, -1.7])  # b = 3.4 > 2: the formulas' side
    b, l = -2 * R0 @ t, R0 @ R0
    I1, I3, I5 = oracle64.moments(b, l, 4, 2)
    ref = _quad_moments(b, l, 1)
    assert np.allclose(I3[:2], ref[1, :2], rtol=1e-9)
    assert np.allclose(I5[:2], ref[2, :2], rtol=1e-9)
    R0 = -R0  # b = -3.4: the reference's sign flip
    b, l = -2 * R0 @ t, R0 @ R0
    _, I3, I5 = oracle64.moments(b, l, 4, 2)
    ref = _quad_moments(b, l, 0)
    assert np.isclose(I3[0], -ref[1, 0], rtol=1e-9) and np.isclose(I5[0], -ref[2, 0], rtol=1e-9)


def test_numerical_vs_analytical_assembly(oracle64):
    """The two inner-integral code paths agree where the 24-point quadrature is
    accurate (well separated fibers)."""
    from oracle.pyoracle import make_params

    M = 6
    pos, ori = random_scene(M, side=30.0, seed=5, min_dist=8.0)
    p = make_params(M)
    A_n, b_n = oracle64.assemble(p, pos, ori, analytical=False)
    A_a, b_a = oracle64.assemble(p, pos, ori, analytical=True)
    off = ~np.eye(A_n.shape[0], dtype=bool)
    # the 24-point rule's own error on the P_5 source modes is ~2e-6 of max|A_ij|;
    # a wrong coefficient in any mode shows up at >= 1e-3
    assert np.abs(A_n - A_a)[off].max() < 1e-5 * np.abs(A_a[off]).max()
    assert rel_err(b_n, b_a) < 1e-9


def test_single_fiber(oracle64):
    from oracle.pyoracle import make_params

    t = np.array([0.6, 0.0, 0.8])
    p = make_params(1)
    A, b = oracle64.assemble(p, [0.0, 0.0, 0.0], t)
    assert np.array_equal(A, np.eye(15)) and not b.any()
    v, w = oracle64.velocities(p, [0.0, 0.0, 0.0], t, np.zeros(15))
    d = -np.log(1e-4 * np.e)
    F = np.array([0.0, 0.0, -1.0])
    assert np.allclose(v, (0.5 / d) * ((d + 2) * F + (d - 2) * (t @ F) * t), rtol=1e-14)
    assert not w.any()


@pytest.mark.parametrize("variant", VARIANTS)
def test_oracle_reproduces_golden_vectors(oracle64, variant):
    """Two timesteps of the reference scene reproduce every golden vector to
    the print precision of the fixtures (F16.8: 5e-9), GMRES iterates included."""
    from oracle.pyoracle import make_params

    M, pos, ori = ref100()
    rec = []
    oracle64.simulate(make_params(M), pos, ori, 2, solver="direct" if variant.endswith("direct") else "gmres",
                      analytical=variant.startswith("analytic"), gmres=(10, 1000, 1e-5), record=rec)
    names = {"BVec": "b", "XVec": "x", "TRANSVel": "v", "ROTVel": "w", "POS": "pos", "ORIENT": "ori"}
    for step in range(2):
        for f in FIELDS:
            err = np.abs(rec[step][names[f]] - golden(variant, step, f)).max()
            assert err < 6e-9, f"{variant} step {step} {f}: {err:.2e}"


def test_golden_solution_satisfies_oracle_system(oracle64):
    """A of the oracle applied to the golden x reproduces the golden b: pins A
    entrywise through the solution (the reference's AMat fixtures are absent)."""
    from oracle.pyoracle import make_params

    M, pos, ori = ref100()
    A, b = oracle64.assemble(make_params(M), pos, ori)
    x = golden("numeric_direct", 0, "XVec")
    r = A @ x - golden("numeric_direct", 0, "BVec")
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-7


def test_planted_mistakes_are_caught(oracle64):
    """Plausible mistakes in the restated formulas fail the golden checks:
    a dropped doublet term, a wrong force sign, swapped mode weights."""
    from oracle.pyoracle import make_params

    M, pos, ori = ref100()
    b_gold = golden("numeric_direct", 0, "BVec")
    # doublet term dropped: eps -> 0 changes K by 2 eps^2 (I/r^3 - 3RR^T/r^5)
    _, b_nodoublet = oracle64.assemble(make_params(M, slenderness=1e-9), pos, ori)
    assert np.abs(b_nodoublet - b_gold).max() > 1e-6
    # wrong sign of the external force
    _, b_flip = oracle64.assemble(make_params(M, force=(0.0, 0.0, 1.0)), pos, ori)
    assert np.abs(b_flip - b_gold).max() > 1e-3
    # transposed orientation/position roles
    _, b_swap = oracle64.assemble(make_params(M), ori, pos)
    assert np.abs(b_swap - b_gold).max() > 1e-3


def test_fp32_build_tracks_fp64(oracle32, oracle64):
    """The FP32 CPU baseline restates the same arithmetic in REAL*4."""
    from oracle.pyoracle import make_params

    M, pos, ori = ref100()
    rec = []
    oracle32.simulate(make_params(M), pos, ori, 2, solver="gmres", record=rec)
    assert rel_err(rec[1]["pos"], golden("numeric_gmres", 1, "POS")) < 1e-5
    assert rel_err(rec[1]["x"], golden("numeric_gmres", 1, "XVec")) < 1e-4
