// rf_device.h -- device-side building blocks shared by the rigid-fiber kernels.
//
// * Quadrature and Legendre tables are compile-time constants for each
//   (N, intervals) instantiation, so every table lookup in the unrolled inner
//   loops becomes an immediate operand (no loads).  Values are computed in FP64
//   and rounded to FP32 exactly as the reference does
//   (fortran/fibers.f95:353-440; src/simulation.cu:153-284).
// * The Stokeslet+doublet kernel K(R) of
//   fortran/compute_inner_integral_numerically.incl:18-23, written as
//   K = a I + c R R^T with a = 1/r + 2 eps^2/r^3, c = 1/r^3 - 6 eps^2/r^5.
// * The moment recurrences of fortran/compute_inner_integral_analytically.incl:12-127,
//   evaluated in FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rf {

constexpr int kMaxTerms = 8;
// sqrt(15)/5, the outer node of the 3-point Gauss-Legendre rule.
constexpr double kGaussNode = 0.77459666924148337703585307995648;

// Monomial coefficients of P_0..P_8 (coef[f][p] multiplies s^p).
struct LegendreCoef {
  double c[9][9];
};
constexpr LegendreCoef kLeg = {{
    {1, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 1, 0, 0, 0, 0, 0, 0, 0},
    {-0.5, 0, 1.5, 0, 0, 0, 0, 0, 0},
    {0, -1.5, 0, 2.5, 0, 0, 0, 0, 0},
    {3.0 / 8, 0, -30.0 / 8, 0, 35.0 / 8, 0, 0, 0, 0},
    {0, 15.0 / 8, 0, -70.0 / 8, 0, 63.0 / 8, 0, 0, 0},
    {-5.0 / 16, 0, 105.0 / 16, 0, -315.0 / 16, 0, 231.0 / 16, 0, 0},
    {0, -35.0 / 16, 0, 315.0 / 16, 0, -693.0 / 16, 0, 429.0 / 16, 0},
    {35.0 / 128, 0, -1260.0 / 128, 0, 6930.0 / 128, 0, -12012.0 / 128, 0, 6435.0 / 128},
}};

__host__ __device__ constexpr double legendre_value(int f, double s) {
  double acc = 0.0, p = 1.0;
  for (int k = 0; k <= f; ++k) {
    acc += kLeg.c[f][k] * p;
    p *= s;
  }
  return acc;
}

// Composite 3-point Gauss-Legendre rule on NQI equal intervals of [-1, 1] and
// the weighted Legendre table W[m][q] = w_q * P_m(s_q), m = 0..N.
template <int N, int NQI>
struct Quad {
  static constexpr int Q = 3 * NQI;
  float s[Q];                 // nodes (FP64 -> FP32, as the reference tables)
  float w[Q];                 // weights
  float P[N + 1][Q];          // P_m(s_q), m = 0..N
  float W[N + 1][Q];          // w_q * P_m(s_q)
  double sd[Q];               // nodes in FP64 (analytical path)
  constexpr Quad() : s(), w(), P(), W(), sd() {
    const double h = 2.0 / NQI;
    const double gp[3] = {-kGaussNode, 0.0, kGaussNode};
    const double gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    for (int iv = 0; iv < NQI; ++iv) {
      const double lo = -1.0 + iv * h;
      for (int k = 0; k < 3; ++k) {
        const double node = (2.0 * lo + h + gp[k] * h) / 2.0;
        s[3 * iv + k] = static_cast<float>(node);
        sd[3 * iv + k] = node;
        w[3 * iv + k] = static_cast<float>(gw[k] / NQI);
      }
    }
    for (int m = 0; m <= N; ++m)
      for (int q = 0; q < Q; ++q) {
        P[m][q] = static_cast<float>(legendre_value(m, sd[q]));
        W[m][q] = w[q] * P[m][q];
      }
  }
};

// Far field of the analytical inner integral.  When the target node lies at
// least kFarField[N] from the source fiber's centre (|R0| >= kFarField[N]),
// the integrand P_m(s) K(R0 - s t) is analytic on a Bernstein ellipse around
// [-1, 1] far enough out that the 8-point Gauss-Legendre rule integrates it,
// for every mode m <= N, to < 1e-8 of the largest G_0 component (worst
// direction at the threshold; the FP32 rounding of the kernel values, ~1e-7,
// dominates).  Higher modes need a larger radius.  The rule's nodes come in
// mirror pairs +-s_k, so even modes take K(+s) + K(-s), odd modes K(+s) - K(-s).
constexpr double kFarField[kMaxTerms + 1] = {0.0, 3.0, 3.0, 3.0, 3.0, 3.0, 3.5, 4.5, 5.0};
constexpr double kGL8Node[4] = {0.18343464249564978, 0.52553240991632899, 0.79666647741362674,
                                0.96028985649753623};
constexpr double kGL8Weight[4] = {0.36268378337836177, 0.31370664587788705, 0.22238103445337434,
                                  0.10122853629037669};
template <int N>
struct FarQuad {
  float s[4];          // positive nodes
  float W[N + 1][4];   // w_k P_m(s_k), m = 0..N
  constexpr FarQuad() : s(), W() {
    for (int k = 0; k < 4; ++k) {
      s[k] = static_cast<float>(kGL8Node[k]);
      for (int m = 0; m <= N; ++m) W[m][k] = static_cast<float>(kGL8Weight[k] * legendre_value(m, kGL8Node[k]));
    }
  }
};

// Node s_q in FP64 and weight w_q in FP32 for a runtime node index (the
// per-lane target node), computed exactly as Quad<> does.
template <int NQI>
__device__ __forceinline__ double node_fp64(int q) {
  const double h = 2.0 / NQI;
  const int iv = q / 3, k = q - 3 * (q / 3);
  const double lo = -1.0 + iv * h;
  const double gp = (k == 0) ? -kGaussNode : ((k == 1) ? 0.0 : kGaussNode);
  return (2.0 * lo + h + gp * h) / 2.0;
}
template <int NQI>
__device__ __forceinline__ float quad_weight_fp32(int q) {
  const int k = q - 3 * (q / 3);
  return static_cast<float>(((k == 1) ? 8.0 / 9.0 : 5.0 / 9.0) / NQI);
}

// Per-run scalars that are not compile-time: d = -ln(eps^2 e), D1, the mode
// factors gamma_k and eigen_k, eps^2 and F/2 (fortran/fibers.f95:141-149,
// fortran/assemble_matrix.incl:101).
struct PhysConsts {
  float D1;
  float eps2;
  float gamma[kMaxTerms];   // gamma_k, k >= 1 (index 0 unused)
  float eigen[kMaxTerms];
  float Fh[3];              // F/2
  double eps2_d;
  double Fh_d[3];
  double d;                 // -ln(eps^2 e)
};

// K(R) components (xx, yy, zz, xy, xz, yz) in FP32.
__device__ __forceinline__ void stokes_kernel(float rx, float ry, float rz, float eps2, float K[6]) {
  const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
  const float inv = rsqrtf(r2);
  const float inv2 = inv * inv;
  const float inv3 = inv2 * inv;
  const float a = fmaf(2.0f * eps2, inv3, inv);                 // 1/r + 2eps^2/r^3
  const float c = inv3 * fmaf(-6.0f * eps2, inv2, 1.0f);        // 1/r^3 - 6eps^2/r^5
  const float cx = c * rx, cy = c * ry, cz = c * rz;
  K[0] = fmaf(cx, rx, a);
  K[1] = fmaf(cy, ry, a);
  K[2] = fmaf(cz, rz, a);
  K[3] = cx * ry;
  K[4] = cx * rz;
  K[5] = cy * rz;
}

// log(P / Q) for normal P, Q > 0 without the division P / Q and without a
// general logarithm: P = 2^eP mP, Q = 2^eQ mQ (exact), the mantissa ratio
// brought into [sqrt(1/2), sqrt(2)) by one exact halving or doubling, then
// log(P / Q) = e ln 2 + 2 atanh(y), y = (mP - mQ) / (mP + mQ), from the odd
// Taylor series 2 sum_{k <= 11} y^(2k+1) / (2k+1): |y| <= 3 - 2 sqrt(2), so the
// first omitted term is below 1e-18 of the sum.  mP - mQ is exact when the
// ratio is near 1, so the result keeps its relative precision there (where
// log(P / Q) of the rounded ratio would not).  Other inputs (zero, subnormal,
// infinite, NaN) take the library logarithm of the ratio.
__device__ __forceinline__ double log_ratio(double P, double Q) {
  constexpr double kMin = 2.2250738585072014e-308;  // smallest normal double
  constexpr double kMax = 1.7976931348623157e308;
  if (!(P >= kMin && Q >= kMin && P <= kMax && Q <= kMax)) return log(P / Q);
  int eP, eQ;
  double mP = frexp(P, &eP), mQ = frexp(Q, &eQ);
  const int adj = (mP > 1.4142135623730951 * mQ) - (mP < 0.70710678118654752 * mQ);
  mP = ldexp(mP, -adj);
  const int e = eP - eQ + adj;
  const double num = mP - mQ, den = mP + mQ;
  // y = num / den: reciprocal refined twice, then one residual correction
  double r = __builtin_amdgcn_rcp(den);
  r = fma(fma(-den, r, 1.0), r, r);
  r = fma(fma(-den, r, 1.0), r, r);
  double y = num * r;
  y = fma(fma(-den, y, num), r, y);
  const double z = y * y;
  double p = 1.0 / 23.0;
  p = fma(p, z, 1.0 / 21.0);
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  const double t = 2.0 * y;
  const double tail = fma(t * z, p, t);  // 2 y (1 + z p)
  // ln 2 as a double plus the remainder, so e ln 2 loses nothing for e != 0
  const double de = static_cast<double>(e);
  return fma(de, 0.69314718055994529, fma(de, 2.3190468138462996e-17, tail));
}

// Moments I1_n, I3_n, I5_n (n = 0..N+2) of int_{-1}^{1} s^n (s^2 + b s + l)^{-m/2} ds
// as the reference computes them: closed forms for n = 0, 1 (collinear branch
// when l - b^2/4 < 1e-7), forward recurrence when l < 10, else a backward
// recurrence started from zeros at n = 27 keeping n < N + 3.
template <int N>
__device__ __forceinline__ void moments(double b, double l, double I1[N + 3], double I3[N + 3], double I5[N + 3]) {
  // Divisions by the recurrence's integer constants are multiplications by
  // reciprocals folded at compile time; the remaining divisions are 1/m and
  // 1/l (1/pu, 1/pl on the collinear branch), the logarithm is log_ratio.
  const double m = l - 0.25 * b * b;
  const double qu = 1.0 + b + l, ql = 1.0 - b + l;  // |X(+-1) - x_j|^2 > 0 for distinct fibers
  const double iuu = rsqrt(qu), iul = rsqrt(ql);
  const double uu = qu * iuu, ul = ql * iul;
  const double pu = 2.0 + b, pl = -2.0 + b;
  // log|pu + 2uu| - log|pl + 2ul|
  I1[0] = log_ratio(fabs(pu + 2.0 * uu), fabs(pl + 2.0 * ul));
  I1[1] = uu - ul - 0.5 * b * I1[0];
  if (m < 1e-7) {
    const double ipu = 1.0 / pu, ipl = 1.0 / pl;
    const double ipu2 = ipu * ipu, ipl2 = ipl * ipl;
    I3[0] = -2.0 * ipu2 + 2.0 * ipl2;
    I5[0] = -4.0 * ipu2 * ipu2 + 4.0 * ipl2 * ipl2;
    I5[1] = (-8.0 / 3.0) * ipu2 * ipu + (8.0 / 3.0) * ipl2 * ipl - 0.5 * b * I5[0];
  } else {
    const double im = 1.0 / m;
    const double iu3 = iuu * iuu * iuu, il3 = iul * iul * iul;
    I3[0] = 0.5 * im * (pu * iuu - pl * iul);
    I5[0] = (im * (1.0 / 6.0)) * (pu * iu3 - pl * il3) + (2.0 / 3.0) * im * I3[0];
    I5[1] = (im * (1.0 / 6.0)) * (-(b + 2.0 * l) * iu3 + (-b + 2.0 * l) * il3) - (1.0 / 3.0) * b * im * I3[0];
  }
  I3[1] = -iuu + iul - 0.5 * b * I3[0];
  if (l < 10.0) {
#pragma unroll
    for (int n = 2; n <= N + 2; ++n) {
      // s^(n-1) at s = +1 and s = -1
      const double sl = (n & 1) ? 1.0 : -1.0;
      const double inv_n = 1.0 / n;                                  // folded
      const double c1 = (1.0 - 2.0 * n) / (2.0 * n);                 // folded
      const double c2 = (n - 1.0) / n;                               // folded
      I1[n] = (uu - sl * ul) * inv_n + c1 * b * I1[n - 1] - c2 * l * I1[n - 2];
      I3[n] = I1[n - 2] - b * I3[n - 1] - l * I3[n - 2];
      I5[n] = I3[n - 2] - b * I5[n - 1] - l * I5[n - 2];
    }
  } else {
    const double il = 1.0 / l;
    const double dm = uu - ul, dp = uu + ul;  // uu - s^(n+1) ul for even / odd n+1
    // Zero start at n = 27 as the reference has it, later where l is large:
    // the start values' error shrinks ~l^(-1/2) per step, and over random
    // (b, l), |b| up to 2 sqrt(l), these starts change every Legendre
    // combination L by < 3e-13 of its largest term, five orders below the
    // FP32 rounding of the G they form (the reference's own start at 27
    // still moves them by ~2e-7 for l in [10, 20], so l < 100 keeps 27).
    const int start = l < 100.0    ? 27
                      : l < 125.0  ? 22
                      : l < 244.0  ? 20
                      : l < 745.0  ? 18
                      : l < 2842.0 ? 16
                      : l < 8673.0 ? 14
                                   : 12;
    // I_n of each family in slot n % 3: a step writes its own slot and reads
    // the other two, so a skipped step (a prefix of the sweep) leaves a zero
    // and the sweep needs no register copies.
    double r1[3] = {0.0, 0.0, 0.0}, r3[3] = {0.0, 0.0, 0.0}, r5[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int n = 27; n >= 2; --n) {
      if (n > 12 && n > start) continue;  // still all zeros
      const int k0 = n % 3, k1 = (n + 1) % 3, k2 = (n + 2) % 3;
      const double rn = n;
      // s^(n+1) at s = -1 is +1 for odd n
      const double d = (n & 1) ? dm : dp;
      const double cA = 1.0 / (rn + 1.0);                                                   // folded
      const double cB = (rn + 2.0) / (rn + 1.0) * ((1.0 - 2.0 * (rn + 2.0)) / (2.0 * (rn + 2.0)));  // folded
      const double cC = (rn + 2.0) / (rn + 1.0);                                             // folded
      r1[k0] = il * (cA * d + cB * b * r1[k1] - cC * r1[k2]);
      r3[k0] = il * (r1[k0] - b * r3[k1] - r3[k2]);
      r5[k0] = il * (r3[k0] - b * r5[k1] - r5[k2]);
      if (n < N + 3) {
        I1[n] = r1[k0];
        I3[n] = r3[k0];
        I5[n] = r5[k0];
      }
    }
  }
}

// G_f = L01 I + L03 R0R0^T - L13 (t R0^T + R0 t^T) + L23 t t^T
//     + 2 eps^2 (L03 I - 3 L05 R0R0^T + 3 L15 (t R0^T + R0 t^T) - 3 L25 t t^T)
// (fortran/compute_inner_integral_analytically.incl:207-260), FP64, from the
// symmetric products of the node shared by every mode: RR = R0 R0^T,
// TR = t R0^T + R0 t^T, TT = t t^T (components xx, yy, zz, xy, xz, yz).
struct NodeProducts {
  double RR[6], TR[6], TT[6];
};

__device__ __forceinline__ NodeProducts node_products(const double R0[3], const double t[3]) {
  constexpr int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  NodeProducts q;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int a = ia[c], bb = ib[c];
    q.RR[c] = R0[a] * R0[bb];
    q.TR[c] = fma(t[a], R0[bb], R0[a] * t[bb]);
    q.TT[c] = t[a] * t[bb];
  }
  return q;
}

template <int F>
__device__ __forceinline__ void analytic_g(const NodeProducts& q, double eps2, const double* I1, const double* I3,
                                           const double* I5, double G[6]) {
  double L01 = 0, L03 = 0, L05 = 0, L13 = 0, L15 = 0, L23 = 0, L25 = 0;
#pragma unroll
  for (int p = 0; p <= F; ++p) {
    constexpr LegendreCoef lc = kLeg;
    const double cf = lc.c[F][p];
    if (cf != 0.0) {
      L01 += cf * I1[p];
      L03 += cf * I3[p];
      L05 += cf * I5[p];
      L13 += cf * I3[p + 1];
      L15 += cf * I5[p + 1];
      L23 += cf * I3[p + 2];
      L25 += cf * I5[p + 2];
    }
  }
  const double two_e2 = 2.0 * eps2;
  const double iso = L01 + two_e2 * L03;           // coefficient of I
  const double crr = L03 - 3.0 * two_e2 * L05;     // of R0 R0^T
  const double ctr = -L13 + 3.0 * two_e2 * L15;    // of (t R0^T + R0 t^T)
  const double ctt = L23 - 3.0 * two_e2 * L25;     // of t t^T
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double g0 = (c < 3) ? fma(crr, q.RR[c], iso) : crr * q.RR[c];
    G[c] = fma(ctt, q.TT[c], fma(ctr, q.TR[c], g0));
  }
}

// Component index of the symmetric 3x3 pair (a, b) in (xx, yy, zz, xy, xz, yz).
__host__ __device__ constexpr int sym_index(int a, int b) {
  return (a == b) ? a : ((a + b == 1) ? 3 : ((a + b == 2) ? 4 : 5));
}

}  // namespace rf
