/*
 * rf_oracle.c -- CPU restatement of the reference rigid-fiber timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the libraries built from it.  The product path (rigidfibers_amd)
 * runs hand-written HIP kernels and never calls into this code.
 *
 * One source, two builds (see oracle/Makefile):
 *   librf_oracle64.so  RFO_REAL=double  -> the FP64 oracle every GPU result
 *                                          is compared against;
 *   librf_oracle32.so  RFO_REAL=float   -> an FP32 OpenMP restatement with the
 *                                          reference Fortran's data types, timed
 *                                          as the CPU baseline ("port").
 *
 * The algorithm follows the reference's Fortran path step by step:
 *   tables        fortran/fibers.f95:353-476
 *   constants     fortran/fibers.f95:141-149
 *   assembly      fortran/assemble_matrix.incl:1-161
 *   numerical G   fortran/compute_inner_integral_numerically.incl:1-40
 *   analytical G  fortran/compute_inner_integral_analytically.incl:1-338
 *   solve         fortran/solve_system.incl:1-59 (sgesv semantics; GMRES with
 *                 the stopping rule of fortran/sgmres.f:1240-1358)
 *   velocities    fortran/update_velocities.incl:1-89
 *   update        fortran/update_fibers_firststep.incl:1-10,
 *                 fortran/update_fibers.incl:1-17
 * The external force is uniform (fortran/fibers.f95:149 carries F/2).
 *
 * Layout: fiber vectors are [3*M] (x0,y0,z0,x1,...).  The system matrix is
 * column-major n x n, n = 3*N*M, row = 3*N*i + 3*k_i + a (target fiber i,
 * mode k_i, component a), column = 3*N*j + 3*k_j + b.
 *
 * Pinning (tests/test_oracle_*.py): tables against closed forms and numpy's
 * Gauss-Legendre rule; the kernel K against its eigen-decomposition; the
 * analytical moments against scipy.integrate.quad; numerical vs analytical
 * assembly for separated fibers; A, b, x, velocities and state against the
 * reference's golden vectors (tests/reference/100_* of the reference).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef RFO_REAL
#define RFO_REAL double
#endif
typedef RFO_REAL real;

#define RFO_MAX_Q 192
#define RFO_MAX_N 8

typedef struct rfo_params {
  int M;               /* number of fibers */
  int N;               /* terms in the force expansion, 1..8 */
  int nqi;             /* quadrature intervals; 3 Gauss points each */
  int threads;         /* OpenMP threads, <=0: runtime default */
  double slenderness;  /* epsilon */
  double dt;           /* timestep */
  double force[3];     /* external force F (the reference carries F/2) */
} rfo_params;

/* Monomial coefficients of the Legendre polynomials P_0..P_8 (the textbook
 * closed forms of fortran/fibers.f95:442-476), coef[f][p] multiplies s^p. */
static const double kLegendreCoef[9][9] = {
    {1, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 1, 0, 0, 0, 0, 0, 0, 0},
    {-0.5, 0, 1.5, 0, 0, 0, 0, 0, 0},
    {0, -1.5, 0, 2.5, 0, 0, 0, 0, 0},
    {3.0 / 8, 0, -30.0 / 8, 0, 35.0 / 8, 0, 0, 0, 0},
    {0, 15.0 / 8, 0, -70.0 / 8, 0, 63.0 / 8, 0, 0, 0},
    {-5.0 / 16, 0, 105.0 / 16, 0, -315.0 / 16, 0, 231.0 / 16, 0, 0},
    {0, -35.0 / 16, 0, 315.0 / 16, 0, -693.0 / 16, 0, 429.0 / 16, 0},
    {35.0 / 128, 0, -1260.0 / 128, 0, 6930.0 / 128, 0, -12012.0 / 128, 0, 6435.0 / 128}};

static double legendre_eval(int f, double s) {
  double acc = 0.0, p = 1.0;
  for (int k = 0; k <= f; ++k) {
    acc += kLegendreCoef[f][k] * p;
    p *= s;
  }
  return acc;
}

/* Quadrature points/weights, P_{k+1}(s_q) table [k*Q + q], lambda and eigen.
 * fortran/fibers.f95:353-440: composite 3-point Gauss-Legendre on nqi equal
 * intervals of [-1,1]; lambda_0 = 2, lambda_k = lambda_{k-1} + 2/(k+1);
 * eigen_k = ((d - e - lambda_k)/2) / (d - lambda_k) with d = -ln(eps^2 e), e = 2. */
int rfo_tables(int nqi, int N, double eps, double *qp, double *qw, double *leg,
               double *lambda, double *eigen) {
  if (nqi < 1 || 3 * nqi > RFO_MAX_Q || N < 1 || N > RFO_MAX_N) return -1;
  const int Q = 3 * nqi;
  const double g = sqrt(15.0) / 5.0;
  const double gp[3] = {-g, 0.0, g};
  const double gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  const double h = 2.0 / nqi;
  for (int iv = 0; iv < nqi; ++iv) {
    const double lo = -1.0 + iv * h;
    for (int p = 0; p < 3; ++p) {
      qp[3 * iv + p] = (2.0 * lo + h + gp[p] * h) / 2.0;
      qw[3 * iv + p] = gw[p] / nqi;
    }
  }
  for (int k = 0; k < N; ++k)
    for (int q = 0; q < Q; ++q) leg[k * Q + q] = legendre_eval(k + 1, qp[q]);
  const double d = -log(eps * eps * exp(1.0));
  const double e = 2.0;
  for (int k = 0; k < N; ++k) {
    lambda[k] = (k == 0) ? 2.0 : lambda[k - 1] + 2.0 / (k + 1);
    eigen[k] = ((d - e - lambda[k]) / 2.0) / (d - lambda[k]);
  }
  return 0;
}

/* Tables in the working precision (the reference rounds its FP64 tables to
 * REAL*4, fortran/fibers.f95:400-401,436-437). */
typedef struct {
  int Q, N;
  real qp[RFO_MAX_Q], qw[RFO_MAX_Q];
  real leg[RFO_MAX_N * RFO_MAX_Q];
  real lambda[RFO_MAX_N], eigen[RFO_MAX_N];
  real c, d, e, D1, eps2;
  real Fh[3]; /* F/2 */
} tables_t;

static int make_tables(const rfo_params *p, tables_t *t) {
  double qp[RFO_MAX_Q], qw[RFO_MAX_Q], leg[RFO_MAX_N * RFO_MAX_Q], lam[RFO_MAX_N], eig[RFO_MAX_N];
  if (rfo_tables(p->nqi, p->N, p->slenderness, qp, qw, leg, lam, eig)) return -1;
  t->Q = 3 * p->nqi;
  t->N = p->N;
  for (int q = 0; q < t->Q; ++q) {
    t->qp[q] = (real)qp[q];
    t->qw[q] = (real)qw[q];
  }
  for (int k = 0; k < t->N * t->Q; ++k) t->leg[k] = (real)leg[k];
  for (int k = 0; k < t->N; ++k) {
    t->lambda[k] = (real)lam[k];
    t->eigen[k] = (real)eig[k];
  }
  /* fortran/fibers.f95:141-145 */
  const real eps = (real)p->slenderness;
  t->c = (real)log((double)(eps * eps) * exp(1.0));
  t->d = -t->c;
  t->e = (real)2.0;
  t->D1 = (real)0.75 / (t->d - (real)2.0);
  t->eps2 = eps * eps;
  for (int a = 0; a < 3; ++a) t->Fh[a] = (real)(0.5 * p->force[a]);
  return 0;
}

static void set_threads(const rfo_params *p) {
#ifdef _OPENMP
  if (p->threads > 0) omp_set_num_threads(p->threads);
#else
  (void)p;
#endif
}

/* Stokeslet + doublet kernel of fortran/compute_inner_integral_numerically.incl:18-23:
 *   K(R) = I/r + R R^T/r^3 + 2 eps^2 (I/r^3 - 3 R R^T/r^5),
 * returned as (xx, yy, zz, xy, xz, yz). */
static inline void stokes_kernel(const real R[3], real eps2, real K[6]) {
  const real r2[3] = {R[0] * R[0], R[1] * R[1], R[2] * R[2]};
  const real inv = (real)1.0 / (real)sqrt(r2[0] + r2[1] + r2[2]);
  const real inv3 = inv * inv * inv;
  const real inv5 = inv3 * inv * inv;
  const real two_e2 = (real)2.0 * eps2;
  K[0] = inv + inv3 * r2[0] + two_e2 * (inv3 - (real)3.0 * inv5 * r2[0]);
  K[1] = inv + inv3 * r2[1] + two_e2 * (inv3 - (real)3.0 * inv5 * r2[1]);
  K[2] = inv + inv3 * r2[2] + two_e2 * (inv3 - (real)3.0 * inv5 * r2[2]);
  K[3] = inv3 * R[0] * R[1] + two_e2 * (real)(-3.0) * inv5 * R[0] * R[1];
  K[4] = inv3 * R[0] * R[2] + two_e2 * (real)(-3.0) * inv5 * R[0] * R[2];
  K[5] = inv3 * R[1] * R[2] + two_e2 * (real)(-3.0) * inv5 * R[1] * R[2];
}

void rfo_stokes_kernel(const real *R, double eps, real *K) {
  stokes_kernel(R, (real)(eps * eps), K);
}

/* Numerical inner integral, fortran/compute_inner_integral_numerically.incl:1-40:
 *   G(q_i)  = sum_{q_j} w_{q_j} K(X_i(q_i) - X_j(q_j)) P_{kj+1}(s_{q_j})
 *   GF(q_i) = sum_{q_j} w_{q_j} K(...) F/2                                   */
static void inner_numerical(const tables_t *t, const real xi[3], const real ti[3],
                            const real xj[3], const real tj[3], int kj, real *G, real *GF) {
  const int Q = t->Q;
  for (int qi = 0; qi < Q; ++qi) {
    real g[6] = {0, 0, 0, 0, 0, 0}, gf[3] = {0, 0, 0};
    real Xi[3];
    for (int a = 0; a < 3; ++a) Xi[a] = xi[a] + t->qp[qi] * ti[a];
    for (int qj = 0; qj < Q; ++qj) {
      real R[3], K[6];
      for (int a = 0; a < 3; ++a) R[a] = Xi[a] - (xj[a] + t->qp[qj] * tj[a]);
      stokes_kernel(R, t->eps2, K);
      const real w = t->qw[qj];
      const real P = t->leg[kj * Q + qj];
      for (int c = 0; c < 6; ++c) g[c] += w * K[c] * P;
      gf[0] += w * (K[0] * t->Fh[0] + K[3] * t->Fh[1] + K[4] * t->Fh[2]);
      gf[1] += w * (K[3] * t->Fh[0] + K[1] * t->Fh[1] + K[5] * t->Fh[2]);
      gf[2] += w * (K[4] * t->Fh[0] + K[5] * t->Fh[1] + K[2] * t->Fh[2]);
    }
    for (int c = 0; c < 6; ++c) G[qi * 6 + c] = g[c];
    for (int a = 0; a < 3; ++a) GF[qi * 3 + a] = gf[a];
  }
}

/* Moments I1_n, I3_n, I5_n = int_{-1}^{1} s^n (s^2 + b s + l)^{-m/2} ds, m = 1,3,5,
 * for n = 0..nmax, as fortran/compute_inner_integral_analytically.incl:12-127
 * computes them: closed forms for n = 0,1 (with the collinear branch when
 * l - b^2/4 < 1e-7), forward recurrence when l < 10, otherwise a backward
 * recurrence started from zeros at n = 27 that keeps n < N + 3.
 * nkeep is the number of force terms N (the reference's store limit); nfwd is
 * the highest index the forward branch computes (f + 2 for source mode f). */
static void moments(real b, real l, int nfwd, int nkeep, real *I1, real *I3, real *I5) {
  const real su = 1, sl = -1;
  const real m = l - (real)0.25 * b * b;
  const real uu = (real)sqrt(su * su + b * su + l);
  const real ul = (real)sqrt(sl * sl + b * sl + l);
  I1[0] = (real)log(fabs((double)((real)2.0 * su + b + (real)2.0 * uu))) -
          (real)log(fabs((double)((real)2.0 * sl + b + (real)2.0 * ul)));
  I1[1] = uu - ul + (-b / (real)2.0) * I1[0];
  const real pu = (real)2.0 * su + b, pl = (real)2.0 * sl + b;
  if (m < (real)1e-7) {
    I3[0] = ((real)-2.0 / (pu * pu)) - ((real)-2.0 / (pl * pl));
    I5[0] = ((real)-4.0 / (pu * pu * pu * pu)) - ((real)-4.0 / (pl * pl * pl * pl));
    I5[1] = ((real)-8.0 / ((real)3.0 * pu * pu * pu)) - ((real)-8.0 / ((real)3.0 * pl * pl * pl)) -
            (b / (real)2.0) * I5[0];
  } else {
    I3[0] = (pu / ((real)2.0 * m * uu)) - (pl / ((real)2.0 * m * ul));
    I5[0] = (pu / ((real)6.0 * m * uu * uu * uu)) - (pl / ((real)6.0 * m * ul * ul * ul)) +
            ((real)2.0 / ((real)3.0 * m)) * I3[0];
    I5[1] = (-(b * su + (real)2.0 * l) / ((real)6.0 * m * uu * uu * uu)) -
            (-(b * sl + (real)2.0 * l) / ((real)6.0 * m * ul * ul * ul)) - (b / ((real)3.0 * m)) * I3[0];
  }
  I3[1] = ((real)-1.0 / uu) - ((real)-1.0 / ul) - b / (real)2.0 * I3[0];
  if (l < (real)10.0) {
    real psu = su, psl = sl;
    for (int n = 2; n <= nfwd; ++n) {
      const real rn = (real)n;
      I1[n] = (psu * uu) / rn - (psl * ul) / rn + (((real)1.0 - (real)2.0 * rn) * b) / ((real)2.0 * rn) * I1[n - 1] -
              ((rn - (real)1.0) * l) / rn * I1[n - 2];
      I3[n] = I1[n - 2] - b * I3[n - 1] - l * I3[n - 2];
      I5[n] = I3[n - 2] - b * I5[n - 1] - l * I5[n - 2];
      psu *= su;
      psl *= sl;
    }
  } else {
    real i1n2 = 0, i1n1 = 0, i3n2 = 0, i3n1 = 0, i5n2 = 0, i5n1 = 0;
    const real invl = (real)1.0 / l;
    for (int n = 27; n >= 2; --n) {
      const real rn = (real)n;
      const real spu = (real)pow((double)su, n + 1), spl = (real)pow((double)sl, n + 1);
      const real i1n0 = (rn + (real)2.0) / ((rn + (real)1.0) * l) *
                        (((spu * uu) / (rn + (real)2.0)) - ((spl * ul) / (rn + (real)2.0)) +
                         (((real)1.0 - (real)2.0 * (rn + (real)2.0)) / ((real)2.0 * (rn + (real)2.0))) * b * i1n1 -
                         i1n2);
      const real i3n0 = invl * (i1n0 - b * i3n1 - i3n2);
      const real i5n0 = invl * (i3n0 - b * i5n1 - i5n2);
      if (n < nkeep + 3) {
        I1[n] = i1n0;
        I3[n] = i3n0;
        I5[n] = i5n0;
      }
      i1n2 = i1n1;
      i1n1 = i1n0;
      i3n2 = i3n1;
      i3n1 = i3n0;
      i5n2 = i5n1;
      i5n1 = i5n0;
    }
  }
}

/* Exported for the moment tests: fills I[3][nmax+1] (I1, I3, I5). */
void rfo_moments(real b, real l, int nfwd, int nkeep, real *I1, real *I3, real *I5) {
  moments(b, l, nfwd, nkeep, I1, I3, I5);
}

/* G(q) = L01 I + L03 R0R0^T - L13 (t R0^T + R0 t^T) + L23 t t^T
 *        + 2 eps^2 (L03 I - 3 L05 R0R0^T + 3 L15 (t R0^T + R0 t^T) - 3 L25 t t^T)
 * (fortran/compute_inner_integral_analytically.incl:207-260), from the moments
 * recombined with the monomial coefficients of P_f. */
static void analytic_block(const real R0[3], const real t[3], int f, real eps2, const real *I1,
                           const real *I3, const real *I5, real G[6]) {
  real L01 = 0, L03 = 0, L05 = 0, L13 = 0, L15 = 0, L23 = 0, L25 = 0;
  for (int p = 0; p <= f; ++p) {
    const real cf = (real)kLegendreCoef[f][p];
    if (cf == 0) continue;
    L01 += cf * I1[p];
    L03 += cf * I3[p];
    L05 += cf * I5[p];
    L13 += cf * I3[p + 1];
    L15 += cf * I5[p + 1];
    L23 += cf * I3[p + 2];
    L25 += cf * I5[p + 2];
  }
  static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  const real two_e2 = (real)2.0 * eps2;
  for (int c = 0; c < 6; ++c) {
    const int a = ia[c], bb = ib[c];
    const real dlt = (a == bb) ? (real)1.0 : (real)0.0;
    const real rr = R0[a] * R0[bb];
    const real tr = t[a] * R0[bb] + R0[a] * t[bb];
    const real tt = t[a] * t[bb];
    G[c] = L01 * dlt + L03 * rr - L13 * tr + L23 * tt +
           two_e2 * (L03 * dlt - (real)3.0 * L05 * rr + (real)3.0 * L15 * tr - (real)3.0 * L25 * tt);
  }
}

/* Analytical inner integral (exact in s_j), fortran/compute_inner_integral_analytically.incl. */
static void inner_analytical(const tables_t *t, const real xi[3], const real ti[3], const real xj[3],
                             const real tj[3], int kj, real *G, real *GF) {
  const int Q = t->Q, f = kj + 1;
  real I1[RFO_MAX_N + 4], I3[RFO_MAX_N + 4], I5[RFO_MAX_N + 4];
  for (int qi = 0; qi < Q; ++qi) {
    real R0[3];
    for (int a = 0; a < 3; ++a) R0[a] = xi[a] + t->qp[qi] * ti[a] - xj[a];
    const real b = (real)-2.0 * (R0[0] * tj[0] + R0[1] * tj[1] + R0[2] * tj[2]);
    const real l = R0[0] * R0[0] + R0[1] * R0[1] + R0[2] * R0[2];
    moments(b, l, f + 2, t->N, I1, I3, I5);
    analytic_block(R0, tj, f, t->eps2, I1, I3, I5, &G[qi * 6]);
    if (kj == 0) {
      real G0[6];
      analytic_block(R0, tj, 0, t->eps2, I1, I3, I5, G0);
      GF[qi * 3 + 0] = G0[0] * t->Fh[0] + G0[3] * t->Fh[1] + G0[4] * t->Fh[2];
      GF[qi * 3 + 1] = G0[3] * t->Fh[0] + G0[1] * t->Fh[1] + G0[5] * t->Fh[2];
      GF[qi * 3 + 2] = G0[4] * t->Fh[0] + G0[5] * t->Fh[1] + G0[2] * t->Fh[2];
    }
  }
}

/* System assembly, fortran/assemble_matrix.incl:1-161, for the rows of the
 * target fibers [i0, i1): A holds those 3N(i1-i0) rows (leading dimension
 * lda, all n columns, column-major), bvec the same rows of b. */
static int assemble_rows(const rfo_params *p, int analytical, const real *pos, const real *ori, int i0, int i1,
                         real *A, size_t lda, real *bvec) {
  tables_t t;
  if (make_tables(p, &t)) return -1;
  const int M = p->M, N = p->N, Q = t.Q;
  const size_t n = (size_t)3 * N * M;
  const size_t roff = (size_t)3 * N * i0, nrows = (size_t)3 * N * (i1 - i0);
  if (i0 < 0 || i1 > M || i0 > i1 || lda < nrows) return -1;
  set_threads(p);
  for (size_t c = 0; c < n; ++c) memset(A + c * lda, 0, nrows * sizeof(real));
  memset(bvec, 0, nrows * sizeof(real));
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = i0; i < i1; ++i) {
    real G[RFO_MAX_Q * 6], GF[RFO_MAX_Q * 3];
    const real *xi = &pos[3 * i], *ti = &ori[3 * i];
    for (int j = 0; j < M; ++j) {
      if (i == j) {
        for (size_t r = (size_t)3 * N * i; r < (size_t)3 * N * (i + 1); ++r) A[(r - roff) + r * lda] = 1;
        continue;
      }
      const real *xj = &pos[3 * j], *tj = &ori[3 * j];
      for (int kj = 0; kj < N; ++kj) {
        if (analytical)
          inner_analytical(&t, xi, ti, xj, tj, kj, G, GF);
        else
          inner_numerical(&t, xi, ti, xj, tj, kj, G, GF);
        for (int ki = 0; ki < N; ++ki) {
          real T[6] = {0, 0, 0, 0, 0, 0}, TF[3] = {0, 0, 0};
          for (int qi = 0; qi < Q; ++qi) {
            const real wP = t.qw[qi] * t.leg[ki * Q + qi];
            for (int c = 0; c < 6; ++c) T[c] += wP * G[qi * 6 + c];
            if (kj == 0)
              for (int a = 0; a < 3; ++a) TF[a] += wP * GF[qi * 3 + a];
          }
          /* Q = Theta t_i (Theta symmetric: xx yy zz xy xz yz) */
          const real Qv[3] = {T[0] * ti[0] + T[3] * ti[1] + T[4] * ti[2],
                              T[3] * ti[0] + T[1] * ti[1] + T[5] * ti[2],
                              T[4] * ti[0] + T[5] * ti[1] + T[2] * ti[2]};
          const real Tm[3][3] = {{T[0], T[3], T[4]}, {T[3], T[1], T[5]}, {T[4], T[5], T[2]}};
          const size_t row0 = (size_t)3 * N * i + 3 * ki - roff;
          const size_t col0 = (size_t)3 * N * j + 3 * kj;
          if (ki == 0) {
            for (int a = 0; a < 3; ++a)
              for (int bb = 0; bb < 3; ++bb) A[(row0 + a) + (col0 + bb) * lda] = t.D1 * ti[a] * Qv[bb];
          } else {
            const real gamma = (real)0.5 * ((real)2.0 * (ki + 1) + (real)1.0) / (t.d + t.e - t.lambda[ki]);
            for (int a = 0; a < 3; ++a)
              for (int bb = 0; bb < 3; ++bb)
                A[(row0 + a) + (col0 + bb) * lda] = gamma * (Tm[a][bb] - t.eigen[ki] * ti[a] * Qv[bb]);
          }
          if (kj == 0) {
            const real QF = TF[0] * ti[0] + TF[1] * ti[1] + TF[2] * ti[2];
            if (ki == 0) {
              for (int a = 0; a < 3; ++a) bvec[row0 + a] -= t.D1 * ti[a] * QF;
            } else {
              const real gamma = (real)0.5 * ((real)2.0 * (ki + 1) + (real)1.0) / (t.d + t.e - t.lambda[ki]);
              for (int a = 0; a < 3; ++a) bvec[row0 + a] -= gamma * (TF[a] - t.eigen[ki] * ti[a] * QF);
            }
          }
        }
      }
    }
  }
  return 0;
}

int rfo_assemble(const rfo_params *p, int analytical, const real *pos, const real *ori, real *A, real *bvec) {
  const size_t n = (size_t)3 * p->N * p->M;
  return assemble_rows(p, analytical, pos, ori, 0, p->M, A, n, bvec);
}

/* Rows of the target fibers [i0, i1) only (a bounded sample for timing). */
int rfo_assemble_rows(const rfo_params *p, int analytical, const real *pos, const real *ori, int i0, int i1,
                      real *A, real *bvec) {
  return assemble_rows(p, analytical, pos, ori, i0, i1, A, (size_t)3 * p->N * (i1 - i0), bvec);
}

/* y = A_rows x for a block of rows stored column-major (nrows x n), swept
 * column by column as sgemv('N') does (fortran/solve_system.incl:36). */
void rfo_matvec_rows(int nrows, int n_, const real *A, const real *x, real *y) {
  const size_t n = (size_t)n_, nr = (size_t)nrows;
  for (size_t r = 0; r < nr; ++r) y[r] = 0;
#pragma omp parallel
  {
#ifdef _OPENMP
    const int nt = omp_get_num_threads(), id = omp_get_thread_num();
#else
    const int nt = 1, id = 0;
#endif
    const size_t lo = nr * id / nt, hi = nr * (id + 1) / nt;
    for (size_t c = 0; c < n; ++c) {
      const real xc = x[c];
      const real *col = A + c * nr;
      for (size_t r = lo; r < hi; ++r) y[r] += col[r] * xc;
    }
  }
}

/* Dense LU with partial pivoting and the two triangular solves (the semantics
 * of LAPACK sgesv called at fortran/solve_system.incl:3).  A (column-major) is
 * overwritten by its factors and b by the solution.  Returns 0, or k+1 if the
 * k-th pivot is exactly zero (LAPACK's INFO). */
int rfo_solve_direct(int n_, real *A, real *b) {
  const size_t n = (size_t)n_;
  for (size_t k = 0; k < n; ++k) {
    size_t piv = k;
    real best = (real)fabs((double)A[k + k * n]);
    for (size_t r = k + 1; r < n; ++r) {
      const real v = (real)fabs((double)A[r + k * n]);
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    if (best == 0) return (int)k + 1;
    if (piv != k) {
      for (size_t c = 0; c < n; ++c) {
        const real tmp = A[k + c * n];
        A[k + c * n] = A[piv + c * n];
        A[piv + c * n] = tmp;
      }
      const real tb = b[k];
      b[k] = b[piv];
      b[piv] = tb;
    }
    const real inv = (real)1.0 / A[k + k * n];
    for (size_t r = k + 1; r < n; ++r) A[r + k * n] *= inv;
#pragma omp parallel for schedule(static) if (n - k > 256)
    for (size_t c = k + 1; c < n; ++c) {
      const real akc = A[k + c * n];
      if (akc != 0)
        for (size_t r = k + 1; r < n; ++r) A[r + c * n] -= A[r + k * n] * akc;
    }
  }
  for (size_t k = 0; k < n; ++k) /* L y = P b (unit lower) */
    for (size_t r = k + 1; r < n; ++r) b[r] -= A[r + k * n] * b[k];
  for (size_t kk = n; kk-- > 0;) { /* U x = y */
    b[kk] /= A[kk + kk * n];
    for (size_t r = 0; r < kk; ++r) b[r] -= A[r + kk * n] * b[kk];
  }
  return 0;
}

static void matvec(size_t n, const real *A, const real *x, real *y) {
#pragma omp parallel for schedule(static) if (n > 512)
  for (size_t r = 0; r < n; ++r) y[r] = 0;
  /* column sweep, as sgemv('N') on a column-major matrix */
#pragma omp parallel
  {
#ifdef _OPENMP
    const int nt = omp_get_num_threads(), id = omp_get_thread_num();
#else
    const int nt = 1, id = 0;
#endif
    const size_t lo = n * id / nt, hi = n * (id + 1) / nt;
    for (size_t c = 0; c < n; ++c) {
      const real xc = x[c];
      const real *col = A + c * n;
      for (size_t r = lo; r < hi; ++r) y[r] += col[r] * xc;
    }
  }
}

static double dot(size_t n, const real *a, const real *b) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
  return s;
}

/* Restarted GMRES(m) with modified Gram-Schmidt, no preconditioner, the true
 * residual recomputed at each restart, stopping when the backward error
 * ||b - A x|| / ||b|| <= tol (fortran/solve_system.incl:18-57 with
 * icntl(5)=0 MGS, icntl(8)=1, sPA=0; fortran/sgmres.f:1240-1358).
 * use_x0 = 0 starts from x = 0 as the shipped Fortran does
 * (fortran/solve_system.incl:16-17); use_x0 = 1 starts from the x passed in
 * (the previous step's solution reproduces the golden *_gmres vectors).
 * Returns 0 on convergence, -4 when maxit is exhausted (sgmres info(1)). */
int rfo_solve_gmres(int n_, const real *A, const real *b, real *x, int use_x0, int m, int maxit, double tol,
                    int *iters, double *relres) {
  const size_t n = (size_t)n_;
  real *V = (real *)malloc(sizeof(real) * n * (size_t)(m + 1));
  real *w = (real *)malloc(sizeof(real) * n);
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = (double *)malloc(sizeof(double) * m), *sn = (double *)malloc(sizeof(double) * m);
  double *g = (double *)malloc(sizeof(double) * (m + 1)), *y = (double *)malloc(sizeof(double) * m);
  int it = 0, status = -4;
  const double bnorm = sqrt(dot(n, b, b));
  if (!use_x0)
    for (size_t i = 0; i < n; ++i) x[i] = 0;
  double res = bnorm;
  if (bnorm == 0.0) {
    status = 0;
    res = 0.0;
    goto done;
  }
  while (it < maxit) {
    /* r = b - A x (true residual at every (re)start) */
    matvec(n, A, x, w);
    for (size_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
    const double beta = sqrt(dot(n, w, w));
    res = beta;
    if (beta / bnorm <= tol) {
      status = 0;
      break;
    }
    for (size_t i = 0; i < n; ++i) V[i] = (real)(w[i] / beta);
    for (int k = 0; k <= m; ++k) g[k] = 0.0;
    g[0] = beta;
    int k_used = 0;
    for (int k = 0; k < m && it < maxit; ++k) {
      ++it;
      matvec(n, A, V + (size_t)k * n, w);
      for (int i = 0; i <= k; ++i) { /* modified Gram-Schmidt */
        const double h = dot(n, w, V + (size_t)i * n);
        H[i + (size_t)k * (m + 1)] = h;
        const real *vi = V + (size_t)i * n;
        for (size_t r = 0; r < n; ++r) w[r] -= (real)h * vi[r];
      }
      const double hn = sqrt(dot(n, w, w));
      H[(k + 1) + (size_t)k * (m + 1)] = hn;
      if (hn > 0)
        for (size_t r = 0; r < n; ++r) V[(size_t)(k + 1) * n + r] = (real)(w[r] / hn);
      for (int i = 0; i < k; ++i) { /* apply previous Givens rotations */
        const double h0 = H[i + (size_t)k * (m + 1)], h1 = H[i + 1 + (size_t)k * (m + 1)];
        H[i + (size_t)k * (m + 1)] = cs[i] * h0 + sn[i] * h1;
        H[i + 1 + (size_t)k * (m + 1)] = -sn[i] * h0 + cs[i] * h1;
      }
      const double h0 = H[k + (size_t)k * (m + 1)], h1 = H[k + 1 + (size_t)k * (m + 1)];
      const double rr = sqrt(h0 * h0 + h1 * h1);
      cs[k] = (rr == 0) ? 1.0 : h0 / rr;
      sn[k] = (rr == 0) ? 0.0 : h1 / rr;
      H[k + (size_t)k * (m + 1)] = rr;
      H[k + 1 + (size_t)k * (m + 1)] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      k_used = k + 1;
      res = fabs(g[k + 1]);
      if (res / bnorm <= tol || hn == 0) break;
    }
    for (int i = k_used - 1; i >= 0; --i) { /* back substitution */
      double s = g[i];
      for (int j = i + 1; j < k_used; ++j) s -= H[i + (size_t)j * (m + 1)] * y[j];
      y[i] = s / H[i + (size_t)i * (m + 1)];
    }
    for (int i = 0; i < k_used; ++i) {
      const real *vi = V + (size_t)i * n;
      for (size_t r = 0; r < n; ++r) x[r] += (real)y[i] * vi[r];
    }
    if (res / bnorm <= tol) {
      /* confirm with the true residual, as sgmres does with icntl(8)=1 */
      matvec(n, A, x, w);
      for (size_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
      res = sqrt(dot(n, w, w));
      if (res / bnorm <= tol) {
        status = 0;
        break;
      }
    }
  }
done:
  if (iters) *iters = it;
  if (relres) *relres = (bnorm > 0) ? res / bnorm : 0.0;
  free(V);
  free(w);
  free(H);
  free(cs);
  free(sn);
  free(g);
  free(y);
  return status;
}

/* Velocities, fortran/update_velocities.incl:1-89:
 *   v_i = (0.5/d) [ (d+2) F + (d-2)(t_i.F) t_i + sum_{j!=i} TFA0_ij ]
 *   w_i = (1.5/d) sum_{j!=i} (I - t_i t_i^T) TFA1_ij
 * with GF(q_i) = sum_{q_j} w K f_j(q_j), f_j(s) = F/2 + sum_k x_{j,k} P_{k+1}(s),
 * TFA0 = sum_{q_i} w GF(q_i), TFA1 = sum_{q_i} w P_1(s_{q_i}) GF(q_i). */
/* Exact inner integral of the velocity kernel: GF(q_i) = int_{-1}^{1} K(R0 - s t_j) f_j(s) ds with
 * f_j(s) = sum_p C_p s^p the monomial form of F/2 + sum_k x_{j,k} P_{k+1}(s), from the moments
 * (same recurrences as the assembly).  The reference's golden 100_analytic_* velocity vectors are
 * computed this way; the shipped fortran/update_velocities.incl always uses the quadrature. */
static void velocity_inner_analytical(const tables_t *t, const real R0[3], const real tj[3], const real C[][3],
                                      real gf[3]) {
  const int N = t->N;
  real I1[RFO_MAX_N + 4], I3[RFO_MAX_N + 4], I5[RFO_MAX_N + 4];
  const real b = (real)-2.0 * (R0[0] * tj[0] + R0[1] * tj[1] + R0[2] * tj[2]);
  const real l = R0[0] * R0[0] + R0[1] * R0[1] + R0[2] * R0[2];
  moments(b, l, N + 2, N, I1, I3, I5);
  const real two_e2 = (real)2.0 * t->eps2;
  gf[0] = gf[1] = gf[2] = 0;
  for (int p = 0; p <= N; ++p) {
    const real *c = C[p];
    const real rc = R0[0] * c[0] + R0[1] * c[1] + R0[2] * c[2];
    const real tc = tj[0] * c[0] + tj[1] * c[1] + tj[2] * c[2];
    for (int a = 0; a < 3; ++a) {
      const real Kc = I1[p] * c[a] + I3[p] * R0[a] * rc - I3[p + 1] * (tj[a] * rc + R0[a] * tc) +
                      I3[p + 2] * tj[a] * tc +
                      two_e2 * (I3[p] * c[a] - (real)3.0 * I5[p] * R0[a] * rc +
                                (real)3.0 * I5[p + 1] * (tj[a] * rc + R0[a] * tc) - (real)3.0 * I5[p + 2] * tj[a] * tc);
      gf[a] += Kc;
    }
  }
}

/* Velocities of the target fibers [i0, i1) (v, w hold 3(i1-i0) values). */
int rfo_velocities_rows(const rfo_params *p, int analytical, const real *pos, const real *ori, const real *x,
                        int i0, int i1, real *v, real *w) {
  tables_t t;
  if (make_tables(p, &t)) return -1;
  const int M = p->M, N = p->N, Q = t.Q;
  if (i0 < 0 || i1 > M || i0 > i1) return -1;
  set_threads(p);
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = i0; i < i1; ++i) {
    const real *xi = &pos[3 * i], *ti = &ori[3 * i];
    real F2[3], vi[3], wi[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) F2[a] = (real)2.0 * t.Fh[a];
    const real tF = ti[0] * F2[0] + ti[1] * F2[1] + ti[2] * F2[2];
    for (int a = 0; a < 3; ++a) vi[a] = (t.d + (real)2.0) * F2[a] + (t.d - (real)2.0) * tF * ti[a];
    for (int j = 0; j < M; ++j) {
      if (j == i) continue;
      const real *xj = &pos[3 * j], *tj = &ori[3 * j];
      real A0[3] = {0, 0, 0}, A1[3] = {0, 0, 0};
      real C[RFO_MAX_N + 1][3];
      if (analytical) {
        for (int q = 0; q <= N; ++q)
          for (int a = 0; a < 3; ++a) {
            real s = (real)kLegendreCoef[0][q] * t.Fh[a];
            for (int k = 0; k < N; ++k) s += (real)kLegendreCoef[k + 1][q] * x[(size_t)3 * N * j + 3 * k + a];
            C[q][a] = s;
          }
      }
      for (int qi = 0; qi < Q; ++qi) {
        real Xi[3], gf[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) Xi[a] = xi[a] + t.qp[qi] * ti[a];
        if (analytical) {
          const real R0[3] = {Xi[0] - xj[0], Xi[1] - xj[1], Xi[2] - xj[2]};
          velocity_inner_analytical(&t, R0, tj, (const real(*)[3])C, gf);
        }
        for (int qj = 0; qj < Q && !analytical; ++qj) {
          real R[3], K[6], f[3];
          for (int a = 0; a < 3; ++a) R[a] = Xi[a] - (xj[a] + t.qp[qj] * tj[a]);
          stokes_kernel(R, t.eps2, K);
          for (int a = 0; a < 3; ++a) f[a] = t.Fh[a];
          for (int k = 0; k < N; ++k) {
            const real P = t.leg[k * Q + qj];
            for (int a = 0; a < 3; ++a) f[a] += x[(size_t)3 * N * j + 3 * k + a] * P;
          }
          const real wq = t.qw[qj];
          gf[0] += wq * (K[0] * f[0] + K[3] * f[1] + K[4] * f[2]);
          gf[1] += wq * (K[3] * f[0] + K[1] * f[1] + K[5] * f[2]);
          gf[2] += wq * (K[4] * f[0] + K[5] * f[1] + K[2] * f[2]);
        }
        for (int a = 0; a < 3; ++a) {
          A0[a] += t.qw[qi] * gf[a];
          A1[a] += t.qw[qi] * gf[a] * t.leg[qi]; /* P_1 */
        }
      }
      for (int a = 0; a < 3; ++a) vi[a] += A0[a];
      const real tA1 = ti[0] * A1[0] + ti[1] * A1[1] + ti[2] * A1[2];
      for (int a = 0; a < 3; ++a) wi[a] += A1[a] - ti[a] * tA1;
    }
    for (int a = 0; a < 3; ++a) {
      v[3 * (i - i0) + a] = vi[a] * ((real)0.5 / t.d);
      w[3 * (i - i0) + a] = wi[a] * ((real)1.5 / t.d);
    }
  }
  return 0;
}

int rfo_velocities(const rfo_params *p, int analytical, const real *pos, const real *ori, const real *x, real *v,
                   real *w) {
  return rfo_velocities_rows(p, analytical, pos, ori, x, 0, p->M, v, w);
}

static void normalize3(real *o) {
  const real len = (real)sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
  for (int a = 0; a < 3; ++a) o[a] /= len;
}

/* Forward Euler first step, fortran/update_fibers_firststep.incl:1-10. */
void rfo_update_first(const rfo_params *p, const real *pos, const real *ori, const real *v, const real *w,
                      real *pos_next, real *ori_next) {
  const real dt = (real)p->dt;
  for (int i = 0; i < p->M; ++i) {
    for (int a = 0; a < 3; ++a) {
      pos_next[3 * i + a] = pos[3 * i + a] + dt * v[3 * i + a];
      ori_next[3 * i + a] = ori[3 * i + a] + dt * w[3 * i + a];
    }
    normalize3(&ori_next[3 * i]);
  }
}

/* Second-order multistep update, fortran/update_fibers.incl:1-17:
 *   y+ = 4/3 y - 1/3 y- + 2/3 dt (2 y' - y-'),  then t+ normalised. */
void rfo_update(const rfo_params *p, const real *pos_prev, const real *pos, const real *ori_prev, const real *ori,
                const real *v_prev, const real *v, const real *w_prev, const real *w, real *pos_next,
                real *ori_next) {
  const real dt = (real)p->dt;
  const real c43 = (real)4.0 / (real)3.0, c13 = (real)1.0 / (real)3.0, c23 = (real)2.0 / (real)3.0;
  for (int i = 0; i < p->M; ++i) {
    for (int a = 0; a < 3; ++a) {
      const int k = 3 * i + a;
      pos_next[k] = c43 * pos[k] - c13 * pos_prev[k] + c23 * dt * ((real)2.0 * v[k] - v_prev[k]);
      ori_next[k] = c43 * ori[k] - c13 * ori_prev[k] + c23 * dt * ((real)2.0 * w[k] - w_prev[k]);
    }
    normalize3(&ori_next[3 * i]);
  }
}

int rfo_real_size(void) { return (int)sizeof(real); }
