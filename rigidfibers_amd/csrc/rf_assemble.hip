// rf_assemble.hip -- assembly of the dense fiber-fiber system A x = b.
//
// Reference: fortran/assemble_matrix.incl:1-161 (block formulas),
// fortran/compute_inner_integral_numerically.incl:1-40 and
// fortran/compute_inner_integral_analytically.incl:1-338 (inner integrals),
// src/kernels/assemble_system_2D.cu:8-367 (the GPU layout of the same system).
//
// Work decomposition.  One workgroup = PAIRS consecutive target fibers i times
// Q target nodes q_i (192 lanes), looping over a group of source fibers j.
//   phase 1  lane (i, q_i) forms G_m(q_i) = sum_{q_j} w P_m(s_{q_j}) K(X_i(q_i) - X_j(q_j))
//            for all modes m = 0..N at once (the reference recomputes K for each
//            source mode), pairing q_j with its mirror -s_{q_j} so even modes use
//            K(q) + K(-q) and odd modes K(q) - K(-q).  Analytical mode forms the
//            same G_m exactly from FP64 moments, once per (i, j, q_i) for all m.
//   phase 2  jobs (i, m, column b) read G_m(.)[(a,b)] from LDS, project over q_i
//            (Theta_{k_i} = sum_{q_i} w P_{k_i+1}(s_{q_i}) G) and write the
//            15-row column segment of A; jobs (i, m = 0) accumulate the RHS.
// A is column-major with leading dimension lda; the workgroup's 8 target
// fibers own 8*15 consecutive rows, so each column segment it writes is
// contiguous.  Diagonal blocks are written as the identity, so A needs no
// memset.  RHS partial sums go to bpart[group][row] (FP64), summed afterwards
// in a fixed order: no atomics, bitwise reproducible.
#include "rf_internal.h"

namespace rf {

constexpr int kPairThreads = 192;

// Modes from the highest down: mode m is the last reader of the moments
// I3_{m+2}, I5_{m+2}, so they die as the sweep goes.
template <int N, int NQI, int M_>
__device__ __forceinline__ void analytic_modes(const NodeProducts& q, double eps2, const double* I1, const double* I3,
                                               const double* I5, float acc[N + 1][6]) {
  if constexpr (M_ >= 0) {
    double G[6];
    analytic_g<M_>(q, eps2, I1, I3, I5, G);
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[M_][c] = static_cast<float>(G[c]);
    analytic_modes<N, NQI, M_ - 1>(q, eps2, I1, I3, I5, acc);
  }
}

// G_m(q_i) for m = 0..N of one ordered pair (i, j) at the lane's target node:
// quadrature over q_j (mirror-paired nodes) or the exact FP64 integral.
template <int N, int NQI, bool ANALYTIC>
__device__ __forceinline__ void inner_integrals(const double xi[3], const double tid_[3], const float ti[3], double sq_d,
                                                float sq, double xj0, double xj1, double xj2, double tj0, double tj1,
                                                double tj2, const PhysConsts& pc, float acc[N + 1][6]) {
  constexpr Quad<N, NQI> quad{};
  constexpr int Q = 3 * NQI;
  constexpr int NM = N + 1;
#pragma unroll
  for (int m = 0; m < NM; ++m)
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[m][c] = 0.0f;
  if constexpr (!ANALYTIC) {
    // R(q_i, q_j) = (x_i - x_j) + s_i t_i - s_j t_j; the centre difference is
    // formed in FP64 and rounded once.
    const float ax = fmaf(sq, ti[0], static_cast<float>(xi[0] - xj0));
    const float ay = fmaf(sq, ti[1], static_cast<float>(xi[1] - xj1));
    const float az = fmaf(sq, ti[2], static_cast<float>(xi[2] - xj2));
    const float tx = static_cast<float>(tj0), ty = static_cast<float>(tj1), tz = static_cast<float>(tj2);
#pragma unroll
    for (int h = 0; h < Q / 2; ++h) {
      const float s = quad.s[h];
      const float ux = s * tx, uy = s * ty, uz = s * tz;
      float Kp[6], Km[6];
      stokes_kernel(ax - ux, ay - uy, az - uz, pc.eps2, Kp);  // node h (s)
      stokes_kernel(ax + ux, ay + uy, az + uz, pc.eps2, Km);  // node Q-1-h (-s)
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const float E = Kp[c] + Km[c];
        const float O = Kp[c] - Km[c];
#pragma unroll
        for (int m = 0; m < NM; ++m) acc[m][c] = fmaf(quad.W[m][h], (m & 1) ? O : E, acc[m][c]);
      }
    }
  } else {
    const double R0[3] = {xi[0] - xj0 + sq_d * tid_[0], xi[1] - xj1 + sq_d * tid_[1], xi[2] - xj2 + sq_d * tid_[2]};
    const double ll = R0[0] * R0[0] + R0[1] * R0[1] + R0[2] * R0[2];
    if (ll >= kFarField[N] * kFarField[N]) {
      // far field: 8-point Gauss-Legendre over the source fiber (FP32)
      constexpr FarQuad<N> far{};
      const float rx = static_cast<float>(R0[0]), ry = static_cast<float>(R0[1]), rz = static_cast<float>(R0[2]);
      const float tx = static_cast<float>(tj0), ty = static_cast<float>(tj1), tz = static_cast<float>(tj2);
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const float s = far.s[h];
        const float ux = s * tx, uy = s * ty, uz = s * tz;
        float Kp[6], Km[6];
        stokes_kernel(rx - ux, ry - uy, rz - uz, pc.eps2, Kp);  // node +s
        stokes_kernel(rx + ux, ry + uy, rz + uz, pc.eps2, Km);  // node -s
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          const float E = Kp[c] + Km[c];
          const float O = Kp[c] - Km[c];
#pragma unroll
          for (int m = 0; m < NM; ++m) acc[m][c] = fmaf(far.W[m][h], (m & 1) ? O : E, acc[m][c]);
        }
      }
    } else {
      const double tjd[3] = {tj0, tj1, tj2};
      const double bb = -2.0 * (R0[0] * tj0 + R0[1] * tj1 + R0[2] * tj2);
      double I1[N + 3], I3[N + 3], I5[N + 3];
      moments<N>(bb, ll, I1, I3, I5);
      analytic_modes<N, NQI, N>(node_products(R0, tjd), pc.eps2_d, I1, I3, I5, acc);
    }
  }
}

template <int N, int NQI, bool ANALYTIC>
__global__ __launch_bounds__(kPairThreads) void assemble_kernel(int M, int64_t lda, const double* __restrict__ pos,
                                                                const double* __restrict__ ori,
                                                                float* __restrict__ A, double* __restrict__ bpart,
                                                                int jgroup, PhysConsts pc) {
  constexpr Quad<N, NQI> quad{};
  constexpr int Q = 3 * NQI;
  constexpr int PAIRS = kPairThreads / Q;
  constexpr int NM = N + 1;
  constexpr int ROWS = 3 * N;
  static_assert(kPairThreads % Q == 0, "Q must divide the workgroup");
  static_assert(Q % 2 == 0, "mirror pairing needs an even node count");

  __shared__ float lds[PAIRS][NM][6][Q];
  __shared__ float ti_lds[PAIRS][3];

  const int tid = threadIdx.x;
  const int p = tid / Q;
  const int qi = tid - p * Q;
  const int i = blockIdx.x * PAIRS + p;
  const bool ivalid = i < M;
  const int j0 = blockIdx.y * jgroup;
  const int j1 = min(M, j0 + jgroup);
  const int64_t n = static_cast<int64_t>(ROWS) * M;

  double xi[3] = {0, 0, 0}, tid_[3] = {0, 0, 0};
  if (ivalid) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      xi[a] = pos[3 * i + a];
      tid_[a] = ori[3 * i + a];
    }
  }
  if (qi == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) ti_lds[p][a] = static_cast<float>(tid_[a]);
  }
  const double sq_d = node_fp64<NQI>(qi);
  const float sq = static_cast<float>(sq_d);
  const float ti[3] = {static_cast<float>(tid_[0]), static_cast<float>(tid_[1]), static_cast<float>(tid_[2])};

  // RHS accumulators of the jobs (i, m = 0); thread tid < PAIRS owns pair tid.
  double bacc[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) bacc[r] = 0.0;

  constexpr int kJobs = PAIRS * (3 * N + 1);

  for (int j = j0; j < j1; ++j) {
    // ---------------- phase 1: inner integrals G_m(q_i) ----------------
    if (ivalid && i != j) {
      const double xj0 = pos[3 * j + 0], xj1 = pos[3 * j + 1], xj2 = pos[3 * j + 2];
      const double tj0 = ori[3 * j + 0], tj1 = ori[3 * j + 1], tj2 = ori[3 * j + 2];
      float acc[NM][6];
      inner_integrals<N, NQI, ANALYTIC>(xi, tid_, ti, sq_d, sq, xj0, xj1, xj2, tj0, tj1, tj2, pc, acc);
#pragma unroll
      for (int m = 0; m < NM; ++m)
#pragma unroll
        for (int c = 0; c < 6; ++c) lds[p][m][c][qi] = acc[m][c];
    }
    __syncthreads();

    // ---------------- phase 2: projections and stores ----------------
    for (int job = tid; job < kJobs; job += kPairThreads) {
      int p2, m, bcol;
      if (job < PAIRS) {
        p2 = job;
        m = 0;
        bcol = 0;
      } else {
        const int jj = job - PAIRS;
        p2 = jj % PAIRS;
        const int rest = jj / PAIRS;
        m = 1 + rest / 3;
        bcol = rest - 3 * (rest / 3);
      }
      const int i2 = blockIdx.x * PAIRS + p2;
      if (i2 >= M) continue;
      const float t0 = ti_lds[p2][0], t1 = ti_lds[p2][1], t2 = ti_lds[p2][2];
      if (m == 0) {
        if (i2 == j) continue;
        // TF_{k_i} = sum_{q_i} w P_{k_i+1} (G_0(q_i) F/2)
        float TF[N][3];
#pragma unroll
        for (int k = 0; k < N; ++k) TF[k][0] = TF[k][1] = TF[k][2] = 0.0f;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const float g0 = lds[p2][0][0][q], g1 = lds[p2][0][1][q], g2 = lds[p2][0][2][q];
          const float g3 = lds[p2][0][3][q], g4 = lds[p2][0][4][q], g5 = lds[p2][0][5][q];
          const float f0 = g0 * pc.Fh[0] + g3 * pc.Fh[1] + g4 * pc.Fh[2];
          const float f1 = g3 * pc.Fh[0] + g1 * pc.Fh[1] + g5 * pc.Fh[2];
          const float f2 = g4 * pc.Fh[0] + g5 * pc.Fh[1] + g2 * pc.Fh[2];
#pragma unroll
          for (int k = 0; k < N; ++k) {
            TF[k][0] = fmaf(quad.W[k + 1][q], f0, TF[k][0]);
            TF[k][1] = fmaf(quad.W[k + 1][q], f1, TF[k][1]);
            TF[k][2] = fmaf(quad.W[k + 1][q], f2, TF[k][2]);
          }
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const float QF = TF[k][0] * t0 + TF[k][1] * t1 + TF[k][2] * t2;
          const float tt[3] = {t0, t1, t2};
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const float v = (k == 0) ? pc.D1 * tt[a] * QF : pc.gamma[k] * (TF[k][a] - pc.eigen[k] * tt[a] * QF);
            bacc[3 * k + a] -= static_cast<double>(v);
          }
        }
        continue;
      }
      const int kj = m - 1;
      const int64_t col = static_cast<int64_t>(ROWS) * j + 3 * kj + bcol;
      float* out = A + col * lda + static_cast<int64_t>(ROWS) * i2;
      if (i2 == j) {
#pragma unroll
        for (int r = 0; r < ROWS; ++r) out[r] = (r == 3 * kj + bcol) ? 1.0f : 0.0f;
        continue;
      }
      // Theta_{k_i}[(a, bcol)] for a = 0..2
      const int c0 = sym_index(0, bcol), c1 = sym_index(1, bcol), c2 = sym_index(2, bcol);
      float T[N][3];
#pragma unroll
      for (int k = 0; k < N; ++k) T[k][0] = T[k][1] = T[k][2] = 0.0f;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const float g0 = lds[p2][m][c0][q], g1 = lds[p2][m][c1][q], g2 = lds[p2][m][c2][q];
#pragma unroll
        for (int k = 0; k < N; ++k) {
          T[k][0] = fmaf(quad.W[k + 1][q], g0, T[k][0]);
          T[k][1] = fmaf(quad.W[k + 1][q], g1, T[k][1]);
          T[k][2] = fmaf(quad.W[k + 1][q], g2, T[k][2]);
        }
      }
      const float tt[3] = {t0, t1, t2};
#pragma unroll
      for (int k = 0; k < N; ++k) {
        // (Theta t_i)_bcol, Theta symmetric
        const float Qb = T[k][0] * t0 + T[k][1] * t1 + T[k][2] * t2;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float v = (k == 0) ? pc.D1 * tt[a] * Qb : pc.gamma[k] * (T[k][a] - pc.eigen[k] * tt[a] * Qb);
          out[3 * k + a] = v;
        }
      }
    }
    __syncthreads();
  }

  if (tid < PAIRS) {
    const int i2 = blockIdx.x * PAIRS + tid;
    if (i2 < M) {
      double* dst = bpart + static_cast<int64_t>(blockIdx.y) * n + static_cast<int64_t>(ROWS) * i2;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) dst[r] = bacc[r];
    }
  }
}

// ---------------------------------------------------------------------------
// Theta form of the system, for the iterative solvers.
//
// Every off-diagonal 3x3 block of A is fixed by t_i and the symmetric
// projection Theta_{k_i,k_j} (fortran/assemble_matrix.incl:59-79,117-137):
//   k_i = 0:  A_blk = D1 t_i Q0^T,  Q0 = Theta_{0,k_j} t_i            (3 numbers)
//   k_i > 0:  A_blk = gamma_k (Theta - eigen_k t_i (Theta t_i)^T)     (6 numbers)
// so 3 + 6(N-1) numbers per (i, j, k_j) carry the 9N entries of a block
// column (27 instead of 45 for N = 5).  They are stored as component arrays
//   T[(i * NC + c) * Lp + j * N + k_j],  NC = 3 + 6(N-1),  Lp >= M*N,
// c = b for Q0_b, c = 3 + 6(k_i-1) + s for Theta_{k_i}[s] (s = xx yy zz xy xz yz),
// zero for j == i (the identity diagonal blocks are implicit).
//
// Workgroup = one target fiber i x 8 consecutive source fibers per iteration
// (lane = (source, q_i)); phase 2 jobs (column b, source, mode) are ordered so
// that 40 consecutive lanes write 40 consecutive floats of each component.

// Velocity operator (when the velocities integrate like the system): the
// reference's velocity sums (fortran/update_velocities.incl:29-86) are
// projections of the same inner integrals G_m(q_i) the assembly forms,
//   sum_j TFA0_ij = sum_j [Phi0_{ij,0} F/2 + sum_k Phi0_{ij,k+1} x_{j,k}]   (weights w)
//   sum_j TFA1_ij = sum_j [Theta_{ij,0,P0} F/2 + sum_k Theta_{ij,0,k} x_{j,k}] (weights w P_1)
// so with Tv != nullptr the kernel also stores, per (i, j, k_j), the
// symmetric Theta_{0,k_j} and Phi0_{k_j+1} = sum_{q_i} w G_{k_j+1}(q_i)
//   Tv[(i * 12 + c) * Lp + j * N + k_j],  c = s (Theta_0), 6 + s (Phi0),
// and in cpart[group][i][6] the constant sums of Phi0_{ij,0} F/2 and
// Theta_{ij,0,P0} F/2 over its sources.

// Column b of a symmetric 3x3 projection (v_a = S[(a, b)]) into its component
// arrays s = xx yy zz xy xz yz at stride Lp: b = 0 writes xx xy xz, b = 1 yy yz,
// b = 2 zz, so the three column jobs together write each component once.
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// Scale of one (target, source) block of the FP16 copy from the bits of g,
// the largest |projection| of the block's rows (modes m >= 1, projections
// k >= 1): the first three components dot three of them with the unit t_i,
// the others are projections themselves, so |T| <= sqrt(3) g < 2 g.  With g
// = 1.f 2^(E - 127), sv = 2^(E - 140) > 2 g / 32768, so the stored values
// T / sv stay below 32768 in magnitude (and dividing by a power of two is
// exact); iv = 1 / sv.  An empty (or vanishing) block gets 0, 0; a
// non-finite g is passed on as the scale.
__device__ __forceinline__ void half_block_scale(unsigned gbits, float& sv, float& iv) {
  const unsigned E = (gbits >> 23) & 0xffu;
  if (E >= 14u && E < 255u) {
    sv = __uint_as_float((E - 13u) << 23);
    iv = __uint_as_float((267u - E) << 23);
  } else {
    sv = (E == 255u) ? __uint_as_float(gbits) : 0.0f;
    iv = 0.0f;
  }
}

// The work of one workgroup: target fiber i, source group gy.
//   phase 1  lane (p, q_i) writes G_m(q_i) of source slot p into LDS as row
//            r = c*CB + (m == 0 ? p : PAIRS + p*N + m-1) of a GEMM operand
//            (c = symmetric component; the mode-0 rows feed the RHS).
//   phase 2  lane r forms C[r][kk] = sum_q W[kk][q] G[r][q] (kk = 0..N: the
//            projections onto w P_kk) for all kk from the mirror sums
//            E = G(q) + G(-q) and differences O = G(q) - G(-q): the weights
//            satisfy W[kk][-q] = (-1)^kk W[kk][q] exactly, so even kk take E,
//            odd kk take O, and the table entries are immediate operands.
//   phase 3  each lane writes 4 consecutive (source, mode) columns of one
//            component array (16-byte stores); the k_i = 0 components
//            (Theta_0 t_i) and the RHS of each source slot are formed here.
// With x0, the workgroup then reads back the rows it stored and forms its
// share of the off-diagonal product, sum_{j in group} A_ij x0_j, into
// apart[gy n + row] (the theta GEMV's arithmetic: FP32 values times FP64
// x0, FP64 sums).  The read is HBM traffic while the other workgroups of the
// CU keep its VALUs busy, which is what makes it cheaper than a matvec.
// HALF (GMRES-IR's FP16 copy T16, S16): phase 2 also takes each (target,
// source) block's largest |projection|, phase 3 stores the block scales, and
// the read-back (run then with or without x0) converts the stored values.
template <int N, int NQI, bool ANALYTIC, bool HALF = false>
__device__ __forceinline__ void assemble_theta_block(int i, int gy, int M, int64_t Lp, const double* __restrict__ pos,
                                                     const double* __restrict__ ori, float* __restrict__ T,
                                                     double* __restrict__ bpart, int jgroup, const PhysConsts& pc,
                                                     float* __restrict__ Tv, double* __restrict__ cpart,
                                                     const double* __restrict__ x0, double* __restrict__ apart,
                                                     _Float16* __restrict__ T16 = nullptr,
                                                     float* __restrict__ S16 = nullptr,
                                                     const SusDev* __restrict__ serving = nullptr) {
  constexpr Quad<N, NQI> quad{};
  constexpr int Q = 3 * NQI;
  constexpr int PAIRS = kPairThreads / Q;
  constexpr int NM = N + 1;
  constexpr int ROWS = 3 * N;
  constexpr int NC = 3 + 6 * (N - 1);
  constexpr int NR = ROWS + 6;  // RHS rows + the two velocity constants
  constexpr int PN = PAIRS * N;  // (source, mode) columns of one sweep
  constexpr int CB = PAIRS + PN;  // GEMM rows per component
  constexpr int RP = 6 * CB;  // GEMM rows
  constexpr int NV = PN / 4;  // float4 chunks per component and sweep
  static_assert(kPairThreads % Q == 0 && Q % 2 == 0 && PAIRS % 4 == 0, "sweep shapes");

  __shared__ __align__(16) float G[RP * Q];
  __shared__ __align__(16) float Pj[NM * RP];
  __shared__ double bred[PAIRS][NR];
  __shared__ float modes[2 * kMaxTerms];  // gamma_k, then eigen_k
  // FP16 copy: each source slot's largest |projection| (float bits), one
  // buffer per sweep parity
  __shared__ unsigned smax[2][PAIRS];

  const int tid = threadIdx.x;
  const int p = tid / Q;
  const int qi = tid - p * Q;
  const int jb = gy * jgroup;
  const int je = min(M, jb + jgroup);
  const int64_t n = static_cast<int64_t>(ROWS) * M;

  const double xi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  const double tid_[3] = {ori[3 * i], ori[3 * i + 1], ori[3 * i + 2]};
  const float ti[3] = {static_cast<float>(tid_[0]), static_cast<float>(tid_[1]), static_cast<float>(tid_[2])};
  const double sq_d = node_fp64<NQI>(qi);
  const float sq = static_cast<float>(sq_d);
  float* Ti = T + static_cast<int64_t>(i) * NC * Lp;
  float* Vi = Tv ? Tv + static_cast<int64_t>(i) * 12 * Lp : nullptr;

  // RHS (and velocity-constant) sums of each source slot, kept in LDS; the
  // mode factors gamma_k, eigen_k as a table the RHS lanes index by k
  for (int r = tid; r < PAIRS * NR; r += kPairThreads) bred[r / NR][r % NR] = 0.0;
  if (tid < 2 * PAIRS) smax[tid / PAIRS][tid % PAIRS] = 0u;
#pragma unroll
  for (int t = 0; t < kMaxTerms; ++t)
    if (tid == t) {
      modes[t] = pc.gamma[t];
      modes[kMaxTerms + t] = pc.eigen[t];
    }
  __syncthreads();


  for (int j0 = jb; j0 < je; j0 += PAIRS) {
    const int j = j0 + p;
    if (j < je && j != i) {
      float acc[NM][6];
      inner_integrals<N, NQI, ANALYTIC>(xi, tid_, ti, sq_d, sq, pos[3 * j], pos[3 * j + 1], pos[3 * j + 2],
                                        ori[3 * j], ori[3 * j + 1], ori[3 * j + 2], pc, acc);
#pragma unroll
      for (int m = 0; m < NM; ++m)
#pragma unroll
        for (int c = 0; c < 6; ++c) G[(c * CB + (m == 0 ? p : PAIRS + p * N + m - 1)) * Q + qi] = acc[m][c];
    }
    __syncthreads();
    // FP16 copy: this sweep's block maxima go to smax[par]; the other buffer,
    // last read in the previous sweep's phase 3 (before the barrier above), is
    // cleared for the next sweep
    const int par = ((j0 - jb) / PAIRS) & 1;
    if constexpr (HALF) {
      if (tid < PAIRS) smax[par ^ 1][tid] = 0u;
    }
    for (int r = tid; r < RP; r += kPairThreads) {
      const float* gr = G + r * Q;
      float E[Q / 2], O[Q / 2];
#pragma unroll
      for (int h = 0; h < Q / 2; ++h) {
        const float x = gr[h], y = gr[Q - 1 - h];
        E[h] = x + y;
        O[h] = x - y;
      }
      float g = 0.0f;
#pragma unroll
      for (int kk = 0; kk < NM; ++kk) {
        float acc = 0.0f;
#pragma unroll
        for (int h = 0; h < Q / 2; ++h) acc = fmaf(quad.W[kk][h], (kk & 1) ? O[h] : E[h], acc);
        Pj[kk * RP + r] = acc;
        if (kk > 0) g = fmaxf(g, fabsf(acc));
      }
      if constexpr (HALF) {  // rows of the modes m >= 1 feed the system components
        const int e = r % CB;
        if (e >= PAIRS) atomicMax(&smax[par][(e - PAIRS) / N], __float_as_uint(g));
      }
    }
    __syncthreads();
    const int npm = min(PAIRS, je - j0) * N;  // valid columns of this sweep
    const bool full = npm == PN;                // every sweep but possibly the last
    const bool self_here = i >= j0 && i < j0 + PAIRS;  // the self pair's columns are in this sweep
    const int sself = (i - j0) * N;
    float* const Tcol = Ti + static_cast<int64_t>(j0) * N;  // this sweep's first column, component 0
    // 4 consecutive columns of one component array (offset `off` from the
    // sweep's first column of component 0): self-pair columns zeroed, columns
    // past the sweep's last source not written.  The conditions are uniform
    // over the workgroup except in the sweep that holds them.
    auto put = [&](float* base, uint32_t off, int pm0, floatx4& val) {
      if (self_here) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (static_cast<unsigned>(pm0 + e - sself) < static_cast<unsigned>(N)) val[e] = 0.0f;
      }
      float* dst = base + off;
      if (full || pm0 + 3 < npm) {
        *reinterpret_cast<floatx4*>(dst) = val;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (pm0 + e < npm) dst[e] = val[e];
      }
    };
    const uint32_t lp32 = static_cast<uint32_t>(Lp);
    // system components: u < 3 Q0_u = (Theta_0 t_i)_u (Theta_0 = the P_1
    // projection), then Theta_{k_i} for k_i = 1..N-1 (projection k_i + 1)
    auto sys_val = [&](int u, int pm0) {
      floatx4 val;
      if (u < 3) {
        const float* b1 = Pj + RP + PAIRS + pm0;
        const floatx4 x0 = *reinterpret_cast<const floatx4*>(b1 + sym_index(0, u) * CB);
        const floatx4 x1 = *reinterpret_cast<const floatx4*>(b1 + sym_index(1, u) * CB);
        const floatx4 x2 = *reinterpret_cast<const floatx4*>(b1 + sym_index(2, u) * CB);
#pragma unroll
        for (int e = 0; e < 4; ++e) val[e] = x0[e] * ti[0] + x1[e] * ti[1] + x2[e] * ti[2];
      } else {
        const int kk = 2 + (u - 3) / 6;
        const int c = (u - 3) - 6 * (kk - 2);
        val = *reinterpret_cast<const floatx4*>(Pj + kk * RP + c * CB + PAIRS + pm0);
      }
      return val;
    };
    // FP16 copy: the sweep's block scales (S16); the values are converted
    // when the workgroup reads its rows back after the sweeps
    if constexpr (HALF) {
      if (tid < PAIRS && j0 + tid < je) {
        float sv, iv;
        half_block_scale(smax[par][tid], sv, iv);
        S16[static_cast<int64_t>(i) * M + j0 + tid] = sv;
      }
    }
    for (int job = tid; job < NC * NV; job += kPairThreads) {
      const int u = job / NV;
      const int pm0 = 4 * (job - u * NV);
      floatx4 val = sys_val(u, pm0);
      put(Tcol, static_cast<uint32_t>(u) * lp32 + static_cast<uint32_t>(pm0), pm0, val);
    }
    // velocity operator: Theta_0 (projection 1), then Phi_0 (projection 0).
    // Analytical form: its jobs start on the lanes after the system
    // components' last partial pass and the RHS jobs sit on the top lanes,
    // so the three waves get the same number of passes (measured faster
    // there, slower for the numerical form, which keeps the plain order).
    if (Vi) {
      float* const Vcol = Vi + static_cast<int64_t>(j0) * N;
      constexpr int kVShift = ANALYTIC ? kPairThreads - (NC * NV) % kPairThreads : 0;
      for (int job = (tid + kVShift) % kPairThreads; job < 12 * NV; job += kPairThreads) {
        const int w = job / NV;
        const int pm0 = 4 * (job - w * NV);
        const int kk = w < 6 ? 1 : 0;
        const int c = w < 6 ? w : w - 6;
        floatx4 val = *reinterpret_cast<const floatx4*>(Pj + kk * RP + c * CB + PAIRS + pm0);
        put(Vcol, static_cast<uint32_t>(w) * lp32 + static_cast<uint32_t>(pm0), pm0, val);
      }
    }
    // RHS (source slot p2, row block k): k < N forms rows 3k..3k+2 of b from
    // the projection of G_0 F/2 onto w P_{k+1}; k = N the velocity constants
    // sum_q w G_0 F/2 and sum_q w P_1 G_0 F/2.
    constexpr int kRShift = ANALYTIC ? PAIRS * (N + 1) % kPairThreads : 0;
    for (int rj = (tid + kRShift) % kPairThreads; rj < PAIRS * (N + 1); rj += kPairThreads) {
      const int p2 = rj / (N + 1), k = rj - p2 * (N + 1), j2 = j0 + p2;
      if (j2 >= je || j2 == i) continue;
      auto proj = [&](int kk, float f[3]) {  // sum_b Theta^(0)_kk[a][b] F_b / 2
        const float* b0 = Pj + kk * RP + p2;
        const float g0 = b0[0 * CB], g1 = b0[1 * CB], g2 = b0[2 * CB];
        const float g3 = b0[3 * CB], g4 = b0[4 * CB], g5 = b0[5 * CB];
        f[0] = g0 * pc.Fh[0] + g3 * pc.Fh[1] + g4 * pc.Fh[2];
        f[1] = g3 * pc.Fh[0] + g1 * pc.Fh[1] + g5 * pc.Fh[2];
        f[2] = g4 * pc.Fh[0] + g5 * pc.Fh[1] + g2 * pc.Fh[2];
      };
      double* br = bred[p2];
      if (k < N) {
        float TF[3];
        proj(k + 1, TF);
        const float gam = modes[k], eig = modes[kMaxTerms + k];
        const float QF = TF[0] * ti[0] + TF[1] * ti[1] + TF[2] * ti[2];
        double v[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
          v[a] = static_cast<double>((k == 0) ? pc.D1 * ti[a] * QF : gam * (TF[a] - eig * ti[a] * QF));
#pragma unroll
        for (int a = 0; a < 3; ++a) br[3 * k + a] -= v[a];
      } else {
        float TV[3], T1[3];
        proj(0, TV);
        proj(1, T1);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          br[ROWS + a] += static_cast<double>(TV[a]);
          br[ROWS + 3 + a] += static_cast<double>(T1[a]);
        }
      }
    }
  }
  __syncthreads();  // bred complete

  if (tid < ROWS) {
    double s = 0.0;
    for (int q = 0; q < PAIRS; ++q) s += bred[q][tid];
    bpart[static_cast<int64_t>(gy) * n + static_cast<int64_t>(ROWS) * i + tid] = s;
  } else if (cpart && tid < NR) {
    double s = 0.0;
    for (int q = 0; q < PAIRS; ++q) s += bred[q][tid];
    cpart[(static_cast<int64_t>(gy) * M + i) * 6 + (tid - ROWS)] = s;
  }

  // Serving: the suspension's fields are read only here, after the sweeps,
  // so that they hold no registers during them.
  double* r0v = nullptr;
  double* npart = nullptr;
  if (serving != nullptr) {
    x0 = serving->x0;
    r0v = serving->V;
    npart = serving->npart;
  }
  double out = 0.0;  // lane tid < ROWS: row tid of the off-diagonal product A x0
  const bool have_x = x0 != nullptr;
  // The workgroup reads back the rows it stored (visible to it after the
  // barrier above): for the product with x0, and for the FP16 copy, which
  // converts them with the block scales the sweeps stored (the values the
  // sweeps wrote, so the copy is theirs bit for bit; the self pair's columns
  // are zero already).
  constexpr int NACC = 1 + 3 * (N - 1);  // s0, then U_k (k >= 1) as in gemv_theta_block
  double acc[NACC];
#pragma unroll
  for (int a = 0; a < NACC; ++a) acc[a] = 0.0;
  if (have_x || HALF) {
    // column (i-row block, e): s0 += Q0 . x_e, U_k += Theta_k x_e
    auto add_column = [&](const float* q, double xa, double xb, double xc) {
      acc[0] = fma(static_cast<double>(q[0]), xa,
                   fma(static_cast<double>(q[Lp]), xb, fma(static_cast<double>(q[2 * Lp]), xc, acc[0])));
#pragma unroll
      for (int k = 1; k < N; ++k) {
        const float* qk = q + (3 + 6 * (k - 1)) * Lp;
        const double txx = qk[0], tyy = qk[Lp], tzz = qk[2 * Lp], txy = qk[3 * Lp], txz = qk[4 * Lp],
                     tyz = qk[5 * Lp];
        double* u = acc + 1 + 3 * (k - 1);
        u[0] = fma(txx, xa, fma(txy, xb, fma(txz, xc, u[0])));
        u[1] = fma(txy, xa, fma(tyy, xb, fma(tyz, xc, u[1])));
        u[2] = fma(txz, xa, fma(tyz, xb, fma(tzz, xc, u[2])));
      }
    };
    // Pairs of consecutive columns per lane with 8-byte loads (the group's
    // first column jb N is even -- jgroup is a multiple of the sources per
    // sweep, a multiple of 4 -- and Lp a multiple of 64, so the loads are
    // aligned), then the remaining column one per lane.
    const int ebeg = jb * N, eend = je * N;
    const int efull = ebeg + (eend - ebeg) / 2 * 2;
    // FP16 copy: the inverse of column e's block scale (1 / sv exactly, sv a
    // power of two; 0 for a zero or non-finite scale, as half_block_scale)
    _Float16* const Hi = HALF ? T16 + static_cast<int64_t>(i) * NC * Lp : nullptr;
    auto inv_scale = [&](int e) {
      const unsigned es = (__float_as_uint(S16[static_cast<int64_t>(i) * M + e / N]) >> 23) & 0xffu;
      return (es >= 1u && es < 255u) ? __uint_as_float((254u - es) << 23) : 0.0f;
    };
    for (int e0 = ebeg + 2 * tid; e0 < efull; e0 += 2 * kPairThreads) {
      const float* q = Ti + e0;
      auto ld = [&](int comp) { return *reinterpret_cast<const floatx2*>(q + static_cast<int64_t>(comp) * Lp); };
      if constexpr (HALF) {
        typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));
        const float iv0 = inv_scale(e0), iv1 = inv_scale(e0 + 1);
#pragma unroll
        for (int comp = 0; comp < NC; ++comp) {
          const floatx2 v = ld(comp);
          halfx2 h;
          h[0] = static_cast<_Float16>(v[0] * iv0);
          h[1] = static_cast<_Float16>(v[1] * iv1);
          *reinterpret_cast<halfx2*>(Hi + static_cast<int64_t>(comp) * Lp + e0) = h;
        }
        if (!have_x) continue;
      }
      double xv[2][3];
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int a = 0; a < 3; ++a) xv[c][a] = x0[3 * static_cast<int64_t>(e0 + c) + a];
      {
        const floatx2 q0 = ld(0), q1 = ld(1), q2 = ld(2);
#pragma unroll
        for (int c = 0; c < 2; ++c)
          acc[0] = fma(static_cast<double>(q0[c]), xv[c][0],
                       fma(static_cast<double>(q1[c]), xv[c][1], fma(static_cast<double>(q2[c]), xv[c][2], acc[0])));
      }
#pragma unroll
      for (int k = 1; k < N; ++k) {
        const int base = 3 + 6 * (k - 1);
        const floatx2 vxx = ld(base), vyy = ld(base + 1), vzz = ld(base + 2), vxy = ld(base + 3),
                      vxz = ld(base + 4), vyz = ld(base + 5);
        double* u = acc + 1 + 3 * (k - 1);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const double txx = vxx[c], tyy = vyy[c], tzz = vzz[c], txy = vxy[c], txz = vxz[c], tyz = vyz[c];
          const double xa = xv[c][0], xb = xv[c][1], xc = xv[c][2];
          u[0] = fma(txx, xa, fma(txy, xb, fma(txz, xc, u[0])));
          u[1] = fma(txy, xa, fma(tyy, xb, fma(tyz, xc, u[1])));
          u[2] = fma(txz, xa, fma(tyz, xb, fma(tzz, xc, u[2])));
        }
      }
    }
    for (int e = efull + tid; e < eend; e += kPairThreads) {
      if constexpr (HALF) {
        const float iv = inv_scale(e);
#pragma unroll
        for (int comp = 0; comp < NC; ++comp)
          Hi[static_cast<int64_t>(comp) * Lp + e] = static_cast<_Float16>(Ti[static_cast<int64_t>(comp) * Lp + e] * iv);
      }
      if (have_x)
        add_column(Ti + e, x0[3 * static_cast<int64_t>(e)], x0[3 * static_cast<int64_t>(e) + 1],
                   x0[3 * static_cast<int64_t>(e) + 2]);
    }
  }
  if (have_x) {  // uniform over the workgroup: the barrier below is reached by all or none
    constexpr int WAVES = kPairThreads / 64;
    static_assert(WAVES * NACC * sizeof(double) <= RP * Q * sizeof(float), "partials fit the freed G");
    double* red = reinterpret_cast<double*>(G);  // G is free after the last sweep
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int a = 0; a < NACC; ++a) {
      double v = acc[a];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) red[wave * NACC + a] = v;
    }
    __syncthreads();
    if (tid < ROWS) {
      const int k = tid / 3, a = tid - 3 * k;
      auto sum = [&](int idx) {
        double t = 0.0;
        for (int w = 0; w < WAVES; ++w) t += red[w * NACC + idx];
        return t;
      };
      if (k == 0) {
        out = static_cast<double>(pc.D1) * tid_[a] * sum(0);
      } else {
        const double U[3] = {sum(1 + 3 * (k - 1)), sum(2 + 3 * (k - 1)), sum(3 + 3 * (k - 1))};
        const double tU = tid_[0] * U[0] + tid_[1] * U[1] + tid_[2] * U[2];
        out = static_cast<double>(pc.gamma[k]) * (U[a] - static_cast<double>(pc.eigen[k]) * tid_[a] * tU);
      }
      if (apart) apart[static_cast<int64_t>(gy) * n + static_cast<int64_t>(ROWS) * i + tid] = out;
    }
  }
  // Serving (all sources in this workgroup, b written directly): the first
  // residual r0 = b - (x0 + A_offdiag x0) (r0 = b without x0) into r0v and the
  // target's squared norms of b and r0 into npart[2 i], npart[2 i + 1].
  if (npart != nullptr) {
    static_assert(2 * ROWS * sizeof(double) <= NM * RP * sizeof(float), "norm partials fit the freed Pj");
    double* nred = reinterpret_cast<double*>(Pj);  // Pj is free after the last sweep
    if (tid < ROWS) {
      double sb = 0.0;
      for (int q = 0; q < PAIRS; ++q) sb += bred[q][tid];  // this row of b, as stored above
      const int64_t row = static_cast<int64_t>(ROWS) * i + tid;
      const double r = (x0 != nullptr) ? sb - (out + x0[row]) : sb;
      r0v[row] = r;
      nred[tid] = sb * sb;
      nred[ROWS + tid] = r * r;
    }
    __syncthreads();
    if (tid == 0) {
      double b2 = 0.0, r2 = 0.0;
      for (int t = 0; t < ROWS; ++t) {
        b2 += nred[t];
        r2 += nred[ROWS + t];
      }
      npart[2 * i] = b2;
      npart[2 * i + 1] = r2;
    }
  }
}

// Waves per SIMD the theta kernel asks the register allocator for: as many
// as the workgroups whose LDS fits a CU's 160 KB can use, at most 3 for the
// analytical form (its FP64 moments need more than the 128 registers of 4
// waves; 3 waves allow 168).
template <int N, int NQI, bool ANALYTIC>
constexpr int theta_min_waves() {
  constexpr int Q = 3 * NQI, PAIRS = kPairThreads / Q, NM = N + 1;
  constexpr int RP = 6 * PAIRS * NM;
  constexpr int bytes = 4 * (RP * Q + NM * RP) + 8 * PAIRS * (3 * N + 6) + 8 * kMaxTerms;
  constexpr int blocks = (160 * 1024) / bytes;
  return (blocks >= 5 && !ANALYTIC) ? 4 : (blocks >= 4 ? 3 : (blocks >= 3 ? 2 : 1));
}

template <int N, int NQI, bool ANALYTIC, bool HALF>
__global__ __launch_bounds__(kPairThreads) __attribute__((amdgpu_waves_per_eu(theta_min_waves<N, NQI, ANALYTIC>(), 8))) void assemble_theta_kernel(
    int M, int64_t Lp, const double* __restrict__ pos, const double* __restrict__ ori, float* __restrict__ T,
    double* __restrict__ bpart, int jgroup, PhysConsts pc, float* __restrict__ Tv, double* __restrict__ cpart,
    const double* __restrict__ x0, double* __restrict__ apart, _Float16* __restrict__ T16, float* __restrict__ S16) {
  assemble_theta_block<N, NQI, ANALYTIC, HALF>(blockIdx.x, blockIdx.y, M, Lp, pos, ori, T, bpart, jgroup, pc, Tv,
                                               cpart, x0, apart, T16, S16);
}

// Serving: workgroup = (suspension, target fiber), all sources in one group,
// so b and the velocity constants are written directly, and with them the
// first GMRES residual (into the basis' first vector) and its norms.
template <int N, int NQI, bool ANALYTIC>
__global__ __launch_bounds__(kPairThreads) __attribute__((amdgpu_waves_per_eu(theta_min_waves<N, NQI, ANALYTIC>(), 8))) void assemble_theta_batch_kernel(
    const SusDev* __restrict__ sus, const int2* __restrict__ targets, PhysConsts pc) {
  const int2 e = targets[blockIdx.x];
  const SusDev& d = sus[e.x];
  assemble_theta_block<N, NQI, ANALYTIC>(e.y, 0, d.M, d.Lp, d.pos, d.ori, d.T, d.b, d.M, pc, d.Tv, d.cpart, nullptr,
                                         nullptr, nullptr, nullptr, &d);
}

// Dense A (column-major, leading dimension n) from the theta form; identity
// diagonal blocks.  One lane per (i, j, k_j).  Only the columns of source
// fibers j0 .. j0 + jcount - 1 (all n rows) are written, the column of source
// fiber j0's first unknown at A[0].
template <int N>
__global__ void expand_theta_kernel(int M, int64_t Lp, const float* __restrict__ T, const double* __restrict__ ori,
                                    PhysConsts pc, int j0, int jcount, float* __restrict__ A) {
  constexpr int NC = 3 + 6 * (N - 1);
  constexpr int ROWS = 3 * N;
  const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t total = static_cast<int64_t>(M) * jcount * N;
  if (id >= total) return;
  const int kj = static_cast<int>(id % N);
  const int64_t ij = id / N;
  const int jl = static_cast<int>(ij % jcount), i = static_cast<int>(ij / jcount);
  const int j = j0 + jl;
  const int64_t n = static_cast<int64_t>(ROWS) * M;
  const int64_t e = static_cast<int64_t>(j) * N + kj;
  const float* Ti = T + static_cast<int64_t>(i) * NC * Lp;
  const float t[3] = {static_cast<float>(ori[3 * i]), static_cast<float>(ori[3 * i + 1]),
                      static_cast<float>(ori[3 * i + 2])};
  for (int b = 0; b < 3; ++b) {
    float* col = A + (static_cast<int64_t>(ROWS) * jl + 3 * kj + b) * n + static_cast<int64_t>(ROWS) * i;
    if (i == j) {
      for (int r = 0; r < ROWS; ++r) col[r] = (r == 3 * kj + b) ? 1.0f : 0.0f;
      continue;
    }
    const float q0 = Ti[b * Lp + e];
    for (int a = 0; a < 3; ++a) col[a] = pc.D1 * t[a] * q0;
    for (int k = 1; k < N; ++k) {
      const float* Tk = Ti + (3 + 6 * (k - 1)) * Lp + e;
      float th[3], tht[3];
      for (int a = 0; a < 3; ++a) th[a] = Tk[sym_index(a, b) * Lp];
      // (Theta t)_b
      float qb = 0.0f;
      for (int a = 0; a < 3; ++a) {
        tht[a] = Tk[sym_index(a, b) * Lp];
        qb += tht[a] * t[a];
      }
      for (int a = 0; a < 3; ++a) col[3 * k + a] = pc.gamma[k] * (th[a] - pc.eigen[k] * t[a] * qb);
    }
  }
}

__global__ void sum_groups_kernel(int groups, int64_t len, const double* __restrict__ part, double* __restrict__ out) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= len) return;
  double s = 0.0;
  for (int g = 0; g < groups; ++g) s += part[static_cast<int64_t>(g) * len + r];
  out[r] = s;
}

__global__ void residual_groups_kernel(int groups, int64_t len, const double* __restrict__ apart,
                                       const double* __restrict__ b, const double* __restrict__ x0,
                                       double* __restrict__ r) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= len) return;
  double s = 0.0;
  for (int g = 0; g < groups; ++g) s += apart[static_cast<int64_t>(g) * len + k];
  r[k] = b[k] - (s + x0[k]);  // identity diagonal blocks
}

PairGrid pair_grid(int M, int pairs_per_tile) {
  PairGrid g;
  g.tiles = (M + pairs_per_tile - 1) / pairs_per_tile;
  // aim at ~16 workgroups per CU over 256 CUs
  int groups = (4096 + g.tiles - 1) / g.tiles;
  if (groups > M) groups = M;
  if (groups < 1) groups = 1;
  g.jgroup = (M + groups - 1) / groups;
  g.groups = (M + g.jgroup - 1) / g.jgroup;
  return g;
}

// Theta-form grid: blockIdx.x = target fiber, blockIdx.y = group of source
// fibers (a multiple of the 8 / 16 / 4 sources one iteration covers).
PairGrid theta_grid(int M, int sources_per_iteration) {
  PairGrid g;
  g.tiles = M;
  int groups = (4096 + M - 1) / M;
  const int max_groups = (M + sources_per_iteration - 1) / sources_per_iteration;
  if (groups > max_groups) groups = max_groups;
  if (groups < 1) groups = 1;
  int jg = (M + groups - 1) / groups;
  jg = (jg + sources_per_iteration - 1) / sources_per_iteration * sources_per_iteration;
  g.jgroup = jg;
  g.groups = (M + jg - 1) / jg;
  return g;
}

int max_pair_groups(int M) {
  int best = 1;
  for (int pairs : {4, 8, 16}) {
    const PairGrid g = pair_grid(M, pairs);
    if (g.groups > best) best = g.groups;
    const PairGrid t = theta_grid(M, pairs);
    if (t.groups > best) best = t.groups;
  }
  return best;
}

int64_t theta_stride(int M, int N) { return (static_cast<int64_t>(M) * N + 63) / 64 * 64; }

template <int N, int NQI, bool AN>
static hipError_t launch_asm_t(int M, int64_t lda, const double* pos, const double* ori, float* A, double* bpart,
                               const PhysConsts& pc, hipStream_t s) {
  constexpr int PAIRS = kPairThreads / (3 * NQI);
  const PairGrid g = pair_grid(M, PAIRS);
  hipLaunchKernelGGL((assemble_kernel<N, NQI, AN>), dim3(g.tiles, g.groups), dim3(kPairThreads), 0, s, M, lda, pos,
                     ori, A, bpart, g.jgroup, pc);
  return hipGetLastError();
}

template <int N, int NQI>
static hipError_t launch_asm_n(bool an, int M, int64_t lda, const double* pos, const double* ori, float* A,
                               double* bpart, const PhysConsts& pc, hipStream_t s) {
  return an ? launch_asm_t<N, NQI, true>(M, lda, pos, ori, A, bpart, pc, s)
            : launch_asm_t<N, NQI, false>(M, lda, pos, ori, A, bpart, pc, s);
}

bool assemble_supported(int N, int NQI) { return N >= 1 && N <= kMaxTerms && (NQI == 4 || NQI == 8 || NQI == 16); }

#define RF_ASM_CASE(NN)                                                                      \
  case NN:                                                                                   \
    if (NQI == 4) return launch_asm_n<NN, 4>(analytical, M, lda, pos, ori, A, bpart, pc, s); \
    if (NQI == 8) return launch_asm_n<NN, 8>(analytical, M, lda, pos, ori, A, bpart, pc, s); \
    if (NQI == 16) return launch_asm_n<NN, 16>(analytical, M, lda, pos, ori, A, bpart, pc, s); \
    return hipErrorInvalidValue;

hipError_t launch_assemble(int N, int NQI, bool analytical, int M, int64_t lda, const double* pos,
                           const double* ori, float* A, double* bpart, const PhysConsts& pc, hipStream_t s) {
  switch (N) {
    RF_ASM_CASE(1)
    RF_ASM_CASE(2)
    RF_ASM_CASE(3)
    RF_ASM_CASE(4)
    RF_ASM_CASE(5)
    RF_ASM_CASE(6)
    RF_ASM_CASE(7)
    RF_ASM_CASE(8)
    default:
      return hipErrorInvalidValue;
  }
}
#undef RF_ASM_CASE

template <int N, int NQI, bool AN>
static hipError_t launch_theta_t(int M, int64_t Lp, const double* pos, const double* ori, float* T, double* bpart,
                                 float* Tv, double* cpart, const double* x0, double* apart, const PhysConsts& pc,
                                 hipStream_t s, int* groups, _Float16* T16, float* S16) {
  constexpr int PAIRS = kPairThreads / (3 * NQI);
  const PairGrid g = theta_grid(M, PAIRS);
  *groups = g.groups;
  if (T16)  // the FP16 copy of GMRES-IR (a separate instantiation: the plain kernel keeps its registers)
    hipLaunchKernelGGL((assemble_theta_kernel<N, NQI, AN, true>), dim3(M, g.groups), dim3(kPairThreads), 0, s, M,
                       Lp, pos, ori, T, bpart, g.jgroup, pc, Tv, cpart, x0, apart, T16, S16);
  else
    hipLaunchKernelGGL((assemble_theta_kernel<N, NQI, AN, false>), dim3(M, g.groups), dim3(kPairThreads), 0, s, M,
                       Lp, pos, ori, T, bpart, g.jgroup, pc, Tv, cpart, x0, apart, nullptr, nullptr);
  return hipGetLastError();
}

template <int N, int NQI>
static hipError_t launch_theta_n(bool an, int M, int64_t Lp, const double* pos, const double* ori, float* T,
                                 double* bpart, float* Tv, double* cpart, const double* x0, double* apart,
                                 const PhysConsts& pc, hipStream_t s, int* groups, _Float16* T16, float* S16) {
  return an ? launch_theta_t<N, NQI, true>(M, Lp, pos, ori, T, bpart, Tv, cpart, x0, apart, pc, s, groups, T16, S16)
            : launch_theta_t<N, NQI, false>(M, Lp, pos, ori, T, bpart, Tv, cpart, x0, apart, pc, s, groups, T16, S16);
}

#define RF_TH_CASE(NN)                                                                             \
  case NN:                                                                                         \
    if (NQI == 4) return launch_theta_n<NN, 4>(analytical, M, Lp, pos, ori, T, bpart, Tv, cpart, x0, apart, pc, s, groups, T16, S16); \
    if (NQI == 8) return launch_theta_n<NN, 8>(analytical, M, Lp, pos, ori, T, bpart, Tv, cpart, x0, apart, pc, s, groups, T16, S16); \
    if (NQI == 16) return launch_theta_n<NN, 16>(analytical, M, Lp, pos, ori, T, bpart, Tv, cpart, x0, apart, pc, s, groups, T16, S16); \
    return hipErrorInvalidValue;

hipError_t launch_assemble_theta(int N, int NQI, bool analytical, int M, int64_t Lp, const double* pos,
                                 const double* ori, float* T, double* bpart, float* Tv, double* cpart,
                                 const double* x0, double* apart, const PhysConsts& pc, hipStream_t s, int* groups,
                                 _Float16* T16, float* S16) {
  switch (N) {
    RF_TH_CASE(1)
    RF_TH_CASE(2)
    RF_TH_CASE(3)
    RF_TH_CASE(4)
    RF_TH_CASE(5)
    RF_TH_CASE(6)
    RF_TH_CASE(7)
    RF_TH_CASE(8)
    default:
      return hipErrorInvalidValue;
  }
}
#undef RF_TH_CASE

// The whole dense A from the theta form, for the LU's factor storage
// (leading dimension lda), with the arithmetic of expand_theta_kernel (the
// identity for i == j).  A workgroup takes TI = 256 / 3N targets x 32 entries
// (entry e: source j = e / N, mode k_j = e % N): it stages their 3 + 6(N-1)
// numbers in LDS, one 128-byte row per (target, component), then writes the
// tile's 96 columns 3 e + b, each a run of TI x 3N contiguous rows 3N i + r,
// one row per lane.
template <int N>
constexpr int exp_targets() { return 256 / (3 * N); }
constexpr int kExpEntries = 32;
constexpr int kExpRow = kExpEntries + 4;  // padded LDS row (16-byte aligned, banks spread)

template <int N>
__global__ __launch_bounds__(256) void expand_theta_dense_kernel(int M, int64_t Lp, const float* __restrict__ T,
                                                                 const double* __restrict__ ori, PhysConsts pc,
                                                                 float* __restrict__ A, int64_t lda) {
  constexpr int NC = 3 + 6 * (N - 1);
  constexpr int ROWS = 3 * N;
  constexpr int TI = exp_targets<N>();
  constexpr int TROWS = TI * ROWS;  // rows of a column in the tile, <= 256
  constexpr int kChunks = TI * NC * (kExpEntries / 4);  // 16-byte chunks of the tile
  constexpr int kPer = (kChunks + 255) / 256;           // per thread
  // the tile's rows, then rows for the last round's chunks past the tile
  __shared__ __align__(16) float Ts[kPer * 256 / (kExpEntries / 4) * kExpRow];
  const int tid = threadIdx.x;
  const int64_t L = static_cast<int64_t>(M) * N;
  // Tile order: workgroup w runs on XCD w % 8; each XCD takes a contiguous
  // run of tiles, target groups fastest, so the column runs that continue
  // each other (and share their first and last 128-byte lines) are written
  // one after another through the same L2.
  const int ny = (M + TI - 1) / TI;
  const int tiles = static_cast<int>((L + kExpEntries - 1) / kExpEntries) * ny;
  const int per = (tiles + 7) / 8;
  const int v = (static_cast<int>(blockIdx.x) & 7) * per + (static_cast<int>(blockIdx.x) >> 3);
  if (v >= tiles) return;  // uniform over the workgroup, before any barrier
  const int64_t e0 = static_cast<int64_t>(v / ny) * kExpEntries;
  const int i0 = (v % ny) * TI;
  // stage: row (ti, c) holds T[i0 + ti][c][e0 .. e0 + 31] (within the padded
  // row of Lp >= e0 + 32 numbers).  Every load of the thread is in flight
  // before the LDS stores, which are unconditional (a guarded store would let
  // the compiler sink each load to it): chunks past the tile, and rows of
  // targets past M, load the tile's first chunk and store it where nothing
  // reads it.
  float4 sv[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    int q = tid + 256 * u;
    int row = q / (kExpEntries / 4);
    if (q >= kChunks || i0 + row / NC >= M) q = 0, row = 0;
    const int part = q - row * (kExpEntries / 4);
    const int ti = row / NC, c = row - ti * NC;
    sv[u] = *reinterpret_cast<const float4*>(T + (static_cast<int64_t>(i0 + ti) * NC + c) * Lp + e0 + 4 * part);
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int q = tid + 256 * u;
    const int row = q / (kExpEntries / 4), part = q - row * (kExpEntries / 4);
    *reinterpret_cast<float4*>(&Ts[row * kExpRow + 4 * part]) = sv[u];
  }
  __syncthreads();
  if (tid >= TROWS) return;  // no barrier follows
  const int ti = tid / ROWS, r = tid - ti * ROWS;
  const int i = i0 + ti;
  if (i >= M) return;
  const int k = r / 3, a = r - 3 * k;
  const float t[3] = {static_cast<float>(ori[3 * i]), static_cast<float>(ori[3 * i + 1]),
                      static_cast<float>(ori[3 * i + 2])};
  float gk = 0.0f, ek = 0.0f;
#pragma unroll
  for (int kk = 1; kk < N; ++kk)
    if (k == kk) {
      gk = pc.gamma[kk];
      ek = pc.eigen[kk];
    }
  const float* Tr = Ts + ti * NC * kExpRow;  // this target's rows
  const int cbase = (k == 0) ? 0 : 3 + 6 * (k - 1);
  const int ecount = static_cast<int>(min(static_cast<int64_t>(kExpEntries), L - e0));
  float* const out = A + static_cast<int64_t>(ROWS) * i0 + tid;
  for (int ee = 0; ee < ecount; ++ee) {
    const int64_t e = e0 + ee;
    const int j = static_cast<int>(e / N), kj = static_cast<int>(e - static_cast<int64_t>(j) * N);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      float v;
      if (i == j) {
        v = (r == 3 * kj + b) ? 1.0f : 0.0f;
      } else if (k == 0) {
        v = pc.D1 * t[a] * Tr[b * kExpRow + ee];
      } else {
        float qb = 0.0f;
#pragma unroll
        for (int a2 = 0; a2 < 3; ++a2) qb += Tr[(cbase + sym_index(a2, b)) * kExpRow + ee] * t[a2];
        v = gk * (Tr[(cbase + sym_index(a, b)) * kExpRow + ee] - ek * t[a] * qb);
      }
      out[(3 * e + b) * lda] = v;
    }
  }
}

hipError_t launch_expand_theta_dense(int N, int M, int64_t Lp, const float* T, const double* ori,
                                     const PhysConsts& pc, float* A, int64_t lda, hipStream_t s) {
  if (M < 1 || lda < static_cast<int64_t>(3) * N * M || Lp % kExpEntries != 0) return hipErrorInvalidValue;
  const int64_t ex = (static_cast<int64_t>(M) * N + kExpEntries - 1) / kExpEntries;
  switch (N) {
#define RF_EXD(NN)                                                                                          \
  case NN: {                                                                                                \
    const int64_t tiles = ex * ((M + exp_targets<NN>() - 1) / exp_targets<NN>());                          \
    if (tiles > INT32_MAX - 8) return hipErrorInvalidValue;                                                 \
    hipLaunchKernelGGL(expand_theta_dense_kernel<NN>, dim3(static_cast<unsigned>((tiles + 7) / 8 * 8)),     \
                       dim3(256), 0, s, M, Lp, T, ori, pc, A, lda);                                         \
    break;                                                                                                  \
  }
    RF_EXD(1)
    RF_EXD(2)
    RF_EXD(3)
    RF_EXD(4)
    RF_EXD(5)
    RF_EXD(6)
    RF_EXD(7)
    RF_EXD(8)
#undef RF_EXD
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_expand_theta(int N, int M, int64_t Lp, const float* T, const double* ori, const PhysConsts& pc,
                               float* A, hipStream_t s, int j0, int jcount) {
  if (jcount < 0) jcount = M - j0;
  if (j0 < 0 || jcount < 1 || j0 + jcount > M) return hipErrorInvalidValue;
  const int64_t total = static_cast<int64_t>(M) * jcount * N;
  const unsigned blocks = static_cast<unsigned>((total + 255) / 256);
  switch (N) {
#define RF_EX(NN) \
  case NN:        \
    hipLaunchKernelGGL(expand_theta_kernel<NN>, dim3(blocks), dim3(256), 0, s, M, Lp, T, ori, pc, j0, jcount, A); \
    break;
    RF_EX(1)
    RF_EX(2)
    RF_EX(3)
    RF_EX(4)
    RF_EX(5)
    RF_EX(6)
    RF_EX(7)
    RF_EX(8)
#undef RF_EX
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int N, int NQI>
static hipError_t launch_theta_batch_n(bool an, int ntargets, const SusDev* sus, const int2* targets,
                                       const PhysConsts& pc, hipStream_t s) {
  if (an)
    hipLaunchKernelGGL((assemble_theta_batch_kernel<N, NQI, true>), dim3(ntargets), dim3(kPairThreads), 0, s, sus,
                       targets, pc);
  else
    hipLaunchKernelGGL((assemble_theta_batch_kernel<N, NQI, false>), dim3(ntargets), dim3(kPairThreads), 0, s, sus,
                       targets, pc);
  return hipGetLastError();
}

hipError_t launch_assemble_theta_batch(int N, int NQI, bool analytical, int ntargets, const SusDev* sus,
                                       const int2* targets, const PhysConsts& pc, hipStream_t s) {
  if (ntargets <= 0) return hipSuccess;
  switch (N) {
#define RF_TB(NN)                                                                                 \
  case NN:                                                                                        \
    if (NQI == 4) return launch_theta_batch_n<NN, 4>(analytical, ntargets, sus, targets, pc, s);   \
    if (NQI == 8) return launch_theta_batch_n<NN, 8>(analytical, ntargets, sus, targets, pc, s);   \
    if (NQI == 16) return launch_theta_batch_n<NN, 16>(analytical, ntargets, sus, targets, pc, s); \
    return hipErrorInvalidValue;
    RF_TB(1)
    RF_TB(2)
    RF_TB(3)
    RF_TB(4)
    RF_TB(5)
    RF_TB(6)
    RF_TB(7)
    RF_TB(8)
#undef RF_TB
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_residual_groups(int groups, int64_t len, const double* apart, const double* b, const double* x0,
                                  double* r, hipStream_t s) {
  const int threads = 256;
  const int64_t blocks = (len + threads - 1) / threads;
  hipLaunchKernelGGL(residual_groups_kernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0, s, groups, len,
                     apart, b, x0, r);
  return hipGetLastError();
}

hipError_t launch_sum_groups(int groups, int64_t len, const double* part, double* out, hipStream_t s) {
  const int threads = 256;
  const int64_t blocks = (len + threads - 1) / threads;
  hipLaunchKernelGGL(sum_groups_kernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0, s, groups, len, part,
                     out);
  return hipGetLastError();
}

}  // namespace rf
