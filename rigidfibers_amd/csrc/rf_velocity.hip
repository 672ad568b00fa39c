// rf_velocity.hip -- fiber velocities from the solved force coefficients, and
// the second-order multistep position/orientation update.
//
// Reference: fortran/update_velocities.incl:1-89 (velocities),
// src/kernels/reset_velocities.cu:6-39 (isolated-fiber term),
// src/kernels/update_velocities_2D.cu:6-166 (GPU form, quadrature always),
// fortran/update_fibers_firststep.incl:1-10, fortran/update_fibers.incl:1-17.
//
// v_i = (0.5/d) [ (d+2) F + (d-2)(t_i.F) t_i + sum_{j!=i} sum_{q_i} w GF_ij(q_i) ]
// w_i = (1.5/d) (I - t_i t_i^T) sum_{j!=i} sum_{q_i} w P_1(s_{q_i}) GF_ij(q_i)
// GF_ij(q_i) = int K(X_i(q_i) - X_j(s)) f_j(s) ds,  f_j(s) = F/2 + sum_k x_{j,k} P_{k+1}(s).
//
// Both sums are linear in GF, so lane (i, q_i) accumulates GF over every source
// fiber of its group in FP64 and the q_i weighting happens once at the end.
// The source data of fiber j (uniform across the workgroup) is read with
// scalar loads: the weighted force samples w_q f_j(s_q) (quadrature) or the
// monomial coefficients of f_j (exact integral, FP64 moments).
#include "rf_internal.h"

namespace rf {

constexpr int kPairThreads = 192;

// wf[j][q][a] = w_q (F/2 + sum_k x_{j,k} P_{k+1}(s_q))_a  (quadrature form)
// fm[j][p][a] = monomial coefficients of F/2 + sum_k x_{j,k} P_{k+1}(s)  (exact form)
template <int N, int NQI>
__global__ void force_density_kernel(int M, const double* __restrict__ x, PhysConsts pc, float* __restrict__ wf,
                                     double* __restrict__ fm) {
  constexpr Quad<N, NQI> quad{};
  constexpr int Q = 3 * NQI;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  double c[N][3];
#pragma unroll
  for (int k = 0; k < N; ++k)
#pragma unroll
    for (int a = 0; a < 3; ++a) c[k][a] = x[static_cast<int64_t>(3 * N) * j + 3 * k + a];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float f = pc.Fh[a];
#pragma unroll
      for (int k = 0; k < N; ++k) f = fmaf(static_cast<float>(c[k][a]), quad.P[k + 1][q], f);
      wf[(static_cast<int64_t>(j) * Q + q) * 3 + a] = quad.w[q] * f;
    }
  }
  constexpr LegendreCoef lc = kLeg;
#pragma unroll
  for (int p = 0; p <= N; ++p)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double s = lc.c[0][p] * pc.Fh_d[a];
#pragma unroll
      for (int k = 0; k < N; ++k) s += lc.c[k + 1][p] * c[k][a];
      fm[(static_cast<int64_t>(j) * (N + 1) + p) * 3 + a] = s;
    }
}

template <int N, int NQI, bool ANALYTIC>
__global__ __launch_bounds__(kPairThreads) void velocity_kernel(int M, const double* __restrict__ pos,
                                                                const double* __restrict__ ori,
                                                                const float* __restrict__ wf,
                                                                const double* __restrict__ fm,
                                                                double* __restrict__ vpart, int jgroup,
                                                                PhysConsts pc) {
  constexpr Quad<N, NQI> quad{};
  constexpr int Q = 3 * NQI;
  constexpr int PAIRS = kPairThreads / Q;
  __shared__ double red[PAIRS][Q][6];

  const int tid = threadIdx.x;
  const int p = tid / Q;
  const int qi = tid - p * Q;
  const int i = blockIdx.x * PAIRS + p;
  const bool ivalid = i < M;
  const int j0 = blockIdx.y * jgroup;
  const int j1 = min(M, j0 + jgroup);

  double xi[3] = {0, 0, 0}, tid_[3] = {0, 0, 0};
  if (ivalid) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      xi[a] = pos[3 * i + a];
      tid_[a] = ori[3 * i + a];
    }
  }
  const double sq_d = node_fp64<NQI>(qi);
  const float sq = static_cast<float>(sq_d);
  const float ti[3] = {static_cast<float>(tid_[0]), static_cast<float>(tid_[1]), static_cast<float>(tid_[2])};
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;

  for (int j = j0; j < j1; ++j) {
    if (!ivalid || j == i) continue;
    const double xj0 = pos[3 * j + 0], xj1 = pos[3 * j + 1], xj2 = pos[3 * j + 2];
    const double tj0 = ori[3 * j + 0], tj1 = ori[3 * j + 1], tj2 = ori[3 * j + 2];
    if constexpr (!ANALYTIC) {
      const float ax = fmaf(sq, ti[0], static_cast<float>(xi[0] - xj0));
      const float ay = fmaf(sq, ti[1], static_cast<float>(xi[1] - xj1));
      const float az = fmaf(sq, ti[2], static_cast<float>(xi[2] - xj2));
      const float tx = static_cast<float>(tj0), ty = static_cast<float>(tj1), tz = static_cast<float>(tj2);
      const float* f = wf + static_cast<int64_t>(j) * Q * 3;
      float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
      for (int h = 0; h < Q / 2; ++h) {
        const float s = quad.s[h];
        const float ux = s * tx, uy = s * ty, uz = s * tz;
        float Kp[6], Km[6];
        stokes_kernel(ax - ux, ay - uy, az - uz, pc.eps2, Kp);
        stokes_kernel(ax + ux, ay + uy, az + uz, pc.eps2, Km);
        const int hm = Q - 1 - h;
        const float fp0 = f[3 * h + 0], fp1 = f[3 * h + 1], fp2 = f[3 * h + 2];
        const float fm0 = f[3 * hm + 0], fm1 = f[3 * hm + 1], fm2 = f[3 * hm + 2];
        s0 += Kp[0] * fp0 + Kp[3] * fp1 + Kp[4] * fp2 + Km[0] * fm0 + Km[3] * fm1 + Km[4] * fm2;
        s1 += Kp[3] * fp0 + Kp[1] * fp1 + Kp[5] * fp2 + Km[3] * fm0 + Km[1] * fm1 + Km[5] * fm2;
        s2 += Kp[4] * fp0 + Kp[5] * fp1 + Kp[2] * fp2 + Km[4] * fm0 + Km[5] * fm1 + Km[2] * fm2;
      }
      g0 += s0;
      g1 += s1;
      g2 += s2;
    } else {
      const double R0[3] = {xi[0] - xj0 + sq_d * tid_[0], xi[1] - xj1 + sq_d * tid_[1],
                            xi[2] - xj2 + sq_d * tid_[2]};
      const double ll = R0[0] * R0[0] + R0[1] * R0[1] + R0[2] * R0[2];
      const double* C = fm + static_cast<int64_t>(j) * (N + 1) * 3;
      if (ll >= kFarField[N] * kFarField[N]) {
        // far field: the assembly's 8-point Gauss-Legendre rule over the
        // source fiber, f_j(+-s_k) from its monomial coefficients
        constexpr FarQuad<N> far{};
        const float rx = static_cast<float>(R0[0]), ry = static_cast<float>(R0[1]), rz = static_cast<float>(R0[2]);
        const float tx = static_cast<float>(tj0), ty = static_cast<float>(tj1), tz = static_cast<float>(tj2);
        double acc[3] = {0, 0, 0};
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          const double sk = kGL8Node[h], wk = kGL8Weight[h];
          double fp[3] = {0, 0, 0}, fn[3] = {0, 0, 0}, pw = 1.0;
#pragma unroll
          for (int q = 0; q <= N; ++q) {
            const double sgn = (q & 1) ? -pw : pw;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              fp[a] = fma(C[3 * q + a], pw, fp[a]);
              fn[a] = fma(C[3 * q + a], sgn, fn[a]);
            }
            pw *= sk;
          }
          const float s = far.s[h];
          const float ux = s * tx, uy = s * ty, uz = s * tz;
          float Kp[6], Km[6];
          stokes_kernel(rx - ux, ry - uy, rz - uz, pc.eps2, Kp);  // node +s
          stokes_kernel(rx + ux, ry + uy, rz + uz, pc.eps2, Km);  // node -s
          acc[0] += wk * (Kp[0] * fp[0] + Kp[3] * fp[1] + Kp[4] * fp[2] + Km[0] * fn[0] + Km[3] * fn[1] + Km[4] * fn[2]);
          acc[1] += wk * (Kp[3] * fp[0] + Kp[1] * fp[1] + Kp[5] * fp[2] + Km[3] * fn[0] + Km[1] * fn[1] + Km[5] * fn[2]);
          acc[2] += wk * (Kp[4] * fp[0] + Kp[5] * fp[1] + Kp[2] * fp[2] + Km[4] * fn[0] + Km[5] * fn[1] + Km[2] * fn[2]);
        }
        g0 += acc[0];
        g1 += acc[1];
        g2 += acc[2];
        continue;
      }
      const double t[3] = {tj0, tj1, tj2};
      const double bb = -2.0 * (R0[0] * tj0 + R0[1] * tj1 + R0[2] * tj2);
      double I1[N + 3], I3[N + 3], I5[N + 3];
      moments<N>(bb, ll, I1, I3, I5);
      const double two_e2 = 2.0 * pc.eps2_d;
      double acc[3] = {0, 0, 0};
#pragma unroll
      for (int q = 0; q <= N; ++q) {
        const double c0 = C[3 * q + 0], c1 = C[3 * q + 1], c2 = C[3 * q + 2];
        const double rc = R0[0] * c0 + R0[1] * c1 + R0[2] * c2;
        const double tc = t[0] * c0 + t[1] * c1 + t[2] * c2;
        const double iso = I1[q] + two_e2 * I3[q];
        const double crr = (I3[q] - 3.0 * two_e2 * I5[q]) * rc;
        const double ctr = -I3[q + 1] + 3.0 * two_e2 * I5[q + 1];
        const double ctt = (I3[q + 2] - 3.0 * two_e2 * I5[q + 2]) * tc;
        const double cc[3] = {c0, c1, c2};
#pragma unroll
        for (int a = 0; a < 3; ++a) acc[a] += iso * cc[a] + crr * R0[a] + ctr * (t[a] * rc + R0[a] * tc) + ctt * t[a];
      }
      g0 += acc[0];
      g1 += acc[1];
      g2 += acc[2];
    }
  }

  const double wq = static_cast<double>(quad_weight_fp32<NQI>(qi));
  const double wp = wq * static_cast<double>(static_cast<float>(sq_d));
  red[p][qi][0] = wq * g0;
  red[p][qi][1] = wq * g1;
  red[p][qi][2] = wq * g2;
  red[p][qi][3] = wp * g0;
  red[p][qi][4] = wp * g1;
  red[p][qi][5] = wp * g2;
  __syncthreads();
  if (tid < PAIRS * 6) {
    const int p2 = tid / 6, c = tid - 6 * (tid / 6);
    const int i2 = blockIdx.x * PAIRS + p2;
    if (i2 < M) {
      double s = 0.0;
      for (int q = 0; q < Q; ++q) s += red[p2][q][c];
      vpart[(static_cast<int64_t>(blockIdx.y) * M + i2) * 6 + c] = s;
    }
  }
}

// v_i, w_i from the per-group sums (fixed order over groups).
__global__ void velocity_finalize_kernel(int M, int groups, const double* __restrict__ ori,
                                         const double* __restrict__ vpart, PhysConsts pc, double* __restrict__ tvel,
                                         double* __restrict__ rvel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  double S[6] = {0, 0, 0, 0, 0, 0};
  for (int g = 0; g < groups; ++g)
#pragma unroll
    for (int c = 0; c < 6; ++c) S[c] += vpart[(static_cast<int64_t>(g) * M + i) * 6 + c];
  const double t[3] = {ori[3 * i], ori[3 * i + 1], ori[3 * i + 2]};
  const double F[3] = {2.0 * pc.Fh_d[0], 2.0 * pc.Fh_d[1], 2.0 * pc.Fh_d[2]};
  const double d = pc.d;
  const double tF = t[0] * F[0] + t[1] * F[1] + t[2] * F[2];
  const double tS = t[0] * S[3] + t[1] * S[4] + t[2] * S[5];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    tvel[3 * i + a] = (0.5 / d) * ((d + 2.0) * F[a] + (d - 2.0) * tF * t[a] + S[a]);
    rvel[3 * i + a] = (1.5 / d) * (S[3 + a] - t[a] * tS);
  }
}

// Velocities from the operator the theta assembly stored: per target i,
//   S_v = c_v + sum_{j,k} Phi0_{ij,k+1} x_{j,k},   S_w = c_w + sum_{j,k} Theta_{ij,0,k} x_{j,k}
// (Tv components s = xx yy zz xy xz yz of Theta_0, then of Phi0), which are the
// sums velocity_kernel forms pair by pair.  One workgroup per target fiber
// streams its 12 component arrays once; FP64 accumulation, fixed-order
// reduction.
constexpr int kVelStoredThreads = 256;

__device__ __forceinline__ void velocity_stored_block(int i, int M, int N, int64_t Lp, const float* __restrict__ Tv,
                                                      const double* __restrict__ cpart, int groups,
                                                      const double* __restrict__ ori, const double* __restrict__ x,
                                                      const PhysConsts& pc, double* __restrict__ tvel,
                                                      double* __restrict__ rvel) {
  __shared__ double red[kVelStoredThreads / 64][6];
  const float* Vi = Tv + static_cast<int64_t>(i) * 12 * Lp;
  const int L = M * N;
  double sw[3] = {0.0, 0.0, 0.0}, sv[3] = {0.0, 0.0, 0.0};
  auto column = [&](int e) {
    const double x0 = x[3 * static_cast<int64_t>(e)], x1 = x[3 * static_cast<int64_t>(e) + 1],
                 x2 = x[3 * static_cast<int64_t>(e) + 2];
    const float* p = Vi + e;
    const double wxx = __builtin_nontemporal_load(p), wyy = __builtin_nontemporal_load(p + Lp),
                 wzz = __builtin_nontemporal_load(p + 2 * Lp), wxy = __builtin_nontemporal_load(p + 3 * Lp),
                 wxz = __builtin_nontemporal_load(p + 4 * Lp), wyz = __builtin_nontemporal_load(p + 5 * Lp);
    const double vxx = __builtin_nontemporal_load(p + 6 * Lp), vyy = __builtin_nontemporal_load(p + 7 * Lp),
                 vzz = __builtin_nontemporal_load(p + 8 * Lp), vxy = __builtin_nontemporal_load(p + 9 * Lp),
                 vxz = __builtin_nontemporal_load(p + 10 * Lp), vyz = __builtin_nontemporal_load(p + 11 * Lp);
    sw[0] = fma(wxx, x0, fma(wxy, x1, fma(wxz, x2, sw[0])));
    sw[1] = fma(wxy, x0, fma(wyy, x1, fma(wyz, x2, sw[1])));
    sw[2] = fma(wxz, x0, fma(wyz, x1, fma(wzz, x2, sw[2])));
    sv[0] = fma(vxx, x0, fma(vxy, x1, fma(vxz, x2, sv[0])));
    sv[1] = fma(vxy, x0, fma(vyy, x1, fma(vyz, x2, sv[1])));
    sv[2] = fma(vxz, x0, fma(vyz, x1, fma(vzz, x2, sv[2])));
  };
  // two columns per pass (24 loads in flight per lane), in the same order
  int e = threadIdx.x;
  for (; e + kVelStoredThreads < L; e += 2 * kVelStoredThreads) {
    column(e);
    column(e + kVelStoredThreads);
  }
  if (e < L) column(e);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double acc[6] = {sv[0], sv[1], sv[2], sw[0], sw[1], sw[2]};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double v = acc[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][c] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double S[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += cpart[(static_cast<int64_t>(g) * M + i) * 6 + c];
#pragma unroll
    for (int w = 0; w < kVelStoredThreads / 64; ++w) s += red[w][c];
    S[c] = s;
  }
  const double t[3] = {ori[3 * i], ori[3 * i + 1], ori[3 * i + 2]};
  const double F[3] = {2.0 * pc.Fh_d[0], 2.0 * pc.Fh_d[1], 2.0 * pc.Fh_d[2]};
  const double d = pc.d;
  const double tF = t[0] * F[0] + t[1] * F[1] + t[2] * F[2];
  const double tS = t[0] * S[3] + t[1] * S[4] + t[2] * S[5];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    tvel[3 * i + a] = (0.5 / d) * ((d + 2.0) * F[a] + (d - 2.0) * tF * t[a] + S[a]);
    rvel[3 * i + a] = (1.5 / d) * (S[3 + a] - t[a] * tS);
  }
}

__global__ __launch_bounds__(kVelStoredThreads) void velocity_stored_kernel(
    int M, int N, int64_t Lp, const float* __restrict__ Tv, const double* __restrict__ cpart, int groups,
    const double* __restrict__ ori, const double* __restrict__ x, PhysConsts pc, double* __restrict__ tvel,
    double* __restrict__ rvel) {
  velocity_stored_block(blockIdx.x, M, N, Lp, Tv, cpart, groups, ori, x, pc, tvel, rvel);
}

// Serving: workgroup = (suspension, target fiber); one group of velocity constants.
__global__ __launch_bounds__(kVelStoredThreads) void velocity_stored_batch_kernel(int N,
                                                                                  const SusDev* __restrict__ sus,
                                                                                  const int2* __restrict__ targets,
                                                                                  PhysConsts pc) {
  const int2 e = targets[blockIdx.x];
  const SusDev& d = sus[e.x];
  velocity_stored_block(e.y, d.M, N, d.Lp, d.Tv, d.cpart, 1, d.ori, d.x, pc, d.tv, d.rv);
}

hipError_t launch_velocities_stored_batch(int N, int ntargets, const SusDev* sus, const int2* targets,
                                          const PhysConsts& pc, hipStream_t s) {
  if (ntargets <= 0) return hipSuccess;
  hipLaunchKernelGGL(velocity_stored_batch_kernel, dim3(ntargets), dim3(kVelStoredThreads), 0, s, N, sus, targets,
                     pc);
  return hipGetLastError();
}

hipError_t launch_velocities_stored(int N, int M, int64_t Lp, const float* Tv, const double* cpart, int groups,
                                    const double* ori, const double* x, const PhysConsts& pc, double* tvel,
                                    double* rvel, hipStream_t s) {
  hipLaunchKernelGGL(velocity_stored_kernel, dim3(M), dim3(kVelStoredThreads), 0, s, M, N, Lp, Tv, cpart, groups, ori,
                     x, pc, tvel, rvel);
  return hipGetLastError();
}

__device__ __forceinline__ void update_fiber(int i, int first, double dt, const double* __restrict__ pos_prev,
                                             const double* __restrict__ pos, const double* __restrict__ ori_prev,
                                             const double* __restrict__ ori, const double* __restrict__ tv_prev,
                                             const double* __restrict__ tv, const double* __restrict__ rv_prev,
                                             const double* __restrict__ rv, double* __restrict__ pos_next,
                                             double* __restrict__ ori_next) {
  double o[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int k = 3 * i + a;
    if (first) {
      pos_next[k] = pos[k] + dt * tv[k];
      o[a] = ori[k] + dt * rv[k];
    } else {
      pos_next[k] = (4.0 / 3.0) * pos[k] - (1.0 / 3.0) * pos_prev[k] + (2.0 / 3.0) * dt * (2.0 * tv[k] - tv_prev[k]);
      o[a] = (4.0 / 3.0) * ori[k] - (1.0 / 3.0) * ori_prev[k] + (2.0 / 3.0) * dt * (2.0 * rv[k] - rv_prev[k]);
    }
  }
  const double inv = 1.0 / sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
#pragma unroll
  for (int a = 0; a < 3; ++a) ori_next[3 * i + a] = o[a] * inv;
}

__global__ void update_kernel(int first, int M, double dt, const double* __restrict__ pos_prev,
                              const double* __restrict__ pos, const double* __restrict__ ori_prev,
                              const double* __restrict__ ori, const double* __restrict__ tv_prev,
                              const double* __restrict__ tv, const double* __restrict__ rv_prev,
                              const double* __restrict__ rv, double* __restrict__ pos_next,
                              double* __restrict__ ori_next) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) update_fiber(i, first, dt, pos_prev, pos, ori_prev, ori, tv_prev, tv, rv_prev, rv, pos_next, ori_next);
}

// Serving: entry (suspension, block of 256 fibers).
__global__ __launch_bounds__(256) void update_batch_kernel(const SusDev* __restrict__ sus,
                                                           const int2* __restrict__ fblocks, double dt) {
  const int2 e = fblocks[blockIdx.x];
  const SusDev& d = sus[e.x];
  const int i = e.y * 256 + threadIdx.x;
  if (i < d.M)
    update_fiber(i, d.first, dt, d.pos_prev, d.pos, d.ori_prev, d.ori, d.tv_prev, d.tv, d.rv_prev, d.rv, d.pos_next,
                 d.ori_next);
}

hipError_t launch_update_batch(int nfblocks, const SusDev* sus, const int2* fblocks, double dt, hipStream_t s) {
  if (nfblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(update_batch_kernel, dim3(nfblocks), dim3(256), 0, s, sus, fblocks, dt);
  return hipGetLastError();
}

template <int N, int NQI>
static hipError_t launch_vel_n(bool an, int M, const double* pos, const double* ori, const double* x, float* wf,
                               double* fm, double* vpart, double* tvel, double* rvel, const PhysConsts& pc,
                               hipStream_t s) {
  constexpr int PAIRS = kPairThreads / (3 * NQI);
  hipLaunchKernelGGL((force_density_kernel<N, NQI>), dim3((M + 63) / 64), dim3(64), 0, s, M, x, pc, wf, fm);
  const PairGrid g = pair_grid(M, PAIRS);
  if (an)
    hipLaunchKernelGGL((velocity_kernel<N, NQI, true>), dim3(g.tiles, g.groups), dim3(kPairThreads), 0, s, M, pos,
                       ori, wf, fm, vpart, g.jgroup, pc);
  else
    hipLaunchKernelGGL((velocity_kernel<N, NQI, false>), dim3(g.tiles, g.groups), dim3(kPairThreads), 0, s, M, pos,
                       ori, wf, fm, vpart, g.jgroup, pc);
  hipLaunchKernelGGL(velocity_finalize_kernel, dim3((M + 127) / 128), dim3(128), 0, s, M, g.groups, ori, vpart, pc,
                     tvel, rvel);
  return hipGetLastError();
}

#define RF_VEL_CASE(NN)                                                                              \
  case NN:                                                                                           \
    if (NQI == 4) return launch_vel_n<NN, 4>(analytical, M, pos, ori, x, wf, fm, vpart, tv, rv, pc, s); \
    if (NQI == 8) return launch_vel_n<NN, 8>(analytical, M, pos, ori, x, wf, fm, vpart, tv, rv, pc, s); \
    if (NQI == 16) return launch_vel_n<NN, 16>(analytical, M, pos, ori, x, wf, fm, vpart, tv, rv, pc, s); \
    return hipErrorInvalidValue;

hipError_t launch_velocities(int N, int NQI, bool analytical, int M, const double* pos, const double* ori,
                             const double* x, float* wf, double* fm, double* vpart, double* tv, double* rv,
                             const PhysConsts& pc, hipStream_t s) {
  switch (N) {
    RF_VEL_CASE(1)
    RF_VEL_CASE(2)
    RF_VEL_CASE(3)
    RF_VEL_CASE(4)
    RF_VEL_CASE(5)
    RF_VEL_CASE(6)
    RF_VEL_CASE(7)
    RF_VEL_CASE(8)
    default:
      return hipErrorInvalidValue;
  }
}
#undef RF_VEL_CASE

hipError_t launch_update(bool first, int M, double dt, const double* pos_prev, const double* pos,
                         const double* ori_prev, const double* ori, const double* tv_prev, const double* tv,
                         const double* rv_prev, const double* rv, double* pos_next, double* ori_next,
                         hipStream_t s) {
  hipLaunchKernelGGL(update_kernel, dim3((M + 127) / 128), dim3(128), 0, s, first ? 1 : 0, M, dt, pos_prev, pos,
                     ori_prev, ori, tv_prev, tv, rv_prev, rv, pos_next, ori_next);
  return hipGetLastError();
}

}  // namespace rf
