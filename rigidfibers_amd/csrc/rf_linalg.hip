// rf_linalg.hip -- HBM-streaming GEMV and the vector kernels of the Krylov
// solvers.
//
// Reference: the GMRES/BiCGStab matvec is ViennaCL's vec_mul_col_kernel
// (lib/viennacl/linalg/cuda/matrix_operations_col.hpp:1251-1278, one thread per
// row over a column-major FP32 matrix) and sgemv in fortran/solve_system.incl:36.
//
// gemv: A is column-major FP32 with a leading dimension lda that is a multiple
// of 64, so every column starts 256-byte aligned.  A workgroup owns 1024 rows
// (4 per lane, one 16-byte load per lane per column: a wave reads 1 KiB
// contiguous per column) and a chunk of columns; x is read with scalar loads.
// Products accumulate in FP64 (the FP64 VALU rate is far above what 8 TB/s of
// FP32 matrix data can feed), chunks are summed afterwards in a fixed order.
#include "rf_internal.h"

namespace rf {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kGemvThreads = 256;
constexpr int kGemvRows = 4 * kGemvThreads;

__global__ __launch_bounds__(kGemvThreads) void gemv_f32_kernel(int64_t n, int64_t lda, const float* __restrict__ A,
                                                                const double* __restrict__ x,
                                                                double* __restrict__ part, int cols_per_chunk,
                                                                const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kGemvRows + 4 * threadIdx.x;
  const int64_t c0 = static_cast<int64_t>(blockIdx.y) * cols_per_chunk;
  int64_t c1 = c0 + cols_per_chunk;
  if (c1 > n) c1 = n;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (r0 < n) {
    // rows r0..r0+3 < lda (lda % 4 == 0), so the 16-byte load stays inside the column
    const float* p = A + c0 * lda + r0;
    int64_t c = c0;
    for (; c + 4 <= c1; c += 4) {
      const f4 v0 = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
      const f4 v1 = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p + lda));
      const f4 v2 = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p + 2 * lda));
      const f4 v3 = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p + 3 * lda));
      const double x0 = x[c], x1 = x[c + 1], x2 = x[c + 2], x3 = x[c + 3];
      a0 = fma(static_cast<double>(v0.x), x0, a0);
      a1 = fma(static_cast<double>(v0.y), x0, a1);
      a2 = fma(static_cast<double>(v0.z), x0, a2);
      a3 = fma(static_cast<double>(v0.w), x0, a3);
      a0 = fma(static_cast<double>(v1.x), x1, a0);
      a1 = fma(static_cast<double>(v1.y), x1, a1);
      a2 = fma(static_cast<double>(v1.z), x1, a2);
      a3 = fma(static_cast<double>(v1.w), x1, a3);
      a0 = fma(static_cast<double>(v2.x), x2, a0);
      a1 = fma(static_cast<double>(v2.y), x2, a1);
      a2 = fma(static_cast<double>(v2.z), x2, a2);
      a3 = fma(static_cast<double>(v2.w), x2, a3);
      a0 = fma(static_cast<double>(v3.x), x3, a0);
      a1 = fma(static_cast<double>(v3.y), x3, a1);
      a2 = fma(static_cast<double>(v3.z), x3, a2);
      a3 = fma(static_cast<double>(v3.w), x3, a3);
      p += 4 * lda;
    }
    for (; c < c1; ++c) {
      const f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
      const double xc = x[c];
      a0 = fma(static_cast<double>(v.x), xc, a0);
      a1 = fma(static_cast<double>(v.y), xc, a1);
      a2 = fma(static_cast<double>(v.z), xc, a2);
      a3 = fma(static_cast<double>(v.w), xc, a3);
      p += lda;
    }
    double* out = part + static_cast<int64_t>(blockIdx.y) * n + r0;
    if (r0 + 3 < n) {
      out[0] = a0;
      out[1] = a1;
      out[2] = a2;
      out[3] = a3;
    } else {
      out[0] = a0;
      if (r0 + 1 < n) out[1] = a1;
      if (r0 + 2 < n) out[2] = a2;
    }
  }
}

// y[r] = sum_k part[k][r]  (mode 0)   or   y[r] = b[r] - sum_k part[k][r]  (mode 1)
__global__ void gemv_reduce_kernel(int64_t n, int chunks, const double* __restrict__ part, double* __restrict__ y,
                                   const double* __restrict__ b, int mode, const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int k = 0; k < chunks; ++k) s += part[static_cast<int64_t>(k) * n + r];
  y[r] = (mode == 1) ? b[r] - s : s;
}

GemvPlan gemv_plan(int64_t n, int64_t lda) {
  GemvPlan p;
  p.n = n;
  p.lda = lda;
  p.row_tiles = static_cast<int>((n + kGemvRows - 1) / kGemvRows);
  // ~2048 workgroups over 256 CUs; at least 64 columns per chunk.
  int64_t chunks = (2048 + p.row_tiles - 1) / p.row_tiles;
  const int64_t max_chunks = (n + 63) / 64;
  if (chunks > max_chunks) chunks = max_chunks;
  if (chunks < 1) chunks = 1;
  p.cols_per_chunk = static_cast<int>((n + chunks - 1) / chunks);
  p.chunks = static_cast<int>((n + p.cols_per_chunk - 1) / p.cols_per_chunk);
  return p;
}

size_t gemv_partial_elems(const GemvPlan& p) { return static_cast<size_t>(p.chunks) * static_cast<size_t>(p.n); }

hipError_t gemv(const GemvPlan& p, const float* A, const double* x, double* part, double* y, const double* b,
                int mode, hipStream_t s, hipEvent_t ev_begin, hipEvent_t ev_end, const int* skip) {
  if (ev_begin) (void)hipEventRecord(ev_begin, s);
  hipLaunchKernelGGL(gemv_f32_kernel, dim3(p.row_tiles, p.chunks), dim3(kGemvThreads), 0, s, p.n, p.lda, A, x, part,
                     p.cols_per_chunk, skip);
  if (ev_end) (void)hipEventRecord(ev_end, s);
  hipLaunchKernelGGL(gemv_reduce_kernel, dim3(static_cast<unsigned>((p.n + 255) / 256)), dim3(256), 0, s, p.n,
                     p.chunks, part, y, b, mode, skip);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// y = A x with A in theta form (see rf_assemble.hip): for target fiber i
//   s0    = sum_{j,k_j} Q0_{ij,k_j} . x_{j,k_j}
//   U_k   = sum_{j,k_j} Theta_{ij,k,k_j} x_{j,k_j}                  (k >= 1)
//   y_i,0 = x_i,0 + D1 t_i s0
//   y_i,k = x_i,k + gamma_k (U_k - eigen_k t_i (t_i . U_k))
// using (Theta t_i)^T x = t_i . (Theta x) for the symmetric Theta.  One
// workgroup per target fiber streams its NC component arrays (each lane one
// (j, k_j) entry per sweep: every load instruction reads 256 contiguous
// bytes); FP64 accumulation, workgroup reduction, 15 outputs.

template <int N>
__device__ __forceinline__ void gemv_theta_block(int i, int M, int64_t Lp, const float* __restrict__ T,
                                                 const double* __restrict__ ori, const double* __restrict__ x,
                                                 double* __restrict__ y, const double* __restrict__ b, int mode,
                                                 const PhysConsts& pc) {
  constexpr int NC = 3 + 6 * (N - 1);
  constexpr int NACC = 1 + 3 * (N - 1);
  __shared__ double red[256 / 64][NACC];
  const float* Ti = T + static_cast<int64_t>(i) * NC * Lp;
  const int L = M * N;
  double acc[NACC];
#pragma unroll
  for (int a = 0; a < NACC; ++a) acc[a] = 0.0;
  for (int e = threadIdx.x; e < L; e += 256) {
    const double x0 = x[3 * static_cast<int64_t>(e)], x1 = x[3 * static_cast<int64_t>(e) + 1],
                 x2 = x[3 * static_cast<int64_t>(e) + 2];
    const float* p = Ti + e;
    const float q0 = __builtin_nontemporal_load(p), q1 = __builtin_nontemporal_load(p + Lp),
                q2 = __builtin_nontemporal_load(p + 2 * Lp);
    acc[0] = fma(static_cast<double>(q0), x0, fma(static_cast<double>(q1), x1, fma(static_cast<double>(q2), x2, acc[0])));
#pragma unroll
    for (int k = 1; k < N; ++k) {
      const float* pk = p + (3 + 6 * (k - 1)) * Lp;
      const double txx = __builtin_nontemporal_load(pk), tyy = __builtin_nontemporal_load(pk + Lp),
                   tzz = __builtin_nontemporal_load(pk + 2 * Lp), txy = __builtin_nontemporal_load(pk + 3 * Lp),
                   txz = __builtin_nontemporal_load(pk + 4 * Lp), tyz = __builtin_nontemporal_load(pk + 5 * Lp);
      double* u = acc + 1 + 3 * (k - 1);
      u[0] = fma(txx, x0, fma(txy, x1, fma(txz, x2, u[0])));
      u[1] = fma(txy, x0, fma(tyy, x1, fma(tyz, x2, u[1])));
      u[2] = fma(txz, x0, fma(tyz, x1, fma(tzz, x2, u[2])));
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NACC; ++a) {
    double v = acc[a];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][a] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3 * N) {
    const int r = threadIdx.x, k = r / 3, a = r - 3 * k;
    const double t[3] = {ori[3 * i], ori[3 * i + 1], ori[3 * i + 2]};
    double out;
    if (k == 0) {
      const double s0 = red[0][0] + red[1][0] + red[2][0] + red[3][0];
      out = static_cast<double>(pc.D1) * t[a] * s0;
    } else {
      double U[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int idx = 1 + 3 * (k - 1) + c;
        U[c] = red[0][idx] + red[1][idx] + red[2][idx] + red[3][idx];
      }
      const double tU = t[0] * U[0] + t[1] * U[1] + t[2] * U[2];
      out = static_cast<double>(pc.gamma[k]) * (U[a] - static_cast<double>(pc.eigen[k]) * t[a] * tU);
    }
    const int64_t row = static_cast<int64_t>(3 * N) * i + r;
    out += x[row];  // identity diagonal block
    y[row] = (mode == 1) ? b[row] - out : out;
  }
}

template <int N>
__global__ __launch_bounds__(256) void gemv_theta_kernel(int M, int64_t Lp, const float* __restrict__ T,
                                                         const double* __restrict__ ori,
                                                         const double* __restrict__ x, double* __restrict__ y,
                                                         const double* __restrict__ b, int mode, PhysConsts pc,
                                                         const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  gemv_theta_block<N>(blockIdx.x, M, Lp, T, ori, x, y, b, mode, pc);
}

// Serving: workgroup = (suspension, target fiber) of the suspensions with a
// product in this iteration.
template <int N>
__global__ __launch_bounds__(256) void gemv_theta_batch_kernel(const SusDev* __restrict__ sus,
                                                               const SusIter* __restrict__ it,
                                                               const int2* __restrict__ targets, PhysConsts pc) {
  const int2 e = targets[blockIdx.x];
  const SusIter& c = it[e.x];
  if (c.gx == nullptr) return;
  const SusDev& d = sus[e.x];
  gemv_theta_block<N>(e.y, d.M, d.Lp, d.T, d.ori, c.gx, c.gy, d.b, c.gmode, pc);
}

hipError_t launch_gemv_theta_batch(int N, int ntargets, const SusDev* sus, const SusIter* it, const int2* targets,
                                   const PhysConsts& pc, hipStream_t s) {
  if (ntargets <= 0) return hipSuccess;
  switch (N) {
#define RF_GB(NN)                                                                                            \
  case NN:                                                                                                   \
    hipLaunchKernelGGL(gemv_theta_batch_kernel<NN>, dim3(ntargets), dim3(256), 0, s, sus, it, targets, pc); \
    break;
    RF_GB(1)
    RF_GB(2)
    RF_GB(3)
    RF_GB(4)
    RF_GB(5)
    RF_GB(6)
    RF_GB(7)
    RF_GB(8)
#undef RF_GB
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t gemv_theta(int N, int M, int64_t Lp, const float* T, const double* ori, const double* x, double* y,
                      const double* b, int mode, const PhysConsts& pc, hipStream_t s, hipEvent_t ev_begin,
                      hipEvent_t ev_end, const int* skip) {
  if (ev_begin) (void)hipEventRecord(ev_begin, s);
  switch (N) {
#define RF_GT(NN)                                                                                               \
  case NN:                                                                                                      \
    hipLaunchKernelGGL(gemv_theta_kernel<NN>, dim3(M), dim3(256), 0, s, M, Lp, T, ori, x, y, b, mode, pc, skip); \
    break;
    RF_GT(1)
    RF_GT(2)
    RF_GT(3)
    RF_GT(4)
    RF_GT(5)
    RF_GT(6)
    RF_GT(7)
    RF_GT(8)
#undef RF_GT
    default:
      return hipErrorInvalidValue;
  }
  if (ev_end) (void)hipEventRecord(ev_end, s);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Multi-dot: out[k] = <V_k, w>, k < K.  Grid (blocks, K); block (b, k) reduces
// rows [b*rpb, (b+1)*rpb) of vector k; a second kernel sums the blocks.

constexpr int kDotThreads = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) t += sh[w];
  return t;
}

int mdot_blocks(int64_t n) {
  int64_t b = (n + 2047) / 2048;
  if (b > 256) b = 256;
  if (b < 1) b = 1;
  return static_cast<int>(b);
}

__global__ __launch_bounds__(kDotThreads) void mdot_partial_kernel(int64_t n, int64_t rpb,
                                                                   const double* __restrict__ V, int64_t ldv,
                                                                   const double* __restrict__ w,
                                                                   double* __restrict__ part,
                                                                   const int* __restrict__ skip) {
  __shared__ double sh[kDotThreads / 64];
  if (skip != nullptr && *skip) return;
  const int k = blockIdx.y;
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * rpb;
  int64_t hi = lo + rpb;
  if (hi > n) hi = n;
  const double* v = V + static_cast<int64_t>(k) * ldv;
  double s = 0.0;
  for (int64_t r = lo + threadIdx.x; r < hi; r += kDotThreads) s = fma(v[r], w[r], s);
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[static_cast<int64_t>(k) * gridDim.x + blockIdx.x] = t;
}

__global__ void mdot_final_kernel(int nb, const double* __restrict__ part, double* __restrict__ out) {
  const int k = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 64) s += part[static_cast<int64_t>(k) * nb + b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (threadIdx.x == 0) out[k] = s;
}

hipError_t mdot(int64_t n, int K, const double* V, int64_t ldv, const double* w, double* part, double* out,
                hipStream_t s) {
  if (K <= 0) return hipSuccess;
  const int nb = mdot_blocks(n);
  const int64_t rpb = (n + nb - 1) / nb;
  hipLaunchKernelGGL(mdot_partial_kernel, dim3(nb, K), dim3(kDotThreads), 0, s, n, rpb, V, ldv, w, part,
                     static_cast<const int*>(nullptr));
  hipLaunchKernelGGL(mdot_final_kernel, dim3(K), dim3(64), 0, s, nb, part, out);
  return hipGetLastError();
}

__global__ void axpby_kernel(int64_t n, double alpha, const double* __restrict__ x, double beta,
                             double* __restrict__ y) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  y[r] = (beta == 0.0) ? alpha * x[r] : fma(alpha, x[r], beta * y[r]);
}

hipError_t axpby(int64_t n, double alpha, const double* x, double beta, double* y, hipStream_t s) {
  hipLaunchKernelGGL(axpby_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, n, alpha, x, beta,
                     y);
  return hipGetLastError();
}

__global__ void fill_kernel(int64_t n, double v, double* __restrict__ y) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < n) y[r] = v;
}

hipError_t fill(int64_t n, double value, double* y, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, n, value, y);
  return hipGetLastError();
}

__global__ void nonfinite_kernel(int64_t n, const double* __restrict__ x, int* __restrict__ flag) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < n && !isfinite(x[r])) flag[0] = 1;  // benign race: any writer stores 1
}

hipError_t nonfinite_count(int64_t n, const double* x, int* flag_dev, hipStream_t s) {
  hipLaunchKernelGGL(nonfinite_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, n, x,
                     flag_dev);
  return hipGetLastError();
}

// Vector work of one lockstep GMRES iteration over 256-element blocks of
// every suspension's vectors (entry (suspension, block) of the table).
constexpr int kVecBlock = 256;

__global__ __launch_bounds__(kVecBlock) void iter_prep_batch_kernel(const SusDev* __restrict__ sus,
                                                                    const SusIter* __restrict__ it,
                                                                    const int2* __restrict__ vblocks) {
  const int2 e = vblocks[blockIdx.x];
  const SusDev& d = sus[e.x];
  const SusIter& c = it[e.x];
  const int r = e.y * kVecBlock + threadIdx.x;
  if (r >= d.n) return;
  if (c.zero_x) d.x[r] = 0.0;
  if (c.ycount > 0) {
    double s = 0.0;
    for (int k = 0; k < c.ycount; ++k) s = fma(d.V[static_cast<int64_t>(k) * d.n + r], c.ycoef[k], s);
    d.x[r] += s;
  }
  if (c.svec) c.svec[r] *= c.sval;
}

// Block partial sums: phase 0/1 <V_j, w> for j < K (blockIdx.y = j), phase 2 <w, w>.
__global__ __launch_bounds__(kVecBlock) void dots_batch_kernel(int phase, int nvb, const SusDev* __restrict__ sus,
                                                               const SusIter* __restrict__ it,
                                                               const int2* __restrict__ vblocks,
                                                               double* __restrict__ part) {
  __shared__ double sh[kVecBlock / 64];
  const int2 e = vblocks[blockIdx.x];
  const SusDev& d = sus[e.x];
  const SusIter& c = it[e.x];
  const int j = blockIdx.y;
  if (c.w == nullptr || (phase < 2 && j >= c.K) || (phase == 2 && j > 0)) return;
  const int r = e.y * kVecBlock + threadIdx.x;
  double v = 0.0;
  if (r < d.n) {
    const double wr = c.w[r];
    v = (phase < 2) ? d.V[static_cast<int64_t>(j) * d.n + r] * wr : wr * wr;
  }
  const double t = block_sum(v, sh);
  if (threadIdx.x == 0) part[static_cast<int64_t>(j) * nvb + blockIdx.x] = t;
}



// Sum of a suspension's block partials of row j, in block order (the same
// order wherever it is formed, so every consumer sees the same value).
__device__ __forceinline__ double sum_block_parts(const double* __restrict__ part, int nvb, int j, int first,
                                                  int count) {
  const double* p = part + static_cast<int64_t>(j) * nvb + first;
  double v = 0.0;
  for (int b = 0; b < count; ++b) v += p[b];
  return v;
}

// Gram-Schmidt pass `phase` (0 or 1) on one 256-row block of one suspension:
// h_j (j < K) from the pass's block partials, w -= sum_j h_j V_j on the
// block's rows, then the block partials of the next reduction on the same
// rows: <V_j, w> after pass 0, <w, w> after pass 1.  Partials of reduction p
// live in part + p kdim nvb, row j at j nvb + block.
__global__ __launch_bounds__(kVecBlock) void orth_dots_batch_kernel(int phase, int nvb, int kdim,
                                                                    const SusDev* __restrict__ sus,
                                                                    const SusIter* __restrict__ it,
                                                                    const int2* __restrict__ vblocks,
                                                                    double* __restrict__ part) {
  extern __shared__ double dsh[];  // h[kdim], then the waves' partials [kVecBlock / 64][kdim]
  double* hsh = dsh;
  double* wsh = dsh + kdim;
  const int2 e = vblocks[blockIdx.x];
  const SusDev& d = sus[e.x];
  const SusIter& c = it[e.x];
  if (c.w == nullptr) return;  // the whole block belongs to one suspension
  const int K = c.K;
  const int64_t region = static_cast<int64_t>(kdim) * nvb;
  for (int j = threadIdx.x; j < K; j += kVecBlock)
    hsh[j] = sum_block_parts(part + phase * region, nvb, j, d.vb_first, d.vb_count);
  __syncthreads();
  const int r = e.y * kVecBlock + threadIdx.x;
  const bool in = r < d.n;
  double wr = 0.0;
  if (in) {
    double sacc = 0.0;
    for (int k = 0; k < K; ++k) sacc = fma(d.V[static_cast<int64_t>(k) * d.n + r], hsh[k], sacc);
    wr = c.w[r] - sacc;
    c.w[r] = wr;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int J = phase == 0 ? K : 1;
  for (int j = 0; j < J; ++j) {
    double v = in ? (phase == 0 ? d.V[static_cast<int64_t>(j) * d.n + r] * wr : wr * wr) : 0.0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) wsh[wv * kdim + j] = v;
  }
  __syncthreads();
  double* pout = part + (phase + 1) * region;
  for (int j = threadIdx.x; j < J; j += kVecBlock) {
    double t = 0.0;
    for (int q = 0; q < kVecBlock / 64; ++q) t += wsh[q * kdim + j];
    pout[static_cast<int64_t>(j) * nvb + blockIdx.x] = t;
  }
}

hipError_t launch_orth_dots_batch(int nvb, int kdim, int phase, const SusDev* sus, const SusIter* it,
                                  const int2* vblocks, double* part, hipStream_t s) {
  if (nvb <= 0) return hipSuccess;
  const size_t shmem = static_cast<size_t>(kdim) * (1 + kVecBlock / 64) * sizeof(double);
  hipLaunchKernelGGL(orth_dots_batch_kernel, dim3(nvb), dim3(kVecBlock), shmem, s, phase, nvb, kdim, sus, it,
                     vblocks, part);
  return hipGetLastError();
}

// Control step of the serving lockstep (one workgroup per suspension), the
// same GMRES(m) as the single-simulation solver: true residual at every
// restart, classical Gram-Schmidt twice (the two passes' coefficients summed),
// Givens rotations, stop on ||r|| <= tol ||b||, a cycle whose estimate meets
// it ending the solve with its update of x (confirm: with the true residual
// of that x).  The lanes form the iteration's reductions from the block
// partials (at the start step, the norms of b and of the first residual from
// the assembly's per-target partials), lane 0 advances the state and writes
// the next iteration's operations into it[q] and whether it has any into
// active[q].
__global__ __launch_bounds__(64) void gmres_control_batch_kernel(int m, int kdim, int nvb, double tol, int maxit,
                                                                 int confirm, const SusDev* __restrict__ sus,
                                                                 SusIter* __restrict__ it, SusCtl* __restrict__ ctl,
                                                                 const double* __restrict__ part,
                                                                 int* __restrict__ active) {
  extern __shared__ double hsh[];  // h0[K], h1[K], ||w||^2
  __shared__ double nsh[2][64];     // start: the lanes' sums of the per-target norm partials
  const int q = blockIdx.x;
  const SusDev& d = sus[q];
  SusIter& c = it[q];
  SusCtl& C = ctl[q];
  const int K = c.K;
  const int64_t region = static_cast<int64_t>(kdim) * nvb;
  const bool start = C.phase == 3;
  if (c.w != nullptr)
    for (int t = threadIdx.x; t <= 2 * K; t += 64) {
      const int p = t < K ? 0 : (t < 2 * K ? 1 : 2);
      const int j = t < K ? t : (t < 2 * K ? t - K : 0);
      hsh[t] = sum_block_parts(part + p * region, nvb, j, d.vb_first, d.vb_count);
    }
  if (start) {  // lane l sums targets l, l + 64, ... (a fixed order)
    double b2 = 0.0, r2 = 0.0;
    for (int i = threadIdx.x; i < d.M; i += 64) {
      b2 += d.npart[2 * i];
      r2 += d.npart[2 * i + 1];
    }
    nsh[0][threadIdx.x] = b2;
    nsh[1][threadIdx.x] = r2;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double* h0 = hsh;
  const double* h1 = hsh + K;
  const double w2 = (c.w != nullptr) ? hsh[2 * K] : 0.0;
  int zero_x = 0, ycount = 0;
  double* svec = nullptr;
  double sval = 1.0;
  if (start) {  // V_0 = b - A x0 (b when cold) from the assembly
    double b2 = 0.0, r2 = 0.0;
    for (int l = 0; l < 64; ++l) {
      b2 += nsh[0][l];
      r2 += nsh[1][l];
    }
    C.bnorm = sqrt(b2);
    const double beta = sqrt(r2);
    C.res = beta;
    if (!isfinite(C.bnorm) || !isfinite(beta)) {
      C.nonfinite = 1;
      C.phase = 2;
    } else if (C.bnorm == 0.0) {  // x = 0 is the solution
      zero_x = 1;
      C.converged = 1;
      C.phase = 2;
    } else {
      zero_x = C.cold;
      if (beta / C.bnorm <= tol) {
        C.converged = 1;
        C.phase = 2;
      } else if (C.it >= maxit) {
        C.phase = 2;
      } else {
        svec = d.V;
        sval = 1.0 / beta;
        for (int i = 0; i <= m; ++i) C.g[i] = 0.0;
        C.g[0] = beta;
        C.k = 0;
        C.phase = 1;
      }
    }
  } else if (C.phase == 0) {  // w = b - A x
    const double beta = sqrt(w2);
    C.res = beta;
    if (!isfinite(beta)) {
      C.nonfinite = 1;
      C.phase = 2;
    } else if (beta / C.bnorm <= tol) {
      C.converged = 1;
      C.phase = 2;
    } else if (C.it >= maxit) {
      C.phase = 2;
    } else {
      svec = d.V;
      sval = 1.0 / beta;
      for (int i = 0; i <= m; ++i) C.g[i] = 0.0;
      C.g[0] = beta;
      C.k = 0;
      C.phase = 1;
    }
  } else if (C.phase == 1) {  // w = A V_k orthogonalised twice against V_0..V_k
    const int k = C.k;
    ++C.it;
    const double hn = sqrt(w2);
    if (!isfinite(hn)) {
      C.nonfinite = 1;
      C.phase = 2;
    } else {
      double* Hk = C.H + static_cast<int64_t>(k) * (m + 1);
      for (int i = 0; i <= k; ++i) Hk[i] = h0[i] + h1[i];
      Hk[k + 1] = hn;
      for (int i = 0; i < k; ++i) {
        const double a0 = Hk[i], a1 = Hk[i + 1];
        Hk[i] = C.cs[i] * a0 + C.sn[i] * a1;
        Hk[i + 1] = -C.sn[i] * a0 + C.cs[i] * a1;
      }
      const double a0 = Hk[k], a1 = Hk[k + 1];
      const double rr = hypot(a0, a1);
      C.cs[k] = (rr == 0.0) ? 1.0 : a0 / rr;
      C.sn[k] = (rr == 0.0) ? 0.0 : a1 / rr;
      Hk[k] = rr;
      Hk[k + 1] = 0.0;
      C.g[k + 1] = -C.sn[k] * C.g[k];
      C.g[k] = C.cs[k] * C.g[k];
      C.res = fabs(C.g[k + 1]);
      if (C.res / C.bnorm <= tol || hn == 0.0 || k + 1 == m || C.it >= maxit) {
        const int ku = k + 1;  // y = H^-1 g, then x += V y and the true residual
        for (int i = ku - 1; i >= 0; --i) {
          double acc = C.g[i];
          for (int j = i + 1; j < ku; ++j) acc -= C.H[i + static_cast<int64_t>(j) * (m + 1)] * C.y[j];
          C.y[i] = acc / C.H[i + static_cast<int64_t>(i) * (m + 1)];
        }
        ycount = ku;
        if (!confirm && C.res / C.bnorm <= tol) {  // the update of x is the last operation
          C.converged = 1;
          C.phase = 2;
        } else {
          C.phase = 0;
        }
      } else {
        svec = d.V + static_cast<int64_t>(k + 1) * d.n;
        sval = 1.0 / hn;
        C.k = k + 1;
      }
    }
  }
  SusIter nx{};
  nx.zero_x = zero_x;
  nx.ycount = ycount;
  nx.ycoef = C.y;
  nx.svec = svec;
  nx.sval = sval;
  if (C.phase == 0) {
    nx.gx = d.x;
    nx.gy = d.V;
    nx.gmode = 1;
    nx.w = d.V;
  } else if (C.phase == 1) {
    nx.gx = d.V + static_cast<int64_t>(C.k) * d.n;
    nx.gy = d.V + static_cast<int64_t>(C.k + 1) * d.n;
    nx.w = nx.gy;
    nx.K = C.k + 1;
  }
  c = nx;
  active[q] = (C.phase != 2 || zero_x || ycount > 0 || svec != nullptr) ? 1 : 0;
}

hipError_t launch_gmres_control_batch(int ns, int m, int kdim, int nvb, double tol, int maxit, int confirm,
                                      const SusDev* sus, SusIter* it, SusCtl* ctl, const double* part, int* active,
                                      hipStream_t s) {
  if (ns <= 0) return hipSuccess;
  hipLaunchKernelGGL(gmres_control_batch_kernel, dim3(ns), dim3(64), (2 * kdim + 1) * sizeof(double), s, m, kdim,
                     nvb, tol, maxit, confirm, sus, it, ctl, part, active);
  return hipGetLastError();
}

hipError_t launch_iter_prep_batch(int nvb, const SusDev* sus, const SusIter* it, const int2* vblocks,
                                  hipStream_t s) {
  if (nvb <= 0) return hipSuccess;
  hipLaunchKernelGGL(iter_prep_batch_kernel, dim3(nvb), dim3(kVecBlock), 0, s, sus, it, vblocks);
  return hipGetLastError();
}


hipError_t launch_dots_partial_batch(int nvb, int kmax, const SusDev* sus, const SusIter* it, const int2* vblocks,
                                    double* part, hipStream_t s) {
  if (nvb <= 0 || kmax <= 0) return hipSuccess;
  hipLaunchKernelGGL(dots_batch_kernel, dim3(nvb, kmax), dim3(kVecBlock), 0, s, 0, nvb, sus, it, vblocks, part);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// FP16 storage of the theta form (GMRES-IR inner solves), written by the
// theta assembly.  Every (target i, source j) block -- the 27 x N numbers of
// a 3N x 3N block of A -- shares one FP32 scale S[i][j], a power of two at
// or above max|T| / 32768 (from a bound on the block's inner integrals), so
// the stored values use the FP16 range whatever the block's distance (far
// blocks would otherwise sink into FP16's subnormals).  Same component
// layout as T.

// y = A16 x (mode 0) or b - A16 x (mode 1): the theta GEMV on the FP16 copy.
// A workgroup takes kG16Targets consecutive targets, so that the x entries it
// streams serve all of them (alone, x's 24 bytes per entry were as many bytes
// as the entry's 27 FP16 values: 6.38 -> 6.79 TB/s at M = 5000 in a copy of
// the kernel); each lane takes two consecutive entries (one 4-byte load per
// component, a wave reads 256 contiguous bytes) and folds each target's block
// scale into x.  Each target's sums see the same operations in the same order
// as with one target per workgroup.
constexpr int kG16Targets = 2;
template <int N>
__global__ __launch_bounds__(256) void gemv_theta16_kernel(int M, int64_t Lp, const _Float16* __restrict__ T16,
                                                           const float* __restrict__ S,
                                                           const double* __restrict__ ori,
                                                           const double* __restrict__ x, double* __restrict__ y,
                                                           const double* __restrict__ b, int mode, PhysConsts pc,
                                                           const int* __restrict__ skip) {
  typedef _Float16 h2 __attribute__((ext_vector_type(2)));
  constexpr int TT = kG16Targets;
  if (skip != nullptr && *skip) return;
  constexpr int NC = 3 + 6 * (N - 1);
  constexpr int NACC = 1 + 3 * (N - 1);
  __shared__ double red[256 / 64][TT][NACC];
  const int i0 = TT * static_cast<int>(blockIdx.x);
  const int nt = min(TT, M - i0);  // targets of this workgroup (uniform)
  const int L = M * N;
  double acc[TT][NACC];
#pragma unroll
  for (int t = 0; t < TT; ++t)
#pragma unroll
    for (int a = 0; a < NACC; ++a) acc[t][a] = 0.0;
  // entry of scaled x (xa, xb, xc) and FP16 values hv into target t's sums
  auto entry = [&](double* ac, const double hv[NC], double xa, double xb, double xc) {
    ac[0] = fma(hv[0], xa, fma(hv[1], xb, fma(hv[2], xc, ac[0])));
#pragma unroll
    for (int k = 1; k < N; ++k) {
      const double* t = hv + 3 + 6 * (k - 1);  // xx yy zz xy xz yz
      double* u = ac + 1 + 3 * (k - 1);
      u[0] = fma(t[0], xa, fma(t[3], xb, fma(t[4], xc, u[0])));
      u[1] = fma(t[3], xa, fma(t[1], xb, fma(t[5], xc, u[1])));
      u[2] = fma(t[4], xa, fma(t[5], xb, fma(t[2], xc, u[2])));
    }
  };
  const int Lpair = L & ~1;
  for (int e = 2 * threadIdx.x; e < Lpair; e += 512) {
    double xv[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) xv[q] = x[3 * static_cast<int64_t>(e) + q];
#pragma unroll
    for (int t = 0; t < TT; ++t) {
      if (t >= nt) break;
      const _Float16* Hi = T16 + static_cast<int64_t>(i0 + t) * NC * Lp;
      const float* Si = S + static_cast<int64_t>(i0 + t) * M;
      h2 v[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) v[c] = __builtin_nontemporal_load(reinterpret_cast<const h2*>(Hi + c * Lp + e));
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        double hv[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) hv[c] = static_cast<double>(static_cast<float>(v[c][q]));
        const double sj = static_cast<double>(Si[(e + q) / N]);
        entry(acc[t], hv, sj * xv[3 * q], sj * xv[3 * q + 1], sj * xv[3 * q + 2]);
      }
    }
  }
  if ((L & 1) && threadIdx.x == 0) {
    const int e = L - 1;
    for (int t = 0; t < nt; ++t) {
      const _Float16* Hi = T16 + static_cast<int64_t>(i0 + t) * NC * Lp;
      double hv[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) hv[c] = static_cast<double>(static_cast<float>(Hi[c * Lp + e]));
      const double sj = static_cast<double>(S[static_cast<int64_t>(i0 + t) * M + e / N]);
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
        if (tt == t)
          entry(acc[tt], hv, sj * x[3 * static_cast<int64_t>(e)], sj * x[3 * static_cast<int64_t>(e) + 1],
                sj * x[3 * static_cast<int64_t>(e) + 2]);
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < TT; ++t)
#pragma unroll
    for (int a = 0; a < NACC; ++a) {
      double v = acc[t][a];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) red[wave][t][a] = v;
    }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < nt * 3 * N) {
    const int t = threadIdx.x / (3 * N), r = threadIdx.x - t * 3 * N, k = r / 3, a = r - 3 * k;
    const int i = i0 + t;
    const double tv[3] = {ori[3 * i], ori[3 * i + 1], ori[3 * i + 2]};
    double out;
    if (k == 0) {
      const double s0 = red[0][t][0] + red[1][t][0] + red[2][t][0] + red[3][t][0];
      out = static_cast<double>(pc.D1) * tv[a] * s0;
    } else {
      double U[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int idx = 1 + 3 * (k - 1) + c;
        U[c] = red[0][t][idx] + red[1][t][idx] + red[2][t][idx] + red[3][t][idx];
      }
      const double tU = tv[0] * U[0] + tv[1] * U[1] + tv[2] * U[2];
      out = static_cast<double>(pc.gamma[k]) * (U[a] - static_cast<double>(pc.eigen[k]) * tv[a] * tU);
    }
    const int64_t row = static_cast<int64_t>(3 * N) * i + r;
    out += x[row];  // identity diagonal block
    y[row] = (mode == 1) ? b[row] - out : out;
  }
}

hipError_t gemv_theta16(int N, int M, int64_t Lp, const _Float16* T16, const float* S, const double* ori,
                        const double* x, double* y, const double* b, int mode, const PhysConsts& pc, hipStream_t s,
                        hipEvent_t ev_begin, hipEvent_t ev_end, const int* skip) {
  if (ev_begin) (void)hipEventRecord(ev_begin, s);
  switch (N) {
#define RF_G16(NN)                                                                                            \
  case NN:                                                                                                    \
    hipLaunchKernelGGL(gemv_theta16_kernel<NN>, dim3((M + kG16Targets - 1) / kG16Targets), dim3(256), 0, s, M, Lp, \
                       T16, S, ori, x, y, b, mode, \
                       pc, skip);                                                                             \
    break;
    RF_G16(1)
    RF_G16(2)
    RF_G16(3)
    RF_G16(4)
    RF_G16(5)
    RF_G16(6)
    RF_G16(7)
    RF_G16(8)
#undef RF_G16
    default:
      return hipErrorInvalidValue;
  }
  if (ev_end) (void)hipEventRecord(ev_end, s);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Single-simulation GMRES(m), a restart cycle with its control on the
// device (the reference's solver reads and writes the Hessenberg entries on
// the host at every step, lib/viennacl/linalg/gmres.hpp:246,281-285).  A
// step is 7 launches after its product: the dots and the update of each
// Gram-Schmidt pass (the update forms the coefficients from the dots' block
// partials itself), ||w||^2, the control step and the normalisation.  Each
// coefficient is the sum of its block partials in block order, the same in
// the update and in the control.

size_t cycle_ctl_doubles(int m) {
  return sizeof(CycleHead) / sizeof(double) + static_cast<size_t>(m + 1) * m + 4 * static_cast<size_t>(m) + 1;
}

CycleCtl cycle_ctl(double* mem, int m) {
  static_assert(sizeof(CycleHead) % sizeof(double) == 0, "CycleHead is a whole number of doubles");
  CycleCtl c;
  c.head = reinterpret_cast<CycleHead*>(mem);
  c.H = mem + sizeof(CycleHead) / sizeof(double);
  c.cs = c.H + static_cast<int64_t>(m + 1) * m;
  c.sn = c.cs + m;
  c.g = c.sn + m;
  c.y = c.g + m + 1;
  c.hdone = nullptr;
  return c;
}

size_t cycle_part_doubles(int64_t n, int m) { return 3 * static_cast<size_t>(m + 1) * mdot_blocks(n); }

__device__ __forceinline__ double sum_parts(const double* __restrict__ p, int nb) {
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += p[b];
  return s;
}

__global__ void cycle_start_kernel(CycleCtl c, int m, double beta) {
  for (int i = threadIdx.x; i <= m; i += blockDim.x) c.g[i] = (i == 0) ? beta : 0.0;
  if (threadIdx.x == 0) {
    c.head->done = 0;
    c.head->k_used = 0;
    c.head->nonfinite = 0;
    c.head->pad = 0;
    c.head->scale = 1.0;
    c.head->res = beta;
    if (c.hdone) *c.hdone = 0;
  }
}

// w -= sum_{k < K} V_k h_k with h_k the sum of the dots' block partials of V_k.
__global__ __launch_bounds__(256) void cgs_update_kernel(int64_t n, int K, const double* __restrict__ V, int64_t ldv,
                                                         const double* __restrict__ part, int nb,
                                                         double* __restrict__ w, const int* __restrict__ skip) {
  extern __shared__ double h[];
  if (*skip) return;
  for (int k = threadIdx.x; k < K; k += blockDim.x) h[k] = sum_parts(part + static_cast<int64_t>(k) * nb, nb);
  __syncthreads();
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int k = 0; k < K; ++k) s = fma(V[static_cast<int64_t>(k) * ldv + r], h[k], s);
  w[r] = fma(-1.0, s, w[r]);
}

// The control of step k: H(:, k) = h0 + h1 and ||w||, the earlier rotations
// and a new one, g and the estimate |g_{k+1}|; at the cycle's end y = H^-1 g.
__global__ __launch_bounds__(64) void arnoldi_step_kernel(int k, int klast, int m, int nb,
                                                          const double* __restrict__ part, double tol, double bnorm,
                                                          CycleCtl c) {
  extern __shared__ double hs[];  // h0[K], h1[K], ||w||^2
  CycleHead& hd = *c.head;
  if (hd.done) return;
  const int K = k + 1;
  const int64_t region = static_cast<int64_t>(m + 1) * nb;
  for (int t = threadIdx.x; t <= 2 * K; t += blockDim.x) {
    const int p = t < K ? 0 : (t < 2 * K ? 1 : 2);
    const int j = t < K ? t : (t < 2 * K ? t - K : 0);
    hs[t] = sum_parts(part + p * region + static_cast<int64_t>(j) * nb, nb);
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double hn = sqrt(hs[2 * K]);
  if (!isfinite(hn)) {
    hd.nonfinite = 1;
    hd.k_used = k;
    hd.done = 1;
    if (c.hdone) *c.hdone = 1;
    return;
  }
  double* Hk = c.H + static_cast<int64_t>(k) * (m + 1);
  for (int i = 0; i <= k; ++i) Hk[i] = hs[i] + hs[K + i];
  Hk[k + 1] = hn;
  for (int i = 0; i < k; ++i) {
    const double a0 = Hk[i], a1 = Hk[i + 1];
    Hk[i] = c.cs[i] * a0 + c.sn[i] * a1;
    Hk[i + 1] = -c.sn[i] * a0 + c.cs[i] * a1;
  }
  const double a0 = Hk[k], a1 = Hk[k + 1];
  const double rr = hypot(a0, a1);
  c.cs[k] = (rr == 0.0) ? 1.0 : a0 / rr;
  c.sn[k] = (rr == 0.0) ? 0.0 : a1 / rr;
  Hk[k] = rr;
  Hk[k + 1] = 0.0;
  c.g[k + 1] = -c.sn[k] * c.g[k];
  c.g[k] = c.cs[k] * c.g[k];
  const double res = fabs(c.g[k + 1]);
  hd.res = res;
  if (res / bnorm <= tol || hn == 0.0 || k == klast) {
    for (int i = k; i >= 0; --i) {
      double acc = c.g[i];
      for (int j = i + 1; j <= k; ++j) acc -= c.H[i + static_cast<int64_t>(j) * (m + 1)] * c.y[j];
      c.y[i] = acc / c.H[i + static_cast<int64_t>(i) * (m + 1)];
    }
    hd.k_used = K;
    hd.done = 1;
    if (c.hdone) *c.hdone = 1;
  } else {
    hd.scale = 1.0 / hn;
  }
}

__global__ void arnoldi_scale_kernel(int64_t n, double* __restrict__ w, const CycleHead* __restrict__ hd) {
  if (hd->done) return;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < n) w[r] = hd->scale * w[r];
}

__global__ void cycle_update_kernel(int64_t n, const double* __restrict__ V, CycleCtl c, double* __restrict__ x) {
  const int K = c.head->k_used;
  if (c.head->nonfinite) return;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int k = 0; k < K; ++k) s = fma(V[static_cast<int64_t>(k) * n + r], c.y[k], s);
  x[r] = fma(1.0, s, x[r]);
}

hipError_t cycle_start(const CycleCtl& c, int m, double beta, hipStream_t s) {
  hipLaunchKernelGGL(cycle_start_kernel, dim3(1), dim3(64), 0, s, c, m, beta);
  return hipGetLastError();
}

hipError_t arnoldi_step(int64_t n, int k, int klast, int m, const double* V, double* w, double* part, double tol,
                        double bnorm, const CycleCtl& c, hipStream_t s) {
  if (k < 0 || k > klast || klast >= m) return hipErrorInvalidValue;
  const int nb = mdot_blocks(n);
  const int64_t rpb = (n + nb - 1) / nb;
  const int64_t region = static_cast<int64_t>(m + 1) * nb;
  const int K = k + 1;
  const unsigned vb = static_cast<unsigned>((n + 255) / 256);
  const int* skip = &c.head->done;  // address arithmetic only
  for (int pass = 0; pass < 2; ++pass) {
    double* pp = part + pass * region;
    hipLaunchKernelGGL(mdot_partial_kernel, dim3(nb, K), dim3(kDotThreads), 0, s, n, rpb, V, n, w, pp, skip);
    hipLaunchKernelGGL(cgs_update_kernel, dim3(vb), dim3(256), K * sizeof(double), s, n, K, V, n, pp, nb, w, skip);
  }
  hipLaunchKernelGGL(mdot_partial_kernel, dim3(nb, 1), dim3(kDotThreads), 0, s, n, rpb, w, n, w, part + 2 * region,
                     skip);
  hipLaunchKernelGGL(arnoldi_step_kernel, dim3(1), dim3(64), (2 * K + 1) * sizeof(double), s, k, klast, m, nb, part,
                     tol, bnorm, c);
  hipLaunchKernelGGL(arnoldi_scale_kernel, dim3(vb), dim3(256), 0, s, n, w, c.head);
  return hipGetLastError();
}

hipError_t cycle_update(int64_t n, const double* V, const CycleCtl& c, double* x, hipStream_t s) {
  hipLaunchKernelGGL(cycle_update_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, n, V, c, x);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// BiCGStab with its scalars on the device.  Each scalar kernel is one wave
// that sums the dots' block partials exactly as mdot_final_kernel does, and
// each vector kernel repeats the axpby sequence of the host form, so the
// iterates are the ones host-side scalars give.

size_t bicg_part_doubles(int64_t n) { return 6 * static_cast<size_t>(mdot_blocks(n)); }

__device__ __forceinline__ double wave_sum_parts(const double* __restrict__ part, int nb) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 64) s += part[b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  return s;
}

__device__ __forceinline__ void bicg_end(const BicgCtl& c) {
  c.head->done = 1;
  c.head->no_restart = 1;
  if (c.hdone) *c.hdone = 1;
}

// Iterations after which BiCGStab restarts from the true residual
// (max_iterations_before_restart, lib/viennacl/linalg/bicgstab.hpp:52).
constexpr int kBicgRestartEvery = 200;

__global__ void bicg_start_kernel(BicgCtl c, double rho, double res) {
  if (threadIdx.x != 0) return;
  BicgHead& h = *c.head;
  h.done = 0;
  h.converged = 0;
  h.nonfinite = 0;
  h.its = 0;
  h.prods = 0;
  h.early_it = 0;
  h.restart = 0;
  h.no_restart = 1;
  h.last_restart = 0;
  h.restarts = 0;
  h.rho = rho;
  h.alpha = h.omega = h.beta = 0.0;
  h.res = res;
  if (c.hdone) *c.hdone = 0;
}

// alpha = rho / <r0, v>; a zero or non-finite <r0, v> ends the solve
__global__ __launch_bounds__(64) void bicg_alpha_kernel(int nb, const double* __restrict__ part, BicgCtl c) {
  BicgHead& h = *c.head;
  if (h.done) return;
  const double r0v = wave_sum_parts(part, nb);
  if (threadIdx.x != 0) return;
  ++h.its;
  ++h.prods;
  if (r0v == 0.0 || !isfinite(r0v)) {
    bicg_end(c);
    return;
  }
  h.alpha = h.rho / r0v;
}

// s = r - alpha v
__global__ void bicg_s_kernel(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                              double* __restrict__ sv, const BicgHead* __restrict__ h) {
  if (h->done) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) sv[i] = fma(-h->alpha, v[i], 1.0 * r[i]);
}

__global__ __launch_bounds__(64) void bicg_snorm_kernel(int it, int nb, const double* __restrict__ part, double tol,
                                                        double bnorm, BicgCtl c) {
  BicgHead& h = *c.head;
  if (h.done) return;
  const double snorm = sqrt(wave_sum_parts(part, nb));
  if (threadIdx.x != 0) return;
  if (!isfinite(snorm)) {
    h.nonfinite = 1;
    bicg_end(c);
  } else if (snorm / bnorm <= tol) {
    h.res = snorm;
    h.converged = 1;
    h.early_it = it;
    bicg_end(c);
  }
}

// x += alpha p when iteration it ended on ||s||
__global__ void bicg_early_kernel(int64_t n, int it, const double* __restrict__ p, double* __restrict__ x,
                                  const BicgHead* __restrict__ h) {
  if (h->early_it != it) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = fma(h->alpha, p[i], 1.0 * x[i]);
}

// omega = <t, s> / <t, t> (0 when <t, t> = 0)
__global__ __launch_bounds__(64) void bicg_omega_kernel(int nb, const double* __restrict__ part, BicgCtl c) {
  BicgHead& h = *c.head;
  if (h.done) return;
  const double ts = wave_sum_parts(part, nb), tt = wave_sum_parts(part + nb, nb);
  if (threadIdx.x != 0) return;
  ++h.prods;
  h.omega = (tt == 0.0) ? 0.0 : ts / tt;
}

// x += alpha p + omega s, r = s - omega t
__global__ void bicg_xr_kernel(int64_t n, const double* __restrict__ p, const double* __restrict__ sv,
                               const double* __restrict__ t, double* __restrict__ x, double* __restrict__ r,
                               const BicgHead* __restrict__ h) {
  if (h->done) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = fma(h->alpha, p[i], 1.0 * x[i]);
  x[i] = fma(h->omega, sv[i], 1.0 * xi);
  r[i] = fma(-h->omega, t[i], 1.0 * (1.0 * sv[i]));
}

// ||r||, the end tests, beta and rho from ||r||^2 and <r0, r>.  A degenerate
// direction (<r0, r> = 0 or omega = 0) or more than kBicgRestartEvery
// iterations since the last restart sets up a restart from the true residual
// at the next iteration (lib/viennacl/linalg/bicgstab.hpp:150-152).
__global__ __launch_bounds__(64) void bicg_end_kernel(int nb, const double* __restrict__ part, double tol,
                                                      double bnorm, BicgCtl c) {
  BicgHead& h = *c.head;
  if (h.done) return;
  const double rr = wave_sum_parts(part, nb), rho_new = wave_sum_parts(part + nb, nb);
  if (threadIdx.x != 0) return;
  const double res = sqrt(rr);
  h.res = res;
  if (!isfinite(res)) {
    h.nonfinite = 1;
    bicg_end(c);
    return;
  }
  if (res / bnorm <= tol) {
    h.converged = 1;
    bicg_end(c);
    return;
  }
  if (rho_new == 0.0 || h.omega == 0.0 || (h.its - 1) - h.last_restart > kBicgRestartEvery) {
    h.restart = 1;
    h.no_restart = 0;
    return;
  }
  h.beta = (rho_new / h.rho) * (h.alpha / h.omega);
  h.rho = rho_new;
}

// p = r + beta (p - omega v) (a pending restart sets p itself)
__global__ void bicg_p_kernel(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                              double* __restrict__ p, const BicgHead* __restrict__ h) {
  if (h->done || h->restart) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double q = fma(-h->omega, v[i], 1.0 * p[i]);
  p[i] = fma(1.0, r[i], h->beta * q);
}

// restart: r0 = p = r (r = b - A x from the caller's product)
__global__ void bicg_restart_copy_kernel(int64_t n, const double* __restrict__ r, double* __restrict__ r0,
                                         double* __restrict__ p, const BicgHead* __restrict__ h) {
  if (h->no_restart) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  r0[i] = r[i];
  p[i] = r[i];
}

// restart: rho = ||r||^2, the restart recorded
__global__ __launch_bounds__(64) void bicg_restart_kernel(int nb, const double* __restrict__ part, BicgCtl c) {
  BicgHead& h = *c.head;
  if (h.no_restart) return;
  const double rr = wave_sum_parts(part, nb);
  if (threadIdx.x != 0) return;
  h.rho = rr;
  h.res = sqrt(rr);
  h.restart = 0;
  h.no_restart = 1;
  h.last_restart = h.its;  // the index of the iteration that starts now
  ++h.restarts;
  if (!isfinite(rr)) {
    h.nonfinite = 1;
    bicg_end(c);
  }
}

hipError_t bicg_start(const BicgCtl& c, double rho, double res, hipStream_t s) {
  hipLaunchKernelGGL(bicg_start_kernel, dim3(1), dim3(64), 0, s, c, rho, res);
  return hipGetLastError();
}

hipError_t bicg_restart(int64_t n, const BicgVecs& w, double* part, const BicgCtl& c, hipStream_t s) {
  const int nb = mdot_blocks(n);
  const int64_t rpb = (n + nb - 1) / nb;
  const unsigned vb = static_cast<unsigned>((n + 255) / 256);
  const int* skip = &c.head->no_restart;  // address arithmetic only
  hipLaunchKernelGGL(bicg_restart_copy_kernel, dim3(vb), dim3(256), 0, s, n, w.r, w.r + n, w.p, c.head);
  hipLaunchKernelGGL(mdot_partial_kernel, dim3(nb, 1), dim3(kDotThreads), 0, s, n, rpb, w.r, n, w.r, part, skip);
  hipLaunchKernelGGL(bicg_restart_kernel, dim3(1), dim3(64), 0, s, nb, static_cast<const double*>(part), c);
  return hipGetLastError();
}

hipError_t bicg_after_v(int64_t n, int it, const BicgVecs& w, double* part, double tol, double bnorm,
                        const BicgCtl& c, hipStream_t s) {
  const int nb = mdot_blocks(n);
  const int64_t rpb = (n + nb - 1) / nb;
  const unsigned vb = static_cast<unsigned>((n + 255) / 256);
  const int* skip = &c.head->done;  // address arithmetic only
  const double* r0 = w.r + n;
  hipLaunchKernelGGL(mdot_partial_kernel, dim3(nb, 1), dim3(kDotThreads), 0, s, n, rpb, r0, n, w.v, part, skip);
  hipLaunchKernelGGL(bicg_alpha_kernel, dim3(1), dim3(64), 0, s, nb, static_cast<const double*>(part), c);
  hipLaunchKernelGGL(bicg_s_kernel, dim3(vb), dim3(256), 0, s, n, w.r, w.v, w.s, c.head);
  hipLaunchKernelGGL(mdot_partial_kernel, dim3(nb, 1), dim3(kDotThreads), 0, s, n, rpb, w.s, n, w.s, part + nb,
                     skip);
  hipLaunchKernelGGL(bicg_snorm_kernel, dim3(1), dim3(64), 0, s, it, nb, static_cast<const double*>(part + nb), tol,
                     bnorm, c);
  hipLaunchKernelGGL(bicg_early_kernel, dim3(vb), dim3(256), 0, s, n, it, w.p, w.x, c.head);
  return hipGetLastError();
}

hipError_t bicg_after_t(int64_t n, const BicgVecs& w, double* part, double tol, double bnorm, const BicgCtl& c,
                        hipStream_t s) {
  const int nb = mdot_blocks(n);
  const int64_t rpb = (n + nb - 1) / nb;
  const unsigned vb = static_cast<unsigned>((n + 255) / 256);
  const int* skip = &c.head->done;
  const double* t = w.s + n;
  // <s, t>, <t, t>
  hipLaunchKernelGGL(mdot_partial_kernel, dim3(nb, 2), dim3(kDotThreads), 0, s, n, rpb, w.s, n, t, part + 2 * nb,
                     skip);
  hipLaunchKernelGGL(bicg_omega_kernel, dim3(1), dim3(64), 0, s, nb, static_cast<const double*>(part + 2 * nb), c);
  hipLaunchKernelGGL(bicg_xr_kernel, dim3(vb), dim3(256), 0, s, n, w.p, w.s, t, w.x, w.r, c.head);
  // ||r||^2, <r0, r>
  hipLaunchKernelGGL(mdot_partial_kernel, dim3(nb, 2), dim3(kDotThreads), 0, s, n, rpb, w.r, n, w.r, part + 4 * nb,
                     skip);
  hipLaunchKernelGGL(bicg_end_kernel, dim3(1), dim3(64), 0, s, nb, static_cast<const double*>(part + 4 * nb), tol,
                     bnorm, c);
  hipLaunchKernelGGL(bicg_p_kernel, dim3(vb), dim3(256), 0, s, n, w.r, w.v, w.p, c.head);
  return hipGetLastError();
}

}  // namespace rf
