// rf_lu.hip -- blocked right-looking LU of the assembled system and the
// solves of the direct solvers.
//
// Reference: magma_sgesv_gpu (src/simulation.cu:403-418) and LAPACK sgesv
// (fortran/solve_system.incl:3): LU with row pivoting, then L y = P b, U x = y.
//
// Factorisation of a copy F of A (A itself is kept for the residuals):
//   for each panel of nb columns at k0:
//     1. diag: one workgroup factors the nb x nb diagonal block in LDS with
//        partial pivoting among its rows (the pivot search is restricted to
//        the diagonal block, which keeps the panel off the critical path of a
//        75 000-row reduction per column);
//     2. swap: the block's row interchanges are applied to every other column;
//     3. trsm: L21 = A21 U11^-1 (one lane per row) and U12 = L11^-1 A12 (one
//        lane per column);
//     4. gemm: A22 -= L21 U12 (the O(n^3) part).
// Every solve is followed by FP64 iterative refinement with residuals of the
// FP32 A (r = b - A x accumulated in FP64), so the returned x carries a
// verified ||b - A x|| / ||b||; a matrix that would need pivoting outside the
// diagonal block shows up as a refinement that does not converge.
#include <cmath>
#include <vector>

#include "rf_lu.h"

namespace rf {

constexpr int kLuNb = 64;      // inner panel (diagonal block in LDS)
constexpr int kLuOuter = 256;  // outer panel: the K of the trailing MFMA update

// ---------------------------------------------------------------------------
// 1. diagonal block: GEPP inside one workgroup.
template <int NB>
__global__ __launch_bounds__(256) void lu_diag_kernel(int64_t n, int64_t lda, float* __restrict__ F, int64_t k0,
                                                      int kb, int* __restrict__ ipiv, int* __restrict__ info) {
  __shared__ float blk[NB][NB + 1];  // [col][row]
  __shared__ int piv_row;
  __shared__ float piv_abs[256];
  __shared__ int piv_idx[256];
  const int tid = threadIdx.x;
  for (int e = tid; e < kb * kb; e += 256) {
    const int c = e / kb, r = e - c * kb;
    blk[c][r] = F[(k0 + c) * lda + k0 + r];
  }
  __syncthreads();
  for (int c = 0; c < kb; ++c) {
    // pivot: argmax |blk[c][r]|, r >= c
    float best = -1.0f;
    int bi = c;
    for (int r = c + tid; r < kb; r += 256) {
      const float v = fabsf(blk[c][r]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
    piv_abs[tid] = best;
    piv_idx[tid] = bi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if (tid < off) {
        const float o = piv_abs[tid + off];
        const int oi = piv_idx[tid + off];
        if (o > piv_abs[tid] || (o == piv_abs[tid] && oi < piv_idx[tid])) {
          piv_abs[tid] = o;
          piv_idx[tid] = oi;
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      piv_row = piv_idx[0];
      ipiv[k0 + c] = static_cast<int>(k0) + piv_idx[0];
      if (piv_abs[0] == 0.0f && info[0] == 0) info[0] = static_cast<int>(k0) + c + 1;
    }
    __syncthreads();
    const int pr = piv_row;
    if (pr != c)
      for (int cc = tid; cc < kb; cc += 256) {
        const float t = blk[cc][c];
        blk[cc][c] = blk[cc][pr];
        blk[cc][pr] = t;
      }
    __syncthreads();
    const float d = blk[c][c];
    const float inv = (d != 0.0f) ? 1.0f / d : 0.0f;
    for (int r = c + 1 + tid; r < kb; r += 256) blk[c][r] *= inv;
    __syncthreads();
    const int rem = kb - c - 1;
    for (int e = tid; e < rem * rem; e += 256) {
      const int cc = c + 1 + e / rem, r = c + 1 + (e - (e / rem) * rem);
      blk[cc][r] = fmaf(-blk[c][r], blk[cc][c], blk[cc][r]);
    }
    __syncthreads();
  }
  for (int e = tid; e < kb * kb; e += 256) {
    const int c = e / kb, r = e - c * kb;
    F[(k0 + c) * lda + k0 + r] = blk[c][r];
  }
}

// 2. row interchanges of block k0 applied to the columns outside [k0, k0+kb).
__global__ void lu_swap_kernel(int64_t n, int64_t lda, float* __restrict__ F, int64_t k0, int kb,
                               const int* __restrict__ ipiv) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n || (c >= k0 && c < k0 + kb)) return;
  float* col = F + c * lda;
  for (int k = 0; k < kb; ++k) {
    const int64_t p = ipiv[k0 + k];
    if (p != k0 + k) {
      const float t = col[k0 + k];
      col[k0 + k] = col[p];
      col[p] = t;
    }
  }
}

// 3a. L21 = A21 U11^-1: lane per row r in [r0, r1), forward over the block's columns.
template <int NB>
__global__ __launch_bounds__(256) void lu_trsm_l21_kernel(int64_t r0, int64_t r1, int64_t lda, float* __restrict__ F,
                                                          int64_t k0, int kb) {
  __shared__ float U[NB][NB + 1];  // [col][row]
  for (int e = threadIdx.x; e < kb * kb; e += 256) {
    const int c = e / kb, r = e - c * kb;
    U[c][r] = F[(k0 + c) * lda + k0 + r];
  }
  __syncthreads();
  const int64_t r = r0 + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (r >= r1) return;
  float x[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) x[c] = (c < kb) ? F[(k0 + c) * lda + r] : 0.0f;
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    if (c < kb) {
      float v = x[c];
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j < c) v = fmaf(-x[j], U[c][j], v);
      x[c] = v / U[c][c];
    }
  }
#pragma unroll
  for (int c = 0; c < NB; ++c)
    if (c < kb) F[(k0 + c) * lda + r] = x[c];
}

// 3b. U12 = L11^-1 A12: lane per column c in [c0, c1) (unit lower L11).
template <int NB>
__global__ __launch_bounds__(256) void lu_trsm_u12_kernel(int64_t c0, int64_t c1, int64_t lda, float* __restrict__ F,
                                                          int64_t k0, int kb) {
  __shared__ float L[NB][NB + 1];  // [col][row]
  for (int e = threadIdx.x; e < kb * kb; e += 256) {
    const int c = e / kb, r = e - c * kb;
    L[c][r] = F[(k0 + c) * lda + k0 + r];
  }
  __syncthreads();
  const int64_t c = c0 + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (c >= c1) return;
  float* col = F + c * lda + k0;
  float x[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) x[r] = (r < kb) ? col[r] : 0.0f;
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    if (r < kb) {
      float v = x[r];
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j < r) v = fmaf(-L[j][r], x[j], v);
      x[r] = v;
    }
  }
#pragma unroll
  for (int r = 0; r < NB; ++r)
    if (r < kb) col[r] = x[r];
}

// ---------------------------------------------------------------------------
// Solves with the factors (FP64 vectors, FP32 factors).

__global__ void lu_apply_pivots_kernel(int64_t n, const int* __restrict__ ipiv, double* __restrict__ y) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int64_t k = 0; k < n; ++k) {
    const int64_t p = ipiv[k];
    if (p != k) {
      const double t = y[k];
      y[k] = y[p];
      y[p] = t;
    }
  }
}

// y[k0:k0+kb] <- T^-1 y[k0:k0+kb], T the unit-lower (lower = 1) or upper diagonal block.
__global__ __launch_bounds__(64) void lu_trsv_block_kernel(int64_t lda, const float* __restrict__ F, int64_t k0,
                                                           int kb, int lower, double* __restrict__ y) {
  __shared__ double ys[kLuNb];
  const int t = threadIdx.x;
  if (t < kb) ys[t] = y[k0 + t];
  __syncthreads();
  if (lower) {
    for (int c = 0; c < kb; ++c) {
      if (t > c && t < kb) ys[t] -= static_cast<double>(F[(k0 + c) * lda + k0 + t]) * ys[c];
      __syncthreads();
    }
  } else {
    for (int c = kb - 1; c >= 0; --c) {
      if (t == c) ys[c] /= static_cast<double>(F[(k0 + c) * lda + k0 + c]);
      __syncthreads();
      if (t < c) ys[t] -= static_cast<double>(F[(k0 + c) * lda + k0 + t]) * ys[c];
      __syncthreads();
    }
  }
  if (t < kb) y[k0 + t] = ys[t];
}

// y[rlo:rhi] -= F[rlo:rhi, k0:k0+kb] y[k0:k0+kb]
__global__ void lu_panel_gemv_kernel(int64_t lda, const float* __restrict__ F, int64_t k0, int kb, int64_t rlo,
                                     int64_t rhi, double* __restrict__ y) {
  const int64_t r = rlo + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= rhi) return;
  double s = 0.0;
  for (int c = 0; c < kb; ++c) s = fma(static_cast<double>(F[(k0 + c) * lda + r]), y[k0 + c], s);
  y[r] -= s;
}

// Two-level right-looking factorisation: inner panels of kLuNb columns are
// factored inside an outer panel of kLuOuter columns (their updates touch only
// the outer panel's columns); the outer panel's rows of U are then solved for
// and the trailing matrix receives one rank-kLuOuter update on MFMA.
static hipError_t lu_factor(LuWorkspace& ws, LuMode mode, hipStream_t s) {
  const int64_t n = ws.n, lda = ws.lda;
  float* F = ws.F;
  auto at = [&](int64_t r, int64_t c) { return F + c * lda + r; };
  auto blocks = [](int64_t len) { return static_cast<unsigned>((len + 255) / 256); };
  hipError_t e = hipSuccess;
  for (int64_t K0 = 0; K0 < n; K0 += kLuOuter) {
    const int64_t KE = std::min<int64_t>(n, K0 + kLuOuter);
    for (int64_t k0 = K0; k0 < KE; k0 += kLuNb) {
      const int kb = static_cast<int>(std::min<int64_t>(kLuNb, KE - k0));
      const int64_t k1 = k0 + kb;
      hipLaunchKernelGGL((lu_diag_kernel<kLuNb>), dim3(1), dim3(256), 0, s, n, lda, F, k0, kb, ws.ipiv, ws.info);
      hipLaunchKernelGGL(lu_swap_kernel, dim3(blocks(n)), dim3(256), 0, s, n, lda, F, k0, kb, ws.ipiv);
      if (k1 < n)
        hipLaunchKernelGGL((lu_trsm_l21_kernel<kLuNb>), dim3(blocks(n - k1)), dim3(256), 0, s, k1, n, lda, F, k0, kb);
      if (k1 < KE) {
        hipLaunchKernelGGL((lu_trsm_u12_kernel<kLuNb>), dim3(blocks(KE - k1)), dim3(256), 0, s, k1, KE, lda, F, k0,
                           kb);
        if ((e = gemm_sub(mode, n - k1, KE - k1, kb, at(k1, k0), lda, at(k0, k1), lda, at(k1, k1), lda, s)) !=
            hipSuccess)
          return e;
      }
    }
    if (KE < n) {
      // rows K0..KE of U for the trailing columns: U12 = L11^-1 A12, blockwise
      for (int64_t k0 = K0; k0 < KE; k0 += kLuNb) {
        const int kb = static_cast<int>(std::min<int64_t>(kLuNb, KE - k0));
        const int64_t k1 = k0 + kb;
        hipLaunchKernelGGL((lu_trsm_u12_kernel<kLuNb>), dim3(blocks(n - KE)), dim3(256), 0, s, KE, n, lda, F, k0, kb);
        if (k1 < KE &&
            (e = gemm_sub(mode, KE - k1, n - KE, kb, at(k1, k0), lda, at(k0, KE), lda, at(k1, KE), lda, s)) !=
                hipSuccess)
          return e;
      }
      if ((e = gemm_sub(mode, n - KE, n - KE, KE - K0, at(KE, K0), lda, at(K0, KE), lda, at(KE, KE), lda, s)) !=
          hipSuccess)
        return e;
    }
  }
  return hipGetLastError();
}

// y <- (LU)^-1 P y
static hipError_t lu_apply_inverse(LuWorkspace& ws, double* y, hipStream_t s) {
  const int64_t n = ws.n, lda = ws.lda;
  hipLaunchKernelGGL(lu_apply_pivots_kernel, dim3(1), dim3(1), 0, s, n, ws.ipiv, y);
  for (int64_t k0 = 0; k0 < n; k0 += kLuNb) {
    const int kb = static_cast<int>(std::min<int64_t>(kLuNb, n - k0));
    hipLaunchKernelGGL(lu_trsv_block_kernel, dim3(1), dim3(64), 0, s, lda, ws.F, k0, kb, 1, y);
    const int64_t rlo = k0 + kb;
    if (rlo < n)
      hipLaunchKernelGGL(lu_panel_gemv_kernel, dim3(static_cast<unsigned>((n - rlo + 255) / 256)), dim3(256), 0, s,
                         lda, ws.F, k0, kb, rlo, n, y);
  }
  const int64_t last = ((n - 1) / kLuNb) * kLuNb;
  for (int64_t k0 = last; k0 >= 0; k0 -= kLuNb) {
    const int kb = static_cast<int>(std::min<int64_t>(kLuNb, n - k0));
    hipLaunchKernelGGL(lu_trsv_block_kernel, dim3(1), dim3(64), 0, s, lda, ws.F, k0, kb, 0, y);
    if (k0 > 0)
      hipLaunchKernelGGL(lu_panel_gemv_kernel, dim3(static_cast<unsigned>((k0 + 255) / 256)), dim3(256), 0, s, lda,
                         ws.F, k0, kb, int64_t(0), k0, y);
  }
  return hipGetLastError();
}

hipError_t lu_workspace_create(LuMode mode, int64_t n, int64_t lda, LuWorkspace* ws) {
  ws->mode = mode;
  ws->n = n;
  ws->lda = lda;
  ws->nb = kLuNb;
  hipError_t e;
  if ((e = hipMalloc(&ws->F, sizeof(float) * static_cast<size_t>(lda) * n)) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->ipiv, sizeof(int) * static_cast<size_t>(n))) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->info, sizeof(int))) != hipSuccess) return e;
  if ((e = hipMalloc(&ws->vec, sizeof(double) * 3 * static_cast<size_t>(n))) != hipSuccess) return e;
  ws->bytes = static_cast<int64_t>(sizeof(float) * static_cast<size_t>(lda) * n + sizeof(int) * (n + 1) +
                                   sizeof(double) * 3 * static_cast<size_t>(n));
  return hipSuccess;
}

int64_t lu_workspace_bytes(const LuWorkspace& ws) { return ws.bytes; }

void lu_workspace_destroy(LuWorkspace* ws) {
  if (ws->F) (void)hipFree(ws->F);
  if (ws->ipiv) (void)hipFree(ws->ipiv);
  if (ws->info) (void)hipFree(ws->info);
  if (ws->vec) (void)hipFree(ws->vec);
  *ws = LuWorkspace{};
}

hipError_t lu_solve(LuWorkspace& ws, LuMode mode, const float* A, int64_t lda, int64_t n, const double* b,
                    double* x, double* gpart, const GemvPlan& plan, double* work, double* dpart, double* hdev,
                    double tol, int maxit, hipStream_t s, LuResult* result) {
  hipError_t e;
  *result = LuResult{};
  if ((e = hipMemcpy2DAsync(ws.F, lda * sizeof(float), A, lda * sizeof(float), n * sizeof(float), n,
                            hipMemcpyDeviceToDevice, s)) != hipSuccess)
    return e;
  if ((e = hipMemsetAsync(ws.info, 0, sizeof(int), s)) != hipSuccess) return e;
  if ((e = lu_factor(ws, mode, s)) != hipSuccess) return e;
  int info = 0;
  if ((e = hipMemcpyAsync(&info, ws.info, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  result->info = info;
  if (info != 0) return hipSuccess;
  double* r = work;
  // ||b||
  if ((e = mdot(n, 1, b, n, b, dpart, hdev, s)) != hipSuccess) return e;
  double bb = 0.0;
  if ((e = hipMemcpyAsync(&bb, hdev, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  const double bnorm = std::sqrt(bb);
  // x = (LU)^-1 P b
  if ((e = axpby(n, 1.0, b, 0.0, x, s)) != hipSuccess) return e;
  if ((e = lu_apply_inverse(ws, x, s)) != hipSuccess) return e;
  double res = 0.0;
  for (int it = 0;; ++it) {
    if ((e = gemv(plan, A, x, gpart, r, b, 1, s)) != hipSuccess) return e;
    if ((e = mdot(n, 1, r, n, r, dpart, hdev, s)) != hipSuccess) return e;
    double rr = 0.0;
    if ((e = hipMemcpyAsync(&rr, hdev, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    res = (bnorm > 0.0) ? std::sqrt(rr) / bnorm : 0.0;
    result->refinement_iterations = it;
    if (!(res > tol) || it >= maxit || !std::isfinite(res)) break;
    if ((e = lu_apply_inverse(ws, r, s)) != hipSuccess) return e;
    if ((e = axpby(n, 1.0, r, 1.0, x, s)) != hipSuccess) return e;
  }
  result->relative_residual = res;
  result->converged = res <= tol;
  return hipSuccess;
}

}  // namespace rf
