// rf_lu.h -- dense LU factorisation and solves for the direct solvers
// (--magma: magma_sgesv_gpu at src/simulation.cu:403-418; sgesv at
// fortran/solve_system.incl:3) and their mixed-precision forms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_internal.h"

namespace rf {

enum class LuMode { kFP32, kBF16, kFP16 };

struct LuWorkspace {
  LuMode mode = LuMode::kFP32;
  int64_t n = 0, lda = 0;
  int nb = 0;
  float* F = nullptr;      // factors (FP32, column-major, lda); the assembled A is kept intact
  int* ipiv = nullptr;     // device pivots (global row index chosen at each column)
  int* info = nullptr;     // device: first zero pivot (1-based) or 0
  double* vec = nullptr;   // 3 x n scratch vectors for the solves and refinement
  int64_t bytes = 0;
};

struct LuResult {
  int info = 0;                   // 0, or k+1 for an exactly zero pivot at column k
  int refinement_iterations = 0;  // FP64-residual refinement sweeps
  double relative_residual = 0.0; // ||b - A x|| / ||b|| with the FP32 A, in FP64
  bool converged = false;
};

// C -= A B, column-major FP32 blocks, products on MFMA in the given mode.
hipError_t gemm_sub(LuMode mode, int64_t m, int64_t n, int64_t k, const float* A, int64_t lda, const float* B,
                    int64_t ldb, float* C, int64_t ldc, hipStream_t s);

hipError_t lu_workspace_create(LuMode mode, int64_t n, int64_t lda, LuWorkspace* ws);
int64_t lu_workspace_bytes(const LuWorkspace& ws);
void lu_workspace_destroy(LuWorkspace* ws);

// Factors a copy of A (FP32 column-major, lda) and solves A x = b with FP64
// iterative refinement of the residual computed from the FP32 A.
// gpart/plan: GEMV workspace; work: >= 2n doubles; dpart/hdev: dot workspace.
hipError_t lu_solve(LuWorkspace& ws, LuMode mode, const float* A, int64_t lda, int64_t n, const double* b,
                    double* x, double* gpart, const GemvPlan& plan, double* work, double* dpart, double* hdev,
                    double tol, int maxit, hipStream_t s, LuResult* result);

}  // namespace rf
