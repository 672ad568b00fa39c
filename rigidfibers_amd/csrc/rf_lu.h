// rf_lu.h -- dense LU factorisation and solves for the direct solvers
// (--magma: magma_sgesv_gpu at src/simulation.cu:403-418; sgesv at
// fortran/solve_system.incl:3) and their mixed-precision forms.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>

#include "rf_internal.h"

namespace rf {

enum class LuMode { kFP32, kBF16, kFP16 };

constexpr int kLuNb = 64;       // diagonal block: pivoting happens inside it
constexpr int kLuPanel = 1024;  // outer panel
// Panels per trailing update (its K = kLuGroup kLuPanel).  Measured at
// n = 75 000 in one run (BF16 / FP16 ms per step, trailing GEMM TFLOP/s):
// 2 -> 495 / 456 (749 / 833), 3 -> 494 / 457 (787 / 878), 4 -> 508 / 470
// (799 / 902: the longer in-group chains on the look-ahead stream cost more
// than the faster update saves).
constexpr int kLuGroup = 3;
// The trailing update beside the look-ahead's panels (BF16 / FP16), once at
// most kLuTailRows rows remain (where the panels' chain is the longer of the two):
// kLuTailWorkgroups workgroups (one per CU) take its tiles, the other CUs run
// the chain's small kernels as they come instead of waiting for a tile to end.
// Measured at n = 75 000 (BF16 / FP16 ms per step, interleaved runs on two
// boxes, against 448.6 / 409.5 and 455.7 / 414.0 without): 24 576 rows, 224
// workgroups -2 / -2; 36 864, 224: -4 / -4.5; 49 152, 224: -4.4 / -2.9;
// 36 864, 192: -8.1 / -7.8 and -8.5 / -8.7; 36 864, 160: -8.0 / -6.5;
// 49 152, 192: -2.8 / -1.5.
constexpr int64_t kLuTailRows = 36864;
constexpr int kLuTailWorkgroups = 192;

struct LuWorkspace {
  LuMode mode = LuMode::kFP32;
  int64_t n = 0, lda = 0;
  float* F = nullptr;      // factors (FP32, column-major, lda); the assembled A is kept intact
  int* perm = nullptr;     // perm[k0 + c] = original row of factored row k0 + c (block-local)
  int* info = nullptr;     // device: first zero pivot (1-based) or 0
  double* vec = nullptr;   // scratch vector of the solves (n doubles)
  // packs, K-contiguous, in the mode's element type (per-32-row / per-32-column
  // scales in the mixed modes); panels are taken in groups of kLuGroup (rf_lu.hip)
  // L21 of a group's panels, rows below the group: kLuGroup kLuPanel per row;
  // two, by the group's parity (the look-ahead stream packs the next group's
  // while the main stream's update reads the current one)
  void* At[2] = {};
  float* sA[2] = {};
  void* Bt = nullptr;      // U12 of the group's panels, columns right of the group: kLuGroup kLuPanel per column
  float* sB = nullptr;
  // the group's own updates (at most kLuGroup - 1 panels deep): rows and
  // columns packs, one pair for the panels' stream (0), one for the main (1)
  void* Pr[2] = {};        // rows: n (0) / kLuPanel (1) of them
  void* Pc[2] = {};        // columns: kLuPanel (0) / n (1) of them
  float* sPr[2] = {};
  float* sPc[2] = {};
  void* Xm = nullptr;      // A12 before its solve (pack of columns)
  void* Ya = nullptr;      // A21 before its solve (pack of rows)
  void* Tinv = nullptr;    // kLuGroup x kLuPanel^2: the forward operators of the group's diagonal blocks
  void* Uinv = nullptr;    // kLuPanel^2: U_D^-1 by columns
  // per panel with rows below it, FP32, row-major kLuPanel^2: the forward
  // operator and U_D^-1 again, for the solves
  float* Tsol = nullptr;
  float* Usol = nullptr;
  hipStream_t aux = nullptr;  // the look-ahead's stream (default priority)
  hipEvent_t next_cols = nullptr, next_done = nullptr;
  double* Dinv = nullptr;  // per diagonal block b: L_bb^-1 then U_bb^-1, 64 x 64 FP64 each, [j][c]
  static constexpr int kEvents = 256;  // event pairs around the trailing updates
  hipEvent_t ev[2 * kEvents] = {};
  int64_t bytes = 0;
};

struct LuResult {
  int info = 0;                   // 0, or k+1 for an exactly zero pivot at column k
  int refinement_iterations = 0;  // FP64-residual refinement sweeps
  double relative_residual = 0.0; // ||b - A x|| / ||b|| with the FP32 A, in FP64
  bool converged = false;
  double factor_ms = 0.0;         // factorisation (device time)
  double gemm_ms = 0.0;           // summed duration of the trailing-update GEMM launches
  double gemm_flop = 0.0;         // their algorithmic flops (2 m n K each)
  int gemm_launches = 0;
};

// C -= A B (store: C = A B) with K-contiguous packs At (m rows, row stride
// ldak elements) and Bt (n rows, stride ldbk) of the mode's element type; K a
// multiple of gemm_stage_k(mode), pack rows 16-byte aligned.  C FP32
// column-major.  sa, sb (both or neither; subtracting form only): the packs
// hold row block r / 32 of A times sa[r / 32] and column block c / 32 of B
// times sb[c / 32] (powers of two), which the update divides out.
int gemm_stage_k(LuMode mode);
// max_workgroups > 0 (subtracting form): at most that many workgroups (rounded
// down to a multiple of 8), each taking several tiles, so that the rest of the
// CUs stay free for work on other streams.
hipError_t gemm_packed(LuMode mode, bool store, int64_t m, int64_t n, int64_t K, const void* At, int64_t ldak,
                       const void* Bt, int64_t ldbk, float* C, int64_t ldc, hipStream_t s,
                       const float* sa = nullptr, const float* sb = nullptr, int max_workgroups = 0);
hipError_t gemm_packed_sub(LuMode mode, int64_t m, int64_t n, int64_t K, const void* At, int64_t ldak,
                           const void* Bt, int64_t ldbk, float* C, int64_t ldc, hipStream_t s,
                           const float* sa = nullptr, const float* sb = nullptr, int max_workgroups = 0);

hipError_t lu_workspace_create(LuMode mode, int64_t n, int64_t lda, LuWorkspace* ws);
int64_t lu_workspace_bytes(const LuWorkspace& ws);
void lu_workspace_destroy(LuWorkspace* ws);

// The system an LU solve factors and refines against: fill writes the FP32 A
// (column-major, leading dimension lda) into the factor storage F;
// residual forms r = b - A x with FP64 sums of the same FP32 A.  Both are
// enqueued on the stream they are given.
struct LuSystem {
  std::function<hipError_t(float* F, int64_t lda, hipStream_t s)> fill;
  std::function<hipError_t(const double* x, double* r, hipStream_t s)> residual;
};

// Factors A (from sys.fill) and solves A x = b with FP64 iterative
// refinement of the residual from sys.residual.  work: >= 2n doubles;
// dpart/hdev: dot workspace.
hipError_t lu_solve(LuWorkspace& ws, LuMode mode, const LuSystem& sys, int64_t n, const double* b, double* x,
                    double* work, double* dpart, double* hdev, double tol, int maxit, hipStream_t s,
                    LuResult* result);

}  // namespace rf
