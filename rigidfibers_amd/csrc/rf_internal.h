// rf_internal.h -- host-side launchers shared between the translation units of
// librigidfibers.so.  Not part of the public C ABI (include/rigidfibers.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rf_device.h"

namespace rf {

// Geometry of one launch over fiber pairs: blockIdx.x tiles the target fibers,
// blockIdx.y the groups of source fibers (each group reduced into its own
// partial-sum slot, so sums over sources are deterministic).
struct PairGrid {
  int tiles;    // target tiles
  int groups;   // source groups
  int jgroup;   // source fibers per group
};
PairGrid pair_grid(int M, int pairs_per_tile);
int max_pair_groups(int M);

// Fiber-pair kernels (rf_assemble.hip / rf_velocity.hip).  All positions and
// orientations are [3*M] FP64 device arrays.
hipError_t launch_assemble(int N, int NQI, bool analytical, int M, int64_t lda, const double* pos,
                           const double* ori, float* A, double* bpart, const PhysConsts& pc, hipStream_t s);
hipError_t launch_sum_groups(int groups, int64_t len, const double* part, double* out, hipStream_t s);

// Theta form (rf_assemble.hip): T holds NC = 3 + 6(N-1) component arrays of
// stride Lp = theta_stride(M, N) per target fiber; *groups receives the number
// of RHS partial slots written.  With Tv != nullptr it also stores the
// velocity operator: 12 component arrays of the same stride per target fiber
// (Theta_0 and Phi0 of every source mode) and the velocity constants
// cpart[group][i][6] (cpart then holds max_pair_groups(M) * M * 6 doubles).
PairGrid theta_grid(int M, int sources_per_iteration);
int64_t theta_stride(int M, int N);
hipError_t launch_assemble_theta(int N, int NQI, bool analytical, int M, int64_t Lp, const double* pos,
                                 const double* ori, float* T, double* bpart, float* Tv, double* cpart,
                                 const PhysConsts& pc, hipStream_t s, int* groups);
// Velocities from the stored operator of the last theta assembly (one pass
// over Tv): tvel, rvel [3*M] FP64.
hipError_t launch_velocities_stored(int N, int M, int64_t Lp, const float* Tv, const double* cpart, int groups,
                                    const double* ori, const double* x, const PhysConsts& pc, double* tvel,
                                    double* rvel, hipStream_t s);
hipError_t launch_expand_theta(int N, int M, int64_t Lp, const float* T, const double* ori, const PhysConsts& pc,
                               float* A, hipStream_t s);
// y = A x (mode 0) or y = b - A x (mode 1) with A in theta form.
hipError_t gemv_theta(int N, int M, int64_t Lp, const float* T, const double* ori, const double* x, double* y,
                      const double* b, int mode, const PhysConsts& pc, hipStream_t s, hipEvent_t ev_begin = nullptr,
                      hipEvent_t ev_end = nullptr);
hipError_t launch_velocities(int N, int NQI, bool analytical, int M, const double* pos, const double* ori,
                             const double* x, float* wf, double* fmono, double* vpart, double* tvel, double* rvel,
                             const PhysConsts& pc, hipStream_t s);
hipError_t launch_update(bool first, int M, double dt, const double* pos_prev, const double* pos,
                         const double* ori_prev, const double* ori, const double* tv_prev, const double* tv,
                         const double* rv_prev, const double* rv, double* pos_next, double* ori_next,
                         hipStream_t s);
bool assemble_supported(int N, int NQI);

// Dense linear algebra (rf_linalg.hip).
struct GemvPlan {
  int64_t n, lda;
  int row_tiles, chunks, cols_per_chunk;
};
GemvPlan gemv_plan(int64_t n, int64_t lda);
size_t gemv_partial_elems(const GemvPlan& p);
// y = A x  (mode 0) or y = b - A x (mode 1); A FP32 column-major.  Optional
// events are recorded around the streaming kernel (not the chunk reduction).
hipError_t gemv(const GemvPlan& p, const float* A, const double* x, double* part, double* y, const double* b,
                int mode, hipStream_t s, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
// out[k] = <V_k, w> for k < K (V column stride ldv), partial buffer of
// mdot_blocks(n) * K doubles.
int mdot_blocks(int64_t n);
hipError_t mdot(int64_t n, int K, const double* V, int64_t ldv, const double* w, double* part, double* out,
                hipStream_t s);
// w[r] += alpha * sum_k V_k[r] * h[k]
hipError_t axpy_multi(int64_t n, int K, const double* V, int64_t ldv, const double* h, double alpha, double* w,
                      hipStream_t s);
// y = alpha * x + beta * y
hipError_t axpby(int64_t n, double alpha, const double* x, double beta, double* y, hipStream_t s);
hipError_t fill(int64_t n, double value, double* y, hipStream_t s);
hipError_t nonfinite_count(int64_t n, const double* x, int* flag_dev, hipStream_t s);

}  // namespace rf
