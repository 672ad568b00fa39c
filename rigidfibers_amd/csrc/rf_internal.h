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
                                 const double* x0, double* apart, const PhysConsts& pc, hipStream_t s, int* groups,
                                 _Float16* T16 = nullptr, float* S16 = nullptr);
// With x0, the theta assembly also writes each group's off-diagonal product
// rows sum_{j in group} A_ij x0_j into apart[group len + row].  With T16 it
// also writes the FP16 copy of T with one FP32 scale per (target, source)
// block, S16[i * M + j] = max |T_ij| / 32768 (the GMRES-IR inner system).
// r = b - (x0 + sum_g apart[g]): the residual of x0 from those partials.
hipError_t launch_residual_groups(int groups, int64_t len, const double* apart, const double* b, const double* x0,
                                  double* r, hipStream_t s);
// Velocities from the stored operator of the last theta assembly (one pass
// over Tv): tvel, rvel [3*M] FP64.
hipError_t launch_velocities_stored(int N, int M, int64_t Lp, const float* Tv, const double* cpart, int groups,
                                    const double* ori, const double* x, const PhysConsts& pc, double* tvel,
                                    double* rvel, hipStream_t s);
// Dense columns of source fibers j0 .. j0 + jcount - 1 (jcount < 0: up to M)
// from the theta form into A (leading dimension n, source fiber j0's first
// column at A[0]).
hipError_t launch_expand_theta(int N, int M, int64_t Lp, const float* T, const double* ori, const PhysConsts& pc,
                               float* A, hipStream_t s, int j0 = 0, int jcount = -1);
// The whole dense A (n x n, leading dimension lda, the identity diagonal
// blocks included) from the theta form, as the LU's factor storage takes it.
hipError_t launch_expand_theta_dense(int N, int M, int64_t Lp, const float* T, const double* ori,
                                     const PhysConsts& pc, float* A, int64_t lda, hipStream_t s);
// The theta GEMV on the FP16 copy (rf_linalg.hip).
hipError_t gemv_theta16(int N, int M, int64_t Lp, const _Float16* T16, const float* S, const double* ori,
                        const double* x, double* y, const double* b, int mode, const PhysConsts& pc, hipStream_t s,
                        hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr, const int* skip = nullptr);
// y = A x (mode 0) or y = b - A x (mode 1) with A in theta form.  The
// products return at once when *skip (device memory) is set.
hipError_t gemv_theta(int N, int M, int64_t Lp, const float* T, const double* ori, const double* x, double* y,
                      const double* b, int mode, const PhysConsts& pc, hipStream_t s, hipEvent_t ev_begin = nullptr,
                      hipEvent_t ev_end = nullptr, const int* skip = nullptr);
hipError_t launch_velocities(int N, int NQI, bool analytical, int M, const double* pos, const double* ori,
                             const double* x, float* wf, double* fmono, double* vpart, double* tvel, double* rvel,
                             const PhysConsts& pc, hipStream_t s);
hipError_t launch_update(bool first, int M, double dt, const double* pos_prev, const double* pos,
                         const double* ori_prev, const double* ori, const double* tv_prev, const double* tv,
                         const double* rv_prev, const double* rv, double* pos_next, double* ori_next,
                         hipStream_t s);
bool assemble_supported(int N, int NQI);

// ---------------------------------------------------------------------------
// Batched launches of the serving engine (rf_serve.hip): one launch covers
// every admitted suspension.  Workgroups find their suspension through tables
// of int2 entries (suspension, fiber) or (suspension, 256-element block).

// Device pointers of one suspension for the current tick.
struct SusDev {
  int M, n;
  int64_t Lp;
  int vb_first, vb_count;          // its range in the vector-block table
  const double* pos;               // current state
  const double* ori;
  float* T;                        // theta-form system (groups of one: b written directly)
  float* Tv;                       // stored velocity operator
  double* cpart;                   // velocity constants [M][6]
  double* b;
  double* x;
  const double* x0;                // the start of this tick's solve: x when warm, else null (x = 0)
  double* npart;                   // [M][2]: per target |b|^2, |b - A x0|^2 (the assembly writes them)
  double* V;                       // Krylov basis, kdim x n; the assembly writes b - A x0 to V_0
  double* tv;                      // velocities of this step
  double* rv;
  const double* pos_prev;
  const double* ori_prev;
  const double* tv_prev;
  const double* rv_prev;
  double* pos_next;
  double* ori_next;
  int first;                       // Euler step
};

// One lockstep GMRES iteration of one suspension.
struct SusIter {
  int zero_x;                      // x = 0 first
  int ycount;                      // then x += V[0..ycount) ycoef
  const double* ycoef;
  double* svec;                    // then svec *= sval (nullptr: none)
  double sval;
  const double* gx;                // then gy = A gx (gmode 0) or b - A gx (1); nullptr: none
  double* gy;
  int gmode;
  double* w;                       // orthogonalised against V[0..K), then ||w||^2
  int K;
};

// Device-side GMRES(m) state of one suspension in the serving lockstep: the
// control step after each iteration forms the iteration's reductions from
// the block partials, advances this state and writes the next SusIter.
struct SusCtl {
  int phase;   // 3: start (V_0 = b - A x0 and the norms from the assembly), 0: true residual next,
               // 1: Arnoldi step next, 2: done
  int k;       // Arnoldi column of the current cycle
  int it;      // Arnoldi steps so far
  int converged;
  int nonfinite;
  int cold;    // start from x = 0 (no usable previous solution)
  double bnorm, res;
  double* H;   // (m + 1) x m Hessenberg matrix, column-major
  double* cs;  // m Givens cosines
  double* sn;  // m Givens sines
  double* g;   // m + 1 rotated right-hand side
  double* y;   // m coefficients of x += V y
};
hipError_t launch_gmres_control_batch(int ns, int m, int kdim, int nvb, double tol, int maxit, int confirm,
                                      const SusDev* sus, SusIter* it, SusCtl* ctl, const double* part, int* active,
                                      hipStream_t s);
hipError_t launch_orth_dots_batch(int nvb, int kdim, int phase, const SusDev* sus, const SusIter* it,
                                  const int2* vblocks, double* part, hipStream_t s);

hipError_t launch_assemble_theta_batch(int N, int NQI, bool analytical, int ntargets, const SusDev* sus,
                                       const int2* targets, const PhysConsts& pc, hipStream_t s);
hipError_t launch_gemv_theta_batch(int N, int ntargets, const SusDev* sus, const SusIter* it, const int2* targets,
                                   const PhysConsts& pc, hipStream_t s);
hipError_t launch_velocities_stored_batch(int N, int ntargets, const SusDev* sus, const int2* targets,
                                          const PhysConsts& pc, hipStream_t s);
hipError_t launch_update_batch(int nfblocks, const SusDev* sus, const int2* fblocks, double dt, hipStream_t s);
// x = 0 / x += V y / svec *= sval, per suspension (before the iteration's GEMV)
hipError_t launch_iter_prep_batch(int nvb, const SusDev* sus, const SusIter* it, const int2* vblocks,
                                  hipStream_t s);
// Block partials of <V_j, w>, j < K, into part[j nvb + block] (the first
// Gram-Schmidt reduction of an iteration; launch_orth_dots_batch forms the
// later ones).
hipError_t launch_dots_partial_batch(int nvb, int kmax, const SusDev* sus, const SusIter* it, const int2* vblocks,
                                    double* part, hipStream_t s);

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
                int mode, hipStream_t s, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr,
                const int* skip = nullptr);
// out[k] = <V_k, w> for k < K (V column stride ldv), partial buffer of
// mdot_blocks(n) * K doubles.
int mdot_blocks(int64_t n);
hipError_t mdot(int64_t n, int K, const double* V, int64_t ldv, const double* w, double* part, double* out,
                hipStream_t s);
// y = alpha * x + beta * y
hipError_t axpby(int64_t n, double alpha, const double* x, double beta, double* y, hipStream_t s);
hipError_t fill(int64_t n, double value, double* y, hipStream_t s);
hipError_t nonfinite_count(int64_t n, const double* x, int* flag_dev, hipStream_t s);

// One restart cycle of single-simulation GMRES(m) with its control on the
// device (rf_linalg.hip): the step that ends the cycle sets head.done (and
// *hdone in host memory), and the products and vector kernels of steps
// enqueued later return at once.
struct CycleHead {
  int done;       // the cycle has ended
  int k_used;     // its Arnoldi steps
  int nonfinite;  // a non-finite ||w|| ended it
  int pad;
  double scale;   // 1 / ||w|| of the last step: normalises V_{k+1}
  double res;     // |g_k|, the cycle's residual estimate
};
struct CycleCtl {
  CycleHead* head;
  double* H;  // (m + 1) x m Hessenberg matrix, column k at H + k (m + 1)
  double* cs;
  double* sn;
  double* g;  // m + 1
  double* y;  // m
  int* hdone;  // pinned host memory (device address): head.done mirrored
};
// Device doubles of a cycle's control (head included), and its layout in them.
size_t cycle_ctl_doubles(int m);
CycleCtl cycle_ctl(double* mem, int m);
// Dot partials of the steps: 3 (m + 1) mdot_blocks(n) doubles.
size_t cycle_part_doubles(int64_t n, int m);
// g = (beta, 0, ..., 0), head cleared.
hipError_t cycle_start(const CycleCtl& c, int m, double beta, hipStream_t s);
// Step k on w = A V_k (= V_{k+1}, V column stride n): Gram-Schmidt twice
// against V_0..V_k, the Hessenberg column and its Givens rotation, and the
// end test ||r|| <= tol ||b||, a breakdown ||w|| = 0 or k = klast, where
// y = H^-1 g is formed; else w /= ||w||.
hipError_t arnoldi_step(int64_t n, int k, int klast, int m, const double* V, double* w, double* part, double tol,
                        double bnorm, const CycleCtl& c, hipStream_t s);
// x += sum_{k < k_used} V_k y_k
hipError_t cycle_update(int64_t n, const double* V, const CycleCtl& c, double* x, hipStream_t s);

// BiCGStab with its scalars on the device (rf_linalg.hip), the same
// arithmetic as lib/viennacl/linalg/bicgstab.hpp:92-160 plus the test of
// ||s|| before the second product.  Iterations the host enqueues after the
// solve's end find head.done set and return at once.
struct BicgHead {
  int done;       // the solve has ended
  int converged;
  int nonfinite;  // ||s|| was not finite
  int its;        // iterations started
  int prods;      // products that ran (the iterations' v and t)
  int early_it;   // the iteration whose ||s|| met the tolerance (0: none)
  int restart;    // the next iteration restarts from the true residual
  int no_restart; // 1 unless a restart is pending and the solve goes on (the skip flag of the restart's kernels)
  int last_restart;  // iteration of the last restart (the start counts as one)
  int restarts;      // restarts so far
  double rho, alpha, omega, beta, res;
};
struct BicgCtl {
  BicgHead* head;
  int* hdone;  // pinned host memory (device address): head.done mirrored
};
struct BicgVecs {
  double* x;
  double* r;   // r0 follows r at stride n
  double* p;
  double* v;
  double* s;   // t follows s at stride n
};
size_t bicg_part_doubles(int64_t n);
hipError_t bicg_start(const BicgCtl& c, double rho, double res, hipStream_t s);
// A pending restart (lib/viennacl/linalg/bicgstab.hpp:119-131): after the
// caller's r = b - A x (a product skipped unless head.no_restart is 0),
// r0 = p = r and rho = ||r||^2; nothing when no restart is pending.
hipError_t bicg_restart(int64_t n, const BicgVecs& w, double* part, const BicgCtl& c, hipStream_t s);
// After v = A p of iteration it (1-based): alpha, s = r - alpha v, ||s||
// and the early end with x += alpha p.
hipError_t bicg_after_v(int64_t n, int it, const BicgVecs& w, double* part, double tol, double bnorm,
                        const BicgCtl& c, hipStream_t s);
// After t = A s: omega, x and r, ||r|| and <r0, r>, the end test, p.
hipError_t bicg_after_t(int64_t n, const BicgVecs& w, double* part, double tol, double bnorm, const BicgCtl& c,
                        hipStream_t s);

}  // namespace rf
