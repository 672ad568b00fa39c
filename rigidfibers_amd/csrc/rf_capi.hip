// rf_capi.hip -- the C ABI of include/rigidfibers.h: simulation state on the
// device, step orchestration (src/simulation.cu:286-329) and the iterative
// solvers (src/simulation.cu:401-445).
//
// Solvers.  GMRES(m) follows the reference semantics (x0 = 0, no
// preconditioner, restart m, stop on ||b - A x|| / ||b|| <= tol with the true
// residual recomputed at every restart and a converged cycle's estimate
// ending the solve: lib/viennacl/linalg/gmres.hpp:169-336;
// fortran/solve_system.incl:18-57, fortran/sgmres.f:1240-1358).  The Krylov
// basis, all vectors and the matvec accumulation are FP64 on the device (A
// stays FP32); the basis is orthogonalised with classical Gram-Schmidt
// applied twice, which spans the same space as the reference's
// MGS/Householder variants, and the (m+1) x m Hessenberg least-squares
// problem is solved on the device as well: the host stays one Arnoldi step
// ahead of the device and never waits for a step before enqueueing the next.
// BiCGStab follows
// lib/viennacl/linalg/bicgstab.hpp:92-160.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rigidfibers.h"
#include "rf_host.h"
#include "rf_internal.h"
#include "rf_lu.h"

namespace {

thread_local std::string g_last_error;

rf_status fail(rf_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define RF_HIP(call)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess)                                                                              \
      return fail(e_ == hipErrorOutOfMemory ? RF_ERR_OUT_OF_MEMORY : RF_ERR_DEVICE,                    \
                  std::string(#call) + ": " + hipGetErrorString(e_));                                  \
  } while (0)

#define RF_TRY(expr)             \
  do {                           \
    rf_status s_ = (expr);       \
    if (s_ != RF_OK) return s_;  \
  } while (0)

}  // namespace

struct rf_simulation {
  rf_config cfg{};
  int M = 0, N = 0, NQI = 0;
  int64_t n = 0, lda = 0;
  bool analytical = false, vel_analytical = false;
  rf::PhysConsts pc{};
  hipStream_t stream = nullptr;
  // phase boundaries of rf_step (0..4) and the assembly kernel's own (5, 6)
  hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int device = 0;

  double* pos[3] = {nullptr, nullptr, nullptr};
  double* ori[3] = {nullptr, nullptr, nullptr};
  double* tv[2] = {nullptr, nullptr};
  double* rv[2] = {nullptr, nullptr};
  int cur = 0, prev = 1, next = 2, vcur = 0, vprev = 1;
  int vlast = 0;  // buffer written by the last rf_update_velocities
  int plast = 0;  // buffer written by the last rf_update_fibers
  bool have_history = false;
  bool have_next = false;
  bool have_solution = false;  // x holds a previous solution (initial guess)
  bool have_previous = false;  // xprev holds the solution before it
  double* xprev = nullptr;     // RF_GUESS_EXTRAPOLATED only

  float* A = nullptr;     // dense system (direct solvers / RF_MATRIX_DENSE)
  bool theta = false;     // system held in theta form (T) instead of A
  float* T = nullptr;
  // orientation buffer of the last theta assembly (its dense blocks depend
  // on t_i), and whether that buffer still holds those orientations
  int tori = 0;
  bool tori_valid = false;
  int64_t Lp = 0;         // theta component stride
  int bgroups = 1;        // RHS partial slots written by the last assembly
  bool vstored = false;   // the theta assembly also stores the velocity operator
  bool vvalid = false;    // ... and it belongs to the current state
  float* Tv = nullptr;    // 12 component arrays of stride Lp per target fiber
  double* cpart = nullptr;
  double* b = nullptr;
  double* x = nullptr;
  double* bpart = nullptr;
  // first residual of a warm-started iterative solve, formed with the
  // assembly (rf_config.separate_residual == 0 on the theta form)
  double* apart = nullptr;   // off-diagonal products A x0, one slot per source group
  double* r0 = nullptr;      // b - A x0
  bool guess_taken = false;  // the assembly took this solve's initial guess
  bool guess_warm = false;
  bool r0_ready = false;     // r0 belongs to the current system and guess
  float* wf = nullptr;
  double* fm = nullptr;
  double* vpart = nullptr;
  _Float16* T16 = nullptr;  // GMRES-IR: FP16 copy of T
  float* S16 = nullptr;     // and its per-block scales

  // Krylov / refinement workspace
  int kdim = 0;           // basis vectors allocated
  double* V = nullptr;    // kdim x n
  double* work = nullptr; // 6 x n (BiCGStab vectors, residuals)
  double* gpart = nullptr;
  double* dpart = nullptr;
  double* hdev = nullptr;
  int* flag = nullptr;
  rf::CycleCtl cyc{};  // GMRES: the device control of a restart cycle
  rf::BicgCtl bicg{};  // BiCGStab: its scalars on the device
  int* hdone = nullptr;  // the done flag of either, mirrored into pinned host memory
  hipEvent_t cev[2] = {nullptr, nullptr};  // the ends of the last two steps
  rf::GemvPlan plan{};
  rf::LuWorkspace lu{};
  rf::LuResult lu_last{};  // factorisation statistics of the last direct solve
  bool direct_fallback = false;  // ... which GMRES finished
  int bicg_restarts = 0;         // BiCGStab: restarts in the last solve

  rf_solver_stats stats{};
  rf_timings timings{};
  // event pairs around the matvec kernels of one solve (timings.gemv)
  static constexpr int kGemvEvents = 256;
  hipEvent_t gev[2 * kGemvEvents] = {};
  int ngemv = 0;
  hipEvent_t hev[2 * kGemvEvents] = {};  // the same for the FP16 products of GMRES-IR
  int nhalf = 0;
  int64_t device_bytes = 0;
  std::vector<void*> allocations;

  template <typename T>
  rf_status alloc(T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(RF_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                            hipGetErrorString(e));
    }
    allocations.push_back(q);
    device_bytes += static_cast<int64_t>(bytes);
    *p = static_cast<T*>(q);
    return RF_OK;
  }
};

namespace {

rf_status validate_config(const rf_config* c) {
  if (!c) return fail(RF_ERR_INVALID_ARGUMENT, "config is NULL");
  if (c->number_of_fibers < 1) return fail(RF_ERR_INVALID_ARGUMENT, "NUMBER_OF_FIBERS must be >= 1");
  if (c->number_of_terms_in_force_expansion < 1 || c->number_of_terms_in_force_expansion > rf::kMaxTerms)
    return fail(RF_ERR_UNSUPPORTED, "NUMBER_OF_TERMS_IN_FORCE_EXPANSION must be in [1..8]");
  if (c->number_of_quadrature_points_per_interval != 3)
    return fail(RF_ERR_UNSUPPORTED, "NUMBER_OF_QUADRATURE_POINTS_PER_INTERVAL must be 3");
  if (!rf::assemble_supported(c->number_of_terms_in_force_expansion, c->number_of_quadrature_intervals))
    return fail(RF_ERR_UNSUPPORTED, "NUMBER_OF_QUADRATURE_INTERVALS must be 4, 8 or 16");
  if (!(c->slenderness > 0.0) || !std::isfinite(c->slenderness))
    return fail(RF_ERR_INVALID_ARGUMENT, "SLENDERNESS must be > 0");
  const double d = -std::log(c->slenderness * c->slenderness * std::exp(1.0));
  if (!(d > 2.0)) return fail(RF_ERR_INVALID_ARGUMENT, "SLENDERNESS too large: -ln(eps^2 e) must exceed 2");
  if (!std::isfinite(c->timestep)) return fail(RF_ERR_INVALID_ARGUMENT, "TIMESTEP must be finite");
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(c->external_force[a])) return fail(RF_ERR_INVALID_ARGUMENT, "external force must be finite");
  if (c->solver < RF_SOLVER_GMRES || c->solver > RF_SOLVER_GMRES_IR)
    return fail(RF_ERR_INVALID_ARGUMENT, "unknown solver");
  if (c->integration != RF_INTEGRATION_NUMERICAL && c->integration != RF_INTEGRATION_ANALYTICAL)
    return fail(RF_ERR_INVALID_ARGUMENT, "unknown integration");
  if (c->matrix_format < RF_MATRIX_AUTO || c->matrix_format > RF_MATRIX_THETA)
    return fail(RF_ERR_INVALID_ARGUMENT, "unknown matrix format");
  if (c->separate_residual != 0 && c->separate_residual != 1)
    return fail(RF_ERR_INVALID_ARGUMENT, "separate_residual must be 0 or 1");
  if (c->confirm_residual != 0 && c->confirm_residual != 1)
    return fail(RF_ERR_INVALID_ARGUMENT, "confirm_residual must be 0 or 1");
  if (c->initial_guess != RF_GUESS_ZERO && c->initial_guess != RF_GUESS_PREVIOUS &&
      c->initial_guess != RF_GUESS_EXTRAPOLATED)
    return fail(RF_ERR_INVALID_ARGUMENT, "unknown initial guess");
  if (c->velocity_integration < RF_VELOCITY_FOLLOW || c->velocity_integration > RF_VELOCITY_ANALYTICAL)
    return fail(RF_ERR_INVALID_ARGUMENT, "unknown velocity integration");
  if (c->velocity_operator != RF_VELOCITY_STORED && c->velocity_operator != RF_VELOCITY_PAIRWISE)
    return fail(RF_ERR_INVALID_ARGUMENT, "unknown velocity operator");
  if (c->solver == RF_SOLVER_GMRES_IR && !(c->inner_tolerance >= 0.0 && c->inner_tolerance < 1.0))
    return fail(RF_ERR_INVALID_ARGUMENT, "GMRES-IR inner tolerance must be in [0, 1)");
  // GMRES settings: its own, GMRES-IR's, and the direct solvers' fallback
  if (c->solver != RF_SOLVER_BICGSTAB) {
    if (c->gmres_restart < 1 || c->gmres_restart > 1024)
      return fail(RF_ERR_INVALID_ARGUMENT, "GMRES RESTART must be in [1..1024]");
    if (c->gmres_max_iterations < 1) return fail(RF_ERR_INVALID_ARGUMENT, "GMRES MAX_ITERATIONS must be >= 1");
    if (!(c->gmres_tolerance > 0.0)) return fail(RF_ERR_INVALID_ARGUMENT, "GMRES TOLERANCE must be > 0");
  }
  if (c->solver == RF_SOLVER_BICGSTAB) {
    if (c->bicgstab_max_iterations < 1) return fail(RF_ERR_INVALID_ARGUMENT, "BICGSTAB MAX_ITERATIONS must be >= 1");
    if (!(c->bicgstab_tolerance > 0.0)) return fail(RF_ERR_INVALID_ARGUMENT, "BICGSTAB TOLERANCE must be > 0");
  }
  if (c->solver == RF_SOLVER_DIRECT_MIXED_BF16 || c->solver == RF_SOLVER_DIRECT_MIXED_FP16) {
    if (!(c->refinement_tolerance > 0.0)) return fail(RF_ERR_INVALID_ARGUMENT, "refinement tolerance must be > 0");
    if (c->refinement_max_iterations < 1)
      return fail(RF_ERR_INVALID_ARGUMENT, "refinement max iterations must be >= 1");
  }
  return RF_OK;
}

void make_consts(const rf_config& c, rf::PhysConsts* pc) {
  std::memset(pc, 0, sizeof(*pc));
  // fortran/fibers.f95:141-145 and :353-440 (lambda, eigen in FP64, then FP32)
  const double eps = c.slenderness;
  const double d = -std::log(eps * eps * std::exp(1.0));
  const double e = 2.0;
  double lambda = 2.0;
  for (int k = 0; k < c.number_of_terms_in_force_expansion; ++k) {
    if (k > 0) lambda += 2.0 / (k + 1);
    const double eigen = ((d - e - lambda) / 2.0) / (d - lambda);
    pc->eigen[k] = static_cast<float>(eigen);
    // gamma_k = 0.5 (2(k+1) + 1) / (d + e - lambda_k), fortran/assemble_matrix.incl:101
    pc->gamma[k] = (k == 0) ? 0.0f : static_cast<float>(0.5 * (2.0 * (k + 1) + 1.0) / (d + e - lambda));
  }
  pc->D1 = static_cast<float>(0.75 / (d - 2.0));
  pc->eps2 = static_cast<float>(eps * eps);
  pc->eps2_d = eps * eps;
  pc->d = d;
  for (int a = 0; a < 3; ++a) {
    pc->Fh_d[a] = 0.5 * c.external_force[a];
    pc->Fh[a] = static_cast<float>(pc->Fh_d[a]);
  }
}

// Krylov basis vectors: GMRES and GMRES-IR, and the direct solvers' GMRES
// fallback.
int solver_kdim(const rf_config& c) { return c.solver == RF_SOLVER_BICGSTAB ? 0 : c.gmres_restart + 1; }

// y = A x or y = b - A x on the simulation's stream, timed with an event pair
// while the pool lasts.
hipError_t sim_gemv(rf_simulation* s, const double* x, double* part, double* y, const double* b, int mode,
                    bool half = false, const int* skip = nullptr) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (half) {
    if (s->nhalf < rf_simulation::kGemvEvents) {
      e0 = s->hev[2 * s->nhalf];
      e1 = s->hev[2 * s->nhalf + 1];
      ++s->nhalf;
    }
    return rf::gemv_theta16(s->N, s->M, s->Lp, s->T16, s->S16, s->ori[s->cur], x, y, b, mode, s->pc, s->stream, e0,
                            e1, skip);
  }
  if (s->ngemv < rf_simulation::kGemvEvents) {
    e0 = s->gev[2 * s->ngemv];
    e1 = s->gev[2 * s->ngemv + 1];
    ++s->ngemv;
  }
  if (s->theta)
    return rf::gemv_theta(s->N, s->M, s->Lp, s->T, s->ori[s->cur], x, y, b, mode, s->pc, s->stream, e0, e1, skip);
  return rf::gemv(s->plan, s->A, x, part, y, b, mode, s->stream, e0, e1, skip);
}

rf_status check_nonfinite(rf_simulation* s, const double* v, int64_t len, const char* what) {
  RF_HIP(hipMemsetAsync(s->flag, 0, sizeof(int), s->stream));
  RF_HIP(rf::nonfinite_count(len, v, s->flag, s->stream));
  int h = 0;
  RF_HIP(hipMemcpyAsync(&h, s->flag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  RF_HIP(hipStreamSynchronize(s->stream));
  if (h) return fail(RF_ERR_NONFINITE, std::string("non-finite values in ") + what);
  return RF_OK;
}

// ||v||_2 on the device, returned to the host (synchronises).
rf_status device_norm(rf_simulation* s, const double* v, double* out) {
  RF_HIP(rf::mdot(s->n, 1, v, s->n, v, s->dpart, s->hdev, s->stream));
  double h = 0.0;
  RF_HIP(hipMemcpyAsync(&h, s->hdev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  RF_HIP(hipStreamSynchronize(s->stream));
  *out = std::sqrt(h);
  return RF_OK;
}


// Restarted GMRES(m), classical Gram-Schmidt twice, FP64 basis.
// GMRES(m) for A x = b: A the FP32 system, or its FP16 copy (half).  Starts
// from x when warm, else from 0; stops at ||b - A x|| <= tol ||b|| or after
// maxit iterations.  Each restart cycle begins with the true residual; a
// cycle whose own estimate |g_k| meets the tolerance ends the solve after
// its update of x (lib/viennacl/linalg/gmres.hpp:330-335), unless confirm
// asks for the true residual of that x first (fortran/sgmres.f with
// icntl(8) = 1).  r0, when given, is b - A x of the warm start.  *out
// receives the iterations, convergence and residual.
rf_status gmres_core(rf_simulation* s, const double* b, double* x, bool warm, double tol, int maxit, bool half,
                     rf_solver_stats* out, const double* r0 = nullptr) {
  const int64_t n = s->n;
  const int m = s->cfg.gmres_restart;
  // GMRES-IR's inner solves leave the true residual to the outer loop
  const bool confirm = !half && s->cfg.confirm_residual;
  double* V = s->V;
  auto Vk = [&](int k) { return V + static_cast<int64_t>(k) * n; };

  double bnorm = 0.0;
  RF_TRY(device_norm(s, b, &bnorm));
  if (!warm) RF_HIP(rf::fill(n, 0.0, x, s->stream));
  *out = rf_solver_stats{0, 0, 0.0};
  if (!std::isfinite(bnorm)) return fail(RF_ERR_NONFINITE, "non-finite right-hand side");
  if (bnorm == 0.0) {
    RF_HIP(rf::fill(n, 0.0, x, s->stream));
    out->converged = 1;
    return RF_OK;
  }
  int it = 0;
  bool first = true, converged = false;
  double res = bnorm;
  while (true) {
    double beta;
    if (first && !warm) {
      RF_HIP(rf::axpby(n, 1.0, b, 0.0, Vk(0), s->stream));
      beta = bnorm;
      first = false;
    } else if (first && r0) {
      first = false;
      RF_HIP(rf::axpby(n, 1.0, r0, 0.0, Vk(0), s->stream));
      RF_TRY(device_norm(s, Vk(0), &beta));
    } else {
      first = false;
      RF_HIP(sim_gemv(s, x, s->gpart, Vk(0), b, 1, half));
      RF_TRY(device_norm(s, Vk(0), &beta));
    }
    res = beta;
    if (!std::isfinite(beta)) return fail(RF_ERR_NONFINITE, "non-finite residual in GMRES");
    if (beta / bnorm <= tol) {
      converged = true;
      break;
    }
    if (it >= maxit) break;
    RF_HIP(rf::axpby(n, 1.0 / beta, Vk(0), 0.0, Vk(0), s->stream));
    // The control of the cycle runs on the device; the host stays one step
    // ahead: it enqueues step k + 1, then waits for the end of step k and
    // reads the done flag that step's control mirrors into host memory.  The
    // product and vector kernels of a step enqueued after the cycle's end
    // return at once (at most one per cycle).
    const int steps = std::min(m, maxit - it);
    int& launched = half ? s->nhalf : s->ngemv;
    const int launched0 = launched;
    RF_HIP(rf::cycle_start(s->cyc, m, beta, s->stream));
    for (int k = 0; k < steps; ++k) {
      RF_HIP(sim_gemv(s, Vk(k), s->gpart, Vk(k + 1), nullptr, 0, half, &s->cyc.head->done));
      RF_HIP(rf::arnoldi_step(n, k, steps - 1, m, V, Vk(k + 1), s->dpart, tol, bnorm, s->cyc, s->stream));
      RF_HIP(hipEventRecord(s->cev[k & 1], s->stream));
      if (k > 0) {
        RF_HIP(hipEventSynchronize(s->cev[(k - 1) & 1]));
        if (*static_cast<volatile int*>(s->hdone)) break;
      }
    }
    RF_HIP(rf::cycle_update(n, V, s->cyc, x, s->stream));
    rf::CycleHead hd{};
    RF_HIP(hipMemcpyAsync(&hd, s->cyc.head, sizeof(hd), hipMemcpyDeviceToHost, s->stream));
    RF_HIP(hipStreamSynchronize(s->stream));
    // products that returned at once are not counted (nor timed)
    launched = std::min(launched, launched0 + hd.k_used + hd.nonfinite);
    it += hd.k_used + hd.nonfinite;
    if (hd.nonfinite) return fail(RF_ERR_NONFINITE, "non-finite Arnoldi vector in GMRES");
    res = hd.res;
    if (!confirm && res / bnorm <= tol) {
      converged = true;
      break;
    }
    // the loop head recomputes the true residual b - A x
  }
  out->iterations = it;
  out->converged = converged ? 1 : 0;
  out->relative_residual = res / bnorm;
  return RF_OK;
}

// Initial guess of the iterative solvers in s->x, which holds the previous
// solution x_n when there is one: PREVIOUS keeps it, EXTRAPOLATED replaces it
// by 2 x_n - x_{n-1} once x_{n-1} is known and keeps x_n for the next solve.
// *warm says whether x is a guess (else the solver starts from zero).
rf_status initial_guess(rf_simulation* s, bool* warm) {
  *warm = s->cfg.initial_guess != RF_GUESS_ZERO && s->have_solution;
  if (*warm && s->cfg.initial_guess == RF_GUESS_EXTRAPOLATED) {
    const int64_t n = s->n;
    if (s->have_previous) {
      RF_HIP(rf::axpby(n, 1.0, s->x, -1.0, s->xprev, s->stream));  // d = x_n - x_{n-1}
      RF_HIP(rf::axpby(n, 1.0, s->xprev, 1.0, s->x, s->stream));   // x = x_n + d
      RF_HIP(rf::axpby(n, 1.0, s->x, -1.0, s->xprev, s->stream));  // xprev = x - d = x_n
    } else {
      RF_HIP(hipMemcpyAsync(s->xprev, s->x, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    }
    s->have_previous = true;
  }
  return RF_OK;
}

// The initial guess of an iterative solve: taken by the assembly when it
// formed the first residual with the system (*r0 then points at it), else
// here.
rf_status take_guess(rf_simulation* s, bool* warm, const double** r0) {
  *r0 = nullptr;
  if (s->guess_taken) {
    *warm = s->guess_warm;
    if (s->r0_ready && s->guess_warm) *r0 = s->r0;
    s->guess_taken = false;
    s->r0_ready = false;
    return RF_OK;
  }
  return initial_guess(s, warm);
}

rf_status solve_gmres(rf_simulation* s) {
  bool warm = false;
  const double* r0 = nullptr;
  RF_TRY(take_guess(s, &warm, &r0));
  return gmres_core(s, s->b, s->x, warm, s->cfg.gmres_tolerance, s->cfg.gmres_max_iterations, false, &s->stats,
                    r0);
}

// GMRES-IR: x <- x + d with d from GMRES on the FP16 copy, A16 d = r, each
// outer residual r = b - A x from the FP32 system in FP64, until
// ||r|| <= tol ||b||.  The inner solves stop at inner_tolerance relative to
// ||r|| (no tighter than the outer target needs).
rf_status solve_gmres_ir(rf_simulation* s) {
  const int64_t n = s->n;
  const double tol = s->cfg.gmres_tolerance;
  const double itol = s->cfg.inner_tolerance > 0.0 ? s->cfg.inner_tolerance : 1e-4;
  const int maxit = s->cfg.gmres_max_iterations;
  double* r = s->work;
  double* d = s->work + n;
  double bnorm = 0.0;
  RF_TRY(device_norm(s, s->b, &bnorm));
  bool warm = false;
  const double* r0 = nullptr;
  RF_TRY(take_guess(s, &warm, &r0));
  if (!warm || bnorm == 0.0) RF_HIP(rf::fill(n, 0.0, s->x, s->stream));
  s->stats = rf_solver_stats{0, 0, 0.0};
  if (!std::isfinite(bnorm)) return fail(RF_ERR_NONFINITE, "non-finite right-hand side");
  if (bnorm == 0.0) {
    s->stats.converged = 1;
    return RF_OK;
  }
  int it = 0;
  while (true) {
    if (r0)
      RF_HIP(rf::axpby(n, 1.0, r0, 0.0, r, s->stream));
    else
      RF_HIP(sim_gemv(s, s->x, s->gpart, r, s->b, 1, false));
    r0 = nullptr;
    double res = 0.0;
    RF_TRY(device_norm(s, r, &res));
    if (!std::isfinite(res)) return fail(RF_ERR_NONFINITE, "non-finite residual in GMRES-IR");
    s->stats.relative_residual = res / bnorm;
    if (res / bnorm <= tol) {
      s->stats.converged = 1;
      break;
    }
    if (it >= maxit) break;
    rf_solver_stats inner{};
    const double target = std::max(itol, 0.5 * tol * bnorm / res);
    RF_TRY(gmres_core(s, r, d, false, target, maxit - it, true, &inner));
    it += std::max(inner.iterations, 1);
    RF_HIP(rf::axpby(n, 1.0, d, 1.0, s->x, s->stream));
  }
  s->stats.iterations = it;
  return RF_OK;
}

// Unpreconditioned BiCGStab, x0 = 0 (lib/viennacl/linalg/bicgstab.hpp:92-160).
rf_status solve_bicgstab(rf_simulation* s) {
  const int64_t n = s->n;
  s->bicg_restarts = 0;
  const int maxit = s->cfg.bicgstab_max_iterations;
  const double tol = s->cfg.bicgstab_tolerance;
  double* r = s->work;
  double* r0 = s->work + n;
  double* p = s->work + 2 * n;
  double* v = s->work + 3 * n;
  double* sv = s->work + 4 * n;
  double* t = s->work + 5 * n;
  double bnorm = 0.0;
  RF_TRY(device_norm(s, s->b, &bnorm));
  bool warm = false;
  const double* r0x = nullptr;
  RF_TRY(take_guess(s, &warm, &r0x));
  s->stats = rf_solver_stats{0, 0, 0.0};
  if (!std::isfinite(bnorm)) return fail(RF_ERR_NONFINITE, "non-finite right-hand side");
  if (bnorm == 0.0 || !warm) RF_HIP(rf::fill(n, 0.0, s->x, s->stream));
  if (bnorm == 0.0) {
    s->stats.converged = 1;
    return RF_OK;
  }
  double rho, res;
  if (warm) {
    if (r0x)
      RF_HIP(rf::axpby(n, 1.0, r0x, 0.0, r, s->stream));
    else
      RF_HIP(sim_gemv(s, s->x, s->gpart, r, s->b, 1));
    RF_TRY(device_norm(s, r, &res));
    rho = res * res;
  } else {
    RF_HIP(rf::axpby(n, 1.0, s->b, 0.0, r, s->stream));
    rho = bnorm * bnorm;
    res = bnorm;
  }
  RF_HIP(rf::axpby(n, 1.0, r, 0.0, r0, s->stream));
  RF_HIP(rf::axpby(n, 1.0, r, 0.0, p, s->stream));
  int it = 0;
  bool converged = res / bnorm <= tol;
  if (!converged && maxit > 0) {
    // The scalars live on the device; the host stays one iteration ahead (it
    // enqueues iteration k + 1, then waits for the end of iteration k and reads
    // the done flag mirrored into host memory), and the kernels of an
    // iteration enqueued after the end return at once.
    const rf::BicgVecs w{s->x, r, p, v, sv};
    const int launched0 = s->ngemv;
    RF_HIP(rf::bicg_start(s->bicg, rho, res, s->stream));
    for (int k = 0; k < maxit; ++k) {
      if (k > 0) {  // a pending restart: r = b - A x (untimed), r0 = p = r, rho = ||r||^2
        if (s->theta)
          RF_HIP(rf::gemv_theta(s->N, s->M, s->Lp, s->T, s->ori[s->cur], s->x, r, s->b, 1, s->pc, s->stream,
                                nullptr, nullptr, &s->bicg.head->no_restart));
        else
          RF_HIP(rf::gemv(s->plan, s->A, s->x, s->gpart, r, s->b, 1, s->stream, nullptr, nullptr,
                          &s->bicg.head->no_restart));
        RF_HIP(rf::bicg_restart(n, w, s->dpart, s->bicg, s->stream));
      }
      RF_HIP(sim_gemv(s, p, s->gpart, v, nullptr, 0, false, &s->bicg.head->done));
      RF_HIP(rf::bicg_after_v(n, k + 1, w, s->dpart, tol, bnorm, s->bicg, s->stream));
      RF_HIP(sim_gemv(s, sv, s->gpart, t, nullptr, 0, false, &s->bicg.head->done));
      RF_HIP(rf::bicg_after_t(n, w, s->dpart, tol, bnorm, s->bicg, s->stream));
      RF_HIP(hipEventRecord(s->cev[k & 1], s->stream));
      if (k > 0) {
        RF_HIP(hipEventSynchronize(s->cev[(k - 1) & 1]));
        if (*static_cast<volatile int*>(s->hdone)) break;
      }
    }
    rf::BicgHead hd{};
    RF_HIP(hipMemcpyAsync(&hd, s->bicg.head, sizeof(hd), hipMemcpyDeviceToHost, s->stream));
    RF_HIP(hipStreamSynchronize(s->stream));
    s->ngemv = std::min(s->ngemv, launched0 + hd.prods);  // products that returned at once are not counted
    s->bicg_restarts = hd.restarts;
    if (hd.nonfinite) return fail(RF_ERR_NONFINITE, "non-finite residual in BiCGStab");
    it = hd.its;
    converged = hd.converged != 0;
    res = hd.res;
  }
  // A converged solve reports its recurrence residual, as the reference's
  // solver does (lib/viennacl/linalg/bicgstab.hpp:146-148,160); the true
  // residual of the returned x when it did not converge, or on request.
  if (!converged || s->cfg.confirm_residual) {
    RF_HIP(sim_gemv(s, s->x, s->gpart, r, s->b, 1));
    RF_TRY(device_norm(s, r, &res));
    converged = converged && res / bnorm <= 10.0 * tol;
  }
  s->stats.iterations = it;
  s->stats.converged = converged ? 1 : 0;
  s->stats.relative_residual = res / bnorm;
  return RF_OK;
}

rf_status solve_direct(rf_simulation* s) {
  const int mode = s->cfg.solver;
  rf::LuResult r{};
  rf::LuSystem sys;
  const int64_t n = s->n;
  if (s->theta) {
    sys.fill = [s](float* F, int64_t lda, hipStream_t st) {
      return rf::launch_expand_theta_dense(s->N, s->M, s->Lp, s->T, s->ori[s->tori], s->pc, F, lda, st);
    };
  } else {
    sys.fill = [s, n](float* F, int64_t lda, hipStream_t st) {
      return hipMemcpy2DAsync(F, lda * sizeof(float), s->A, s->lda * sizeof(float), n * sizeof(float), n,
                              hipMemcpyDeviceToDevice, st);
    };
  }
  sys.residual = [s](const double* x, double* r, hipStream_t) { return sim_gemv(s, x, s->gpart, r, s->b, 1); };
  hipError_t e = rf::lu_solve(s->lu, mode == RF_SOLVER_DIRECT ? rf::LuMode::kFP32
                                     : mode == RF_SOLVER_DIRECT_MIXED_BF16 ? rf::LuMode::kBF16
                                                                          : rf::LuMode::kFP16,
                              sys, s->n, s->b, s->x, s->work, s->dpart, s->hdev, s->cfg.refinement_tolerance,
                              s->cfg.refinement_max_iterations, s->stream, &r);
  if (e != hipSuccess) return fail(RF_ERR_DEVICE, std::string("LU: ") + hipGetErrorString(e));
  s->lu_last = r;
  s->stats.iterations = r.refinement_iterations;
  s->stats.relative_residual = r.relative_residual;
  s->stats.converged = r.converged ? 1 : 0;
  s->direct_fallback = false;
  if (r.info == 0 && r.converged) return RF_OK;
  // The pivots are chosen within 64-row diagonal blocks; when that choice
  // leaves the factors too inaccurate for the refinement (an exact zero
  // pivot, or element growth), GMRES on the dense system finishes the solve
  // from the refined x (from zero after a zero pivot) to the refinement
  // tolerance, so the step never returns an unverified x.
  const bool warm = r.info == 0 && std::isfinite(r.relative_residual);
  rf_solver_stats st{};
  RF_TRY(gmres_core(s, s->b, s->x, warm, s->cfg.refinement_tolerance, s->cfg.gmres_max_iterations, false, &st));
  s->direct_fallback = true;
  s->stats.iterations = r.refinement_iterations + st.iterations;
  s->stats.relative_residual = st.relative_residual;
  s->stats.converged = st.converged;
  if (!std::isfinite(st.relative_residual)) return fail(RF_ERR_NONFINITE, "non-finite solution in the direct solve");
  return RF_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0f;
  return ms;
}

}  // namespace

// Staging buffer of rf_get_system_noalloc, which may not allocate device
// memory: 16 MiB of file-scope device memory that each GPU the library runs
// on holds from the library's load until the process ends.  Host code uses
// it only through the address hipGetSymbolAddress returns for the current
// device (the array's own name is a host-side stand-in).  Every call on a
// GPU shares that GPU's copy, so calls that overlap on one GPU race.
constexpr size_t kScratchFloats = size_t{4} << 20;
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current; looked up
// once per device.
rf_status scratch_address(int device, float** out) {
  if (device < 0 || device >= kMaxDevices) return fail(RF_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(RF_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_ptr[device] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[device];
  return RF_OK;
}

// Makes a device current for a scope and restores the caller's afterwards.
struct DeviceScope {
  int saved = -1;
  bool switched = false;
  rf_status enter(int device) {
    RF_HIP(hipGetDevice(&saved));
    if (saved != device) {
      RF_HIP(hipSetDevice(device));
      switched = true;
    }
    return RF_OK;
  }
  ~DeviceScope() {
    if (switched) (void)hipSetDevice(saved);
  }
};

}  // namespace

namespace rf {
rf_status set_error(rf_status code, const std::string& msg) { return fail(code, msg); }
rf_status check_config(const rf_config* c) { return validate_config(c); }
void physical_constants(const rf_config& c, PhysConsts* pc) { make_consts(c, pc); }
}  // namespace rf

extern "C" {

rf_status rf_config_init(rf_config* cfg) {
  if (!cfg) return fail(RF_ERR_INVALID_ARGUMENT, "config is NULL");
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->number_of_fibers = 100;
  cfg->number_of_terms_in_force_expansion = 5;
  cfg->number_of_quadrature_intervals = 8;
  cfg->number_of_quadrature_points_per_interval = 3;
  cfg->timestep = 0.1;
  cfg->slenderness = 0.01;
  cfg->external_force[0] = 0.0;
  cfg->external_force[1] = 0.0;
  cfg->external_force[2] = -1.0;
  cfg->solver = RF_SOLVER_GMRES;
  cfg->integration = RF_INTEGRATION_NUMERICAL;
  cfg->velocity_integration = RF_VELOCITY_FOLLOW;
  cfg->gmres_restart = 10;
  cfg->gmres_max_iterations = 1000;
  cfg->gmres_tolerance = 1e-5;
  cfg->bicgstab_max_iterations = 1000;
  cfg->bicgstab_tolerance = 1e-5;
  cfg->refinement_tolerance = 1e-10;
  cfg->refinement_max_iterations = 50;
  cfg->initial_guess = RF_GUESS_PREVIOUS;
  return RF_OK;
}

rf_status rf_create(const rf_config* cfg, const double* positions, const double* orientations, rf_simulation** out) {
  if (!out) return fail(RF_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  RF_TRY(validate_config(cfg));
  if (!positions || !orientations) return fail(RF_ERR_INVALID_ARGUMENT, "positions/orientations are NULL");
  const int M = cfg->number_of_fibers;
  for (int64_t k = 0; k < 3 * static_cast<int64_t>(M); ++k)
    if (!std::isfinite(positions[k]) || !std::isfinite(orientations[k]))
      return fail(RF_ERR_INVALID_ARGUMENT, "non-finite initial state");
  rf_simulation* s = new (std::nothrow) rf_simulation();
  if (!s) return fail(RF_ERR_OUT_OF_MEMORY, "host allocation failed");
  auto cleanup = [&](rf_status st) {
    rf_destroy(s);
    return st;
  };
  s->cfg = *cfg;
  s->M = M;
  s->N = cfg->number_of_terms_in_force_expansion;
  s->NQI = cfg->number_of_quadrature_intervals;
  s->n = 3LL * s->N * M;
  s->lda = (s->n + 63) / 64 * 64;
  s->analytical = cfg->integration == RF_INTEGRATION_ANALYTICAL;
  s->vel_analytical = (cfg->velocity_integration == RF_VELOCITY_FOLLOW) ? s->analytical
                                                                        : cfg->velocity_integration == RF_VELOCITY_ANALYTICAL;
  make_consts(*cfg, &s->pc);
  if (hipGetDevice(&s->device) != hipSuccess) return cleanup(fail(RF_ERR_DEVICE, "no HIP device"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(RF_ERR_DEVICE, "hipStreamCreate failed"));
  for (auto& e : s->ev)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(RF_ERR_DEVICE, "hipEventCreate failed"));
  for (auto& e : s->gev)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(RF_ERR_DEVICE, "hipEventCreate failed"));
  for (auto& e : s->hev)
    if (hipEventCreate(&e) != hipSuccess) return cleanup(fail(RF_ERR_DEVICE, "hipEventCreate failed"));
  for (auto& e : s->cev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return cleanup(fail(RF_ERR_DEVICE, "hipEventCreate failed"));

  const int64_t n = s->n;
  const size_t fiber3 = 3 * static_cast<size_t>(M);
  rf_status st = RF_OK;
  for (int k = 0; k < 3 && st == RF_OK; ++k) {
    st = s->alloc(&s->pos[k], fiber3);
    if (st == RF_OK) st = s->alloc(&s->ori[k], fiber3);
  }
  for (int k = 0; k < 2 && st == RF_OK; ++k) {
    st = s->alloc(&s->tv[k], fiber3);
    if (st == RF_OK) st = s->alloc(&s->rv[k], fiber3);
  }
  const bool iterative =
      cfg->solver == RF_SOLVER_GMRES || cfg->solver == RF_SOLVER_BICGSTAB || cfg->solver == RF_SOLVER_GMRES_IR;
  if (cfg->solver == RF_SOLVER_GMRES_IR && cfg->matrix_format == RF_MATRIX_DENSE)
    return cleanup(fail(RF_ERR_UNSUPPORTED, "GMRES-IR needs the theta form"));
  // The theta form unless the dense one is asked for: the direct solvers
  // expand it into their factor storage and refine with the theta GEMV.
  s->theta = cfg->matrix_format != RF_MATRIX_DENSE;
  if (s->theta) {
    const int NC = 3 + 6 * (s->N - 1);
    s->Lp = rf::theta_stride(M, s->N);
    if (st == RF_OK) st = s->alloc(&s->T, static_cast<size_t>(M) * NC * static_cast<size_t>(s->Lp));
    s->vstored = cfg->velocity_operator == RF_VELOCITY_STORED && s->vel_analytical == s->analytical;
    if (s->vstored && st == RF_OK) st = s->alloc(&s->Tv, static_cast<size_t>(M) * 12 * static_cast<size_t>(s->Lp));
    if (s->vstored && st == RF_OK)
      st = s->alloc(&s->cpart, static_cast<size_t>(rf::max_pair_groups(M)) * M * 6);
    if (cfg->solver == RF_SOLVER_GMRES_IR) {
      if (st == RF_OK) st = s->alloc(&s->T16, static_cast<size_t>(M) * NC * static_cast<size_t>(s->Lp));
      if (st == RF_OK) st = s->alloc(&s->S16, static_cast<size_t>(M) * static_cast<size_t>(M));
    }
  } else if (st == RF_OK) {
    st = s->alloc(&s->A, static_cast<size_t>(s->lda) * static_cast<size_t>(n));
  }
  if (st == RF_OK) st = s->alloc(&s->b, static_cast<size_t>(n));
  if (st == RF_OK) st = s->alloc(&s->x, static_cast<size_t>(n));
  if (st == RF_OK && cfg->initial_guess == RF_GUESS_EXTRAPOLATED) st = s->alloc(&s->xprev, static_cast<size_t>(n));
  if (st == RF_OK) st = s->alloc(&s->bpart, static_cast<size_t>(rf::max_pair_groups(M)) * n);
  if (s->theta && iterative && !cfg->separate_residual) {
    if (st == RF_OK) st = s->alloc(&s->apart, static_cast<size_t>(rf::max_pair_groups(M)) * n);
    if (st == RF_OK) st = s->alloc(&s->r0, static_cast<size_t>(n));
  }
  if (st == RF_OK) st = s->alloc(&s->vpart, static_cast<size_t>(rf::max_pair_groups(M)) * M * 6);
  if (st == RF_OK) st = s->alloc(&s->wf, static_cast<size_t>(M) * 3 * 3 * s->NQI);
  if (st == RF_OK) st = s->alloc(&s->fm, static_cast<size_t>(M) * 3 * (s->N + 1));
  s->plan = rf::gemv_plan(n, s->lda);
  if (st == RF_OK) st = s->alloc(&s->gpart, rf::gemv_partial_elems(s->plan));
  s->kdim = solver_kdim(*cfg);
  if (st == RF_OK && s->kdim > 0) st = s->alloc(&s->V, static_cast<size_t>(s->kdim) * n);
  if (st == RF_OK) st = s->alloc(&s->work, 6 * static_cast<size_t>(n));
  const int kmax = (s->kdim > 2 ? s->kdim : 2) + 1;
  size_t dpart_len = static_cast<size_t>(rf::mdot_blocks(n)) * kmax;
  const bool bicg = cfg->solver == RF_SOLVER_BICGSTAB;
  if (s->kdim > 0 || bicg) {
    if (s->kdim > 0) {
      const int m = cfg->gmres_restart;
      dpart_len = std::max(dpart_len, rf::cycle_part_doubles(n, m));
      double* mem = nullptr;
      if (st == RF_OK) st = s->alloc(&mem, rf::cycle_ctl_doubles(m));
      if (st == RF_OK) s->cyc = rf::cycle_ctl(mem, m);
    } else {
      dpart_len = std::max(dpart_len, rf::bicg_part_doubles(n));
      double* mem = nullptr;
      if (st == RF_OK) st = s->alloc(&mem, (sizeof(rf::BicgHead) + sizeof(double) - 1) / sizeof(double));
      if (st == RF_OK) s->bicg.head = reinterpret_cast<rf::BicgHead*>(mem);
    }
    void* hp = nullptr;
    void* dp = nullptr;
    if (st == RF_OK && (hipHostMalloc(&hp, 64, hipHostMallocMapped) != hipSuccess ||
                        hipHostGetDevicePointer(&dp, hp, 0) != hipSuccess)) {
      if (hp) (void)hipHostFree(hp);
      (void)hipGetLastError();
      st = fail(RF_ERR_OUT_OF_MEMORY, "hipHostMalloc of the solver's done flag failed");
    } else if (st == RF_OK) {
      s->hdone = static_cast<int*>(hp);
      *s->hdone = 0;
      s->cyc.hdone = static_cast<int*>(dp);
      s->bicg.hdone = static_cast<int*>(dp);
    }
  }
  if (st == RF_OK) st = s->alloc(&s->dpart, dpart_len);
  if (st == RF_OK) st = s->alloc(&s->hdev, static_cast<size_t>(2 * kmax + 4));
  if (st == RF_OK) st = s->alloc(&s->flag, 1);
  if (st != RF_OK) return cleanup(st);
  if (cfg->solver == RF_SOLVER_DIRECT || cfg->solver == RF_SOLVER_DIRECT_MIXED_BF16 ||
      cfg->solver == RF_SOLVER_DIRECT_MIXED_FP16) {
    const rf::LuMode mode = cfg->solver == RF_SOLVER_DIRECT ? rf::LuMode::kFP32
                            : cfg->solver == RF_SOLVER_DIRECT_MIXED_BF16 ? rf::LuMode::kBF16
                                                                          : rf::LuMode::kFP16;
    hipError_t e = rf::lu_workspace_create(mode, n, s->lda, &s->lu);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return cleanup(fail(e == hipErrorOutOfMemory ? RF_ERR_OUT_OF_MEMORY : RF_ERR_DEVICE,
                          std::string("LU workspace: ") + hipGetErrorString(e)));
    }
    s->device_bytes += rf::lu_workspace_bytes(s->lu);
  }
  if (hipMemcpyAsync(s->pos[s->cur], positions, fiber3 * sizeof(double), hipMemcpyHostToDevice, s->stream) !=
          hipSuccess ||
      hipMemcpyAsync(s->ori[s->cur], orientations, fiber3 * sizeof(double), hipMemcpyHostToDevice, s->stream) !=
          hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess)
    return cleanup(fail(RF_ERR_DEVICE, "initial upload failed"));
  *out = s;
  return RF_OK;
}

void rf_destroy(rf_simulation* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  rf::lu_workspace_destroy(&s->lu);
  for (void* p : s->allocations) (void)hipFree(p);
  for (auto& e : s->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : s->gev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : s->hev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : s->cev)
    if (e) (void)hipEventDestroy(e);
  if (s->hdone) (void)hipHostFree(s->hdone);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

rf_status rf_assemble_system(rf_simulation* s) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  const double* x0 = nullptr;
  s->r0_ready = false;
  if (s->theta) {
    if (s->apart) {  // the solve's initial guess now, its residual with the system
      if (!s->guess_taken) {
        bool warm = false;
        RF_TRY(initial_guess(s, &warm));
        s->guess_taken = true;
        s->guess_warm = warm;
      }
      if (s->guess_warm) x0 = s->x;
    }
    RF_HIP(hipEventRecord(s->ev[5], s->stream));
    RF_HIP(rf::launch_assemble_theta(s->N, s->NQI, s->analytical, s->M, s->Lp, s->pos[s->cur], s->ori[s->cur], s->T,
                                     s->bpart, s->Tv, s->cpart, x0, x0 ? s->apart : nullptr, s->pc, s->stream,
                                     &s->bgroups, s->T16, s->S16));
    RF_HIP(hipEventRecord(s->ev[6], s->stream));
    s->tori = s->cur;
    s->tori_valid = true;
    s->vvalid = s->vstored;
  } else {
    RF_HIP(hipEventRecord(s->ev[5], s->stream));
    RF_HIP(rf::launch_assemble(s->N, s->NQI, s->analytical, s->M, s->lda, s->pos[s->cur], s->ori[s->cur], s->A,
                               s->bpart, s->pc, s->stream));
    RF_HIP(hipEventRecord(s->ev[6], s->stream));
    s->bgroups = rf::pair_grid(s->M, 192 / (3 * s->NQI)).groups;
  }
  RF_HIP(rf::launch_sum_groups(s->bgroups, s->n, s->bpart, s->b, s->stream));
  if (x0) {
    RF_HIP(rf::launch_residual_groups(s->bgroups, s->n, s->apart, s->b, x0, s->r0, s->stream));
    s->r0_ready = true;
  }
  return RF_OK;
}

rf_status rf_solve_system(rf_simulation* s) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  s->ngemv = 0;
  s->nhalf = 0;
  s->lu_last = rf::LuResult{};
  switch (s->cfg.solver) {
    case RF_SOLVER_GMRES:
      RF_TRY(solve_gmres(s));
      break;
    case RF_SOLVER_GMRES_IR:
      RF_TRY(solve_gmres_ir(s));
      break;
    case RF_SOLVER_BICGSTAB:
      RF_TRY(solve_bicgstab(s));
      break;
    default:
      RF_TRY(solve_direct(s));
      break;
  }
  s->have_solution = true;
  return RF_OK;
}

rf_status rf_update_velocities(rf_simulation* s) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  if (s->vstored && s->vvalid)
    RF_HIP(rf::launch_velocities_stored(s->N, s->M, s->Lp, s->Tv, s->cpart, s->bgroups, s->ori[s->cur], s->x, s->pc,
                                        s->tv[s->vcur], s->rv[s->vcur], s->stream));
  else
    RF_HIP(rf::launch_velocities(s->N, s->NQI, s->vel_analytical, s->M, s->pos[s->cur], s->ori[s->cur], s->x,
                                 s->wf, s->fm, s->vpart, s->tv[s->vcur], s->rv[s->vcur], s->pc, s->stream));
  s->vlast = s->vcur;
  return RF_OK;
}

rf_status rf_update_fibers(rf_simulation* s, int32_t first_timestep) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  const bool first = first_timestep != 0 || !s->have_history;
  RF_HIP(rf::launch_update(first, s->M, s->cfg.timestep, s->pos[s->prev], s->pos[s->cur], s->ori[s->prev],
                           s->ori[s->cur], s->tv[s->vprev], s->tv[s->vcur], s->rv[s->vprev], s->rv[s->vcur],
                           s->pos[s->next], s->ori[s->next], s->stream));
  if (s->next == s->tori) s->tori_valid = false;
  s->plast = s->next;
  s->have_next = true;
  return RF_OK;
}

rf_status rf_advance(rf_simulation* s) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  if (!s->have_next) return fail(RF_ERR_INVALID_ARGUMENT, "rf_advance before rf_update_fibers");
  const int old_prev = s->prev;
  s->prev = s->cur;
  s->cur = s->next;
  s->next = old_prev;
  std::swap(s->vcur, s->vprev);
  s->have_history = true;
  s->have_next = false;
  s->vvalid = false;  // the stored velocity operator belongs to the old state
  return RF_OK;
}

rf_status rf_step(rf_simulation* s, int64_t current_timestep) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  RF_HIP(hipEventRecord(s->ev[0], s->stream));
  RF_TRY(rf_assemble_system(s));
  RF_HIP(hipEventRecord(s->ev[1], s->stream));
  RF_TRY(rf_solve_system(s));
  RF_HIP(hipEventRecord(s->ev[2], s->stream));
  RF_TRY(rf_update_velocities(s));
  RF_HIP(hipEventRecord(s->ev[3], s->stream));
  RF_TRY(check_nonfinite(s, s->tv[s->vcur], 3LL * s->M, "translational velocities"));
  RF_TRY(rf_update_fibers(s, current_timestep == 0 ? 1 : 0));
  RF_HIP(hipEventRecord(s->ev[4], s->stream));
  RF_HIP(hipEventSynchronize(s->ev[4]));
  s->timings.assemble_system = elapsed(s->ev[0], s->ev[1]);
  s->timings.solve_system = elapsed(s->ev[1], s->ev[2]);
  s->timings.update_velocities = elapsed(s->ev[2], s->ev[3]);
  s->timings.update_fibers = elapsed(s->ev[3], s->ev[4]);
  s->timings.total = elapsed(s->ev[0], s->ev[4]);
  s->timings.assemble_kernel = elapsed(s->ev[5], s->ev[6]);
  s->timings.gemv = 0.0;
  for (int k = 0; k < s->ngemv; ++k) s->timings.gemv += elapsed(s->gev[2 * k], s->gev[2 * k + 1]);
  s->timings.gemv_launches = s->ngemv;
  s->timings.gemv_half = 0.0;
  for (int k = 0; k < s->nhalf; ++k) s->timings.gemv_half += elapsed(s->hev[2 * k], s->hev[2 * k + 1]);
  s->timings.gemv_half_launches = s->nhalf;
  s->timings.gemm_launches = s->lu_last.gemm_launches;
  s->timings.gemm = s->lu_last.gemm_ms;
  s->timings.gemm_flop = s->lu_last.gemm_flop;
  s->timings.direct_fallback = s->direct_fallback ? 1 : 0;
  s->timings.solver_restarts = s->bicg_restarts;
  RF_TRY(rf_advance(s));
  if (!s->stats.converged) {
    g_last_error = "solver did not reach its tolerance (relative residual " +
                   std::to_string(s->stats.relative_residual) + ")";
    return RF_WARN_NOT_CONVERGED;
  }
  return RF_OK;
}

rf_status rf_get_state(rf_simulation* s, double* positions, double* orientations) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  const size_t bytes = 3 * static_cast<size_t>(s->M) * sizeof(double);
  if (positions) RF_HIP(hipMemcpyAsync(positions, s->pos[s->cur], bytes, hipMemcpyDeviceToHost, s->stream));
  if (orientations) RF_HIP(hipMemcpyAsync(orientations, s->ori[s->cur], bytes, hipMemcpyDeviceToHost, s->stream));
  RF_HIP(hipStreamSynchronize(s->stream));
  return RF_OK;
}

rf_status rf_get_next_state(rf_simulation* s, double* positions, double* orientations) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  const int k = s->plast;
  const size_t bytes = 3 * static_cast<size_t>(s->M) * sizeof(double);
  if (positions) RF_HIP(hipMemcpyAsync(positions, s->pos[k], bytes, hipMemcpyDeviceToHost, s->stream));
  if (orientations) RF_HIP(hipMemcpyAsync(orientations, s->ori[k], bytes, hipMemcpyDeviceToHost, s->stream));
  RF_HIP(hipStreamSynchronize(s->stream));
  return RF_OK;
}

rf_status rf_get_velocities(rf_simulation* s, double* translational, double* rotational) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  const int k = s->vlast;
  const size_t bytes = 3 * static_cast<size_t>(s->M) * sizeof(double);
  if (translational) RF_HIP(hipMemcpyAsync(translational, s->tv[k], bytes, hipMemcpyDeviceToHost, s->stream));
  if (rotational) RF_HIP(hipMemcpyAsync(rotational, s->rv[k], bytes, hipMemcpyDeviceToHost, s->stream));
  RF_HIP(hipStreamSynchronize(s->stream));
  return RF_OK;
}

rf_status rf_set_state(rf_simulation* s, const double* positions, const double* orientations, int32_t history,
                       const double* pp, const double* po, const double* pt, const double* pr) {
  if (!s || !positions || !orientations) return fail(RF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (history && (!pp || !po || !pt || !pr)) return fail(RF_ERR_INVALID_ARGUMENT, "history arrays are NULL");
  const size_t bytes = 3 * static_cast<size_t>(s->M) * sizeof(double);
  RF_HIP(hipMemcpyAsync(s->pos[s->cur], positions, bytes, hipMemcpyHostToDevice, s->stream));
  RF_HIP(hipMemcpyAsync(s->ori[s->cur], orientations, bytes, hipMemcpyHostToDevice, s->stream));
  if (history) {
    RF_HIP(hipMemcpyAsync(s->pos[s->prev], pp, bytes, hipMemcpyHostToDevice, s->stream));
    RF_HIP(hipMemcpyAsync(s->ori[s->prev], po, bytes, hipMemcpyHostToDevice, s->stream));
    RF_HIP(hipMemcpyAsync(s->tv[s->vprev], pt, bytes, hipMemcpyHostToDevice, s->stream));
    RF_HIP(hipMemcpyAsync(s->rv[s->vprev], pr, bytes, hipMemcpyHostToDevice, s->stream));
  }
  RF_HIP(hipStreamSynchronize(s->stream));
  if (s->tori == s->cur || (history && s->tori == s->prev)) s->tori_valid = false;
  s->have_history = history != 0;
  s->have_next = false;
  s->vvalid = false;
  s->have_previous = false;  // extrapolation needs two solutions of consecutive steps
  s->guess_taken = false;    // a residual formed with the old state is void
  s->r0_ready = false;
  return RF_OK;
}

rf_status rf_get_system(rf_simulation* s, float* a_matrix, double* b_vector, double* x_vector) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  const int64_t n = s->n;
  if (a_matrix && s->theta && !s->tori_valid)
    return fail(RF_ERR_INVALID_ARGUMENT, "no theta-form system to expand (none assembled, or the state it was "
                                         "assembled from has been overwritten)");
  if (a_matrix && s->theta) {
    // rebuild the dense blocks in a temporary device buffer
    float* tmp = nullptr;
    hipError_t e = hipMalloc(&tmp, sizeof(float) * static_cast<size_t>(n) * static_cast<size_t>(n));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(RF_ERR_OUT_OF_MEMORY, "no device memory to expand the theta-form matrix");
    }
    e = rf::launch_expand_theta(s->N, s->M, s->Lp, s->T, s->ori[s->tori], s->pc, tmp, s->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(a_matrix, tmp, sizeof(float) * static_cast<size_t>(n) * n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(RF_ERR_DEVICE, std::string("expanding the matrix: ") + hipGetErrorString(e));
  } else if (a_matrix) {
    RF_HIP(hipMemcpy2DAsync(a_matrix, n * sizeof(float), s->A, s->lda * sizeof(float), n * sizeof(float), n,
                            hipMemcpyDeviceToHost, s->stream));
  }
  if (b_vector) RF_HIP(hipMemcpyAsync(b_vector, s->b, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  if (x_vector) RF_HIP(hipMemcpyAsync(x_vector, s->x, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  RF_HIP(hipStreamSynchronize(s->stream));
  return RF_OK;
}

rf_status rf_get_system_noalloc(rf_simulation* s, float* a_matrix, double* b_vector, double* x_vector) {
  if (!s) return fail(RF_ERR_INVALID_ARGUMENT, "simulation is NULL");
  if (!a_matrix || !s->theta) return rf_get_system(s, a_matrix, b_vector, x_vector);  // no staging needed
  if (!s->tori_valid)
    return fail(RF_ERR_INVALID_ARGUMENT, "no theta-form system to expand (none assembled, or the state it was "
                                         "assembled from has been overwritten)");
  const int64_t fiber_cols = 3LL * s->N * s->n;  // floats of one source fiber's columns
  if (fiber_cols > static_cast<int64_t>(kScratchFloats))
    return fail(RF_ERR_UNSUPPORTED, "rf_get_system_noalloc: the 3N columns of one source fiber (" +
                                        std::to_string(fiber_cols * 4) + " bytes) exceed its " +
                                        std::to_string(kScratchFloats * 4) + "-byte staging buffer");
  DeviceScope scope;
  RF_TRY(scope.enter(s->device));
  float* stage = nullptr;
  RF_TRY(scratch_address(s->device, &stage));
  const int chunk = static_cast<int>(std::min<int64_t>(s->M, static_cast<int64_t>(kScratchFloats) / fiber_cols));
  // Stream order keeps each expansion behind the copy of the chunk before it;
  // the expansion writes every element of a chunk, so nothing is cleared.
  for (int j0 = 0; j0 < s->M; j0 += chunk) {
    const int jc = std::min(chunk, s->M - j0);
    RF_HIP(rf::launch_expand_theta(s->N, s->M, s->Lp, s->T, s->ori[s->tori], s->pc, stage, s->stream, j0, jc));
    RF_HIP(hipMemcpyAsync(a_matrix + static_cast<int64_t>(j0) * fiber_cols, stage,
                          sizeof(float) * static_cast<size_t>(jc) * static_cast<size_t>(fiber_cols),
                          hipMemcpyDeviceToHost, s->stream));
  }
  RF_HIP(hipStreamSynchronize(s->stream));
  return rf_get_system(s, nullptr, b_vector, x_vector);
}

int64_t rf_system_size(const rf_simulation* s) { return s ? s->n : 0; }

rf_status rf_get_solver_stats(rf_simulation* s, rf_solver_stats* st) {
  if (!s || !st) return fail(RF_ERR_INVALID_ARGUMENT, "NULL argument");
  *st = s->stats;
  return RF_OK;
}

rf_status rf_get_timings(rf_simulation* s, rf_timings* t) {
  if (!s || !t) return fail(RF_ERR_INVALID_ARGUMENT, "NULL argument");
  *t = s->timings;
  return RF_OK;
}

int64_t rf_device_bytes(const rf_simulation* s) { return s ? s->device_bytes : 0; }

const char* rf_last_error(void) { return g_last_error.c_str(); }

const char* rf_version(void) { return "rigidfibers_amd 0.1.0 (gfx950)"; }

}  // extern "C"
