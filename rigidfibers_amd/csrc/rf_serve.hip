// rf_serve.hip -- serving engine of include/rigidfibers.h: many independent
// suspensions (each the reference's Simulation, src/simulation.h:38-98, with
// its own fiber count, state and step count) advanced together on one GPU.
//
// A worker thread owns the device work.  Requests wait in a FIFO queue and
// are admitted while a slot is free and their device footprint fits the
// memory budget; every admitted suspension keeps its state, theta-form
// system, stored velocity operator, Krylov basis and previous solution
// resident until it finishes.  One tick advances every admitted suspension by
// one timestep (src/simulation.cu:286-329):
//   assemble    one launch, workgroup = (suspension, target fiber);
//   solve       lockstep GMRES(m): every iteration is one launch sequence over
//               all suspensions (vector prep, theta matvec, two Gram-Schmidt
//               passes, norm) and one device->host copy of every suspension's
//               Hessenberg column; each suspension keeps its own Krylov state,
//               restarts and tolerance, and leaves the lockstep when it converges;
//   velocities  one launch over the stored velocity operators;
//   update      one launch over all fibers.
// Launches find their suspension through tables of (suspension, fiber) or
// (suspension, 256-element block) entries rebuilt when the admitted set
// changes.  Reductions are per suspension in a fixed order, so each result is
// bitwise independent of what else is admitted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rigidfibers.h"
#include "rf_host.h"
#include "rf_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  int M = 0, steps = 0;
  int64_t footprint = 0;
  std::vector<double> pos0, ori0;      // initial state until admitted
  std::vector<double> pos_out, ori_out;
  std::vector<double> step_ms;
  rf_request_info info{};
  bool done = false;
};

struct Suspension {
  Request* req = nullptr;
  int M = 0, n = 0;
  int64_t Lp = 0;
  void* base = nullptr;  // one device allocation holding everything below
  int64_t bytes = 0;
  double* pos[3] = {};
  double* ori[3] = {};
  double* tv[2] = {};
  double* rv[2] = {};
  int cur = 0, prev = 1, next = 2, vcur = 0, vprev = 1;
  bool have_history = false, have_solution = false;
  float* T = nullptr;
  float* Tv = nullptr;
  double* cpart = nullptr;
  double* b = nullptr;
  double* x = nullptr;
  double* npart = nullptr;  // [M][2] norms of b and the first residual, per target
  double* V = nullptr;
  int steps_done = 0;
  rf_status status = RF_OK;
};

int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

template <typename T>
bool grow_device(T** p, size_t* cap, size_t need) {
  if (need <= *cap) return true;
  size_t c = std::max<size_t>(need, 2 * *cap);
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, c * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  *cap = c;
  return true;
}

template <typename T>
bool grow_host(T** p, size_t* cap, size_t need) {
  if (need <= *cap) return true;
  size_t c = std::max<size_t>(need, 2 * *cap);
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  *cap = 0;
  if (hipHostMalloc(reinterpret_cast<void**>(p), c * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  *cap = c;
  return true;
}

}  // namespace

struct rf_server {
  rf_server_config cfg{};
  rf::PhysConsts pc{};
  int device = 0;
  hipStream_t stream = nullptr;
  int N = 0, NQI = 0, m = 0, kdim = 0;
  bool analytical = false;
  Clock::time_point t0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::map<int64_t, Request*> requests;
  int64_t next_id = 1;
  bool stop = false;
  std::thread worker;

  std::vector<Suspension*> active;  // worker thread only
  int64_t budget = 0, used = 0, peak = 0;
  int64_t ticks = 0, fiber_steps = 0, completed = 0, n_active = 0;
  double busy = 0.0;

  // launch tables: device and pinned host mirrors
  rf::SusDev* d_sus = nullptr;
  rf::SusDev* h_sus = nullptr;
  rf::SusIter* d_it = nullptr;
  rf::SusIter* h_it = nullptr;
  int2* d_targets = nullptr;
  int2* h_targets = nullptr;
  int2* d_vblocks = nullptr;
  int2* h_vblocks = nullptr;
  int2* d_fblocks = nullptr;
  int2* h_fblocks = nullptr;
  double* d_part = nullptr;
  size_t c_dsus = 0, c_hsus = 0, c_dit = 0, c_hit = 0, c_dt = 0, c_ht = 0, c_dv = 0, c_hv = 0, c_df = 0, c_hf = 0,
         c_part = 0;
  int ntargets = 0, nvb = 0, nfb = 0;
  rf::SusCtl* d_ctl = nullptr;
  rf::SusCtl* h_ctl = nullptr;
  double* d_ctlmem = nullptr;  // Hessenberg, Givens and y of every suspension
  int* d_act = nullptr;
  int* h_act = nullptr;
  size_t c_dctl = 0, c_hctl = 0, c_cmem = 0, c_dact = 0, c_hact = 0;
  hipGraphExec_t gexec = nullptr;  // captured round of GMRES iterations of the admitted set
  std::array<int64_t, 10> gkey{};  // the launch parameters it was captured with

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

int64_t footprint(const rf_server* S, int M, int64_t* Lp_out) {
  const int64_t n = 3LL * S->N * M;
  const int64_t Lp = rf::theta_stride(M, S->N);
  const int64_t NC = 3 + 6 * (S->N - 1);
  int64_t b = 0;
  b += 10 * align256(3LL * M * 8);                 // pos[3], ori[3], tv[2], rv[2]
  b += align256(static_cast<int64_t>(M) * NC * Lp * 4);  // theta system
  b += align256(static_cast<int64_t>(M) * 12 * Lp * 4);  // velocity operator
  b += align256(6LL * M * 8);                      // velocity constants
  b += 2 * align256(n * 8);                        // b, x
  b += align256(2LL * M * 8);                      // norm partials
  b += align256(static_cast<int64_t>(S->kdim) * n * 8);  // Krylov basis
  if (Lp_out) *Lp_out = Lp;
  return b;
}

// Device allocation and upload of an admitted request (worker thread).
rf_status admit(rf_server* S, Request* r, Suspension** out) {
  auto* s = new (std::nothrow) Suspension();
  if (!s) return RF_ERR_OUT_OF_MEMORY;
  s->req = r;
  s->M = r->M;
  s->n = 3 * S->N * r->M;
  s->bytes = footprint(S, r->M, &s->Lp);
  if (hipMalloc(&s->base, s->bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return RF_ERR_OUT_OF_MEMORY;
  }
  char* p = static_cast<char*>(s->base);
  auto take = [&](int64_t bytes) {
    char* q = p;
    p += align256(bytes);
    return q;
  };
  const int64_t f3 = 3LL * s->M * 8;
  for (int k = 0; k < 3; ++k) s->pos[k] = reinterpret_cast<double*>(take(f3));
  for (int k = 0; k < 3; ++k) s->ori[k] = reinterpret_cast<double*>(take(f3));
  for (int k = 0; k < 2; ++k) s->tv[k] = reinterpret_cast<double*>(take(f3));
  for (int k = 0; k < 2; ++k) s->rv[k] = reinterpret_cast<double*>(take(f3));
  const int64_t NC = 3 + 6 * (S->N - 1);
  s->T = reinterpret_cast<float*>(take(static_cast<int64_t>(s->M) * NC * s->Lp * 4));
  s->Tv = reinterpret_cast<float*>(take(static_cast<int64_t>(s->M) * 12 * s->Lp * 4));
  s->cpart = reinterpret_cast<double*>(take(6LL * s->M * 8));
  s->b = reinterpret_cast<double*>(take(static_cast<int64_t>(s->n) * 8));
  s->x = reinterpret_cast<double*>(take(static_cast<int64_t>(s->n) * 8));
  s->npart = reinterpret_cast<double*>(take(2LL * s->M * 8));
  s->V = reinterpret_cast<double*>(take(static_cast<int64_t>(S->kdim) * s->n * 8));
  if (hipMemcpyAsync(s->pos[0], r->pos0.data(), f3, hipMemcpyHostToDevice, S->stream) != hipSuccess ||
      hipMemcpyAsync(s->ori[0], r->ori0.data(), f3, hipMemcpyHostToDevice, S->stream) != hipSuccess ||
      hipStreamSynchronize(S->stream) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(s->base);
    delete s;
    return RF_ERR_DEVICE;
  }
  r->pos0.clear();
  r->pos0.shrink_to_fit();
  r->ori0.clear();
  r->ori0.shrink_to_fit();
  *out = s;
  return RF_OK;
}

// Tables of the admitted set (rebuilt every tick: the state pointers rotate).
bool build_tables(rf_server* S) {
  const int ns = static_cast<int>(S->active.size());
  int nt = 0, nv = 0, nf = 0;
  for (auto* s : S->active) {
    nt += s->M;
    nv += (s->n + 255) / 256;
    nf += (s->M + 255) / 256;
  }
  if (!grow_host(&S->h_sus, &S->c_hsus, ns) || !grow_device(&S->d_sus, &S->c_dsus, ns) ||
      !grow_host(&S->h_it, &S->c_hit, ns) || !grow_device(&S->d_it, &S->c_dit, ns) ||
      !grow_host(&S->h_targets, &S->c_ht, nt) || !grow_device(&S->d_targets, &S->c_dt, nt) ||
      !grow_host(&S->h_vblocks, &S->c_hv, nv) || !grow_device(&S->d_vblocks, &S->c_dv, nv) ||
      !grow_host(&S->h_fblocks, &S->c_hf, nf) || !grow_device(&S->d_fblocks, &S->c_df, nf) ||
      !grow_device(&S->d_part, &S->c_part, 3 * static_cast<size_t>(nv) * S->kdim) ||
      !grow_host(&S->h_ctl, &S->c_hctl, ns) || !grow_device(&S->d_ctl, &S->c_dctl, ns) ||
      !grow_device(&S->d_ctlmem, &S->c_cmem,
                   static_cast<size_t>(ns) * ((S->m + 1) * static_cast<size_t>(S->m) + 4 * S->m + 1)) ||
      !grow_host(&S->h_act, &S->c_hact, ns) || !grow_device(&S->d_act, &S->c_dact, ns))
    return false;
  int ti = 0, vi = 0, fi = 0;
  for (int q = 0; q < ns; ++q) {
    Suspension* s = S->active[q];
    rf::SusDev& d = S->h_sus[q];
    d.M = s->M;
    d.n = s->n;
    d.Lp = s->Lp;
    d.vb_first = vi;
    d.vb_count = (s->n + 255) / 256;
    d.pos = s->pos[s->cur];
    d.ori = s->ori[s->cur];
    d.T = s->T;
    d.Tv = s->Tv;
    d.cpart = s->cpart;
    d.b = s->b;
    d.x = s->x;
    // the solve starts from the previous solution when there is one
    d.x0 = (S->cfg.base.initial_guess != RF_GUESS_ZERO && s->have_solution) ? s->x : nullptr;
    d.npart = s->npart;
    d.V = s->V;
    d.tv = s->tv[s->vcur];
    d.rv = s->rv[s->vcur];
    d.pos_prev = s->pos[s->prev];
    d.ori_prev = s->ori[s->prev];
    d.tv_prev = s->tv[s->vprev];
    d.rv_prev = s->rv[s->vprev];
    d.pos_next = s->pos[s->next];
    d.ori_next = s->ori[s->next];
    d.first = s->have_history ? 0 : 1;
    for (int i = 0; i < s->M; ++i) S->h_targets[ti++] = make_int2(q, i);
    for (int b = 0; b < d.vb_count; ++b) S->h_vblocks[vi++] = make_int2(q, b);
    for (int b = 0; b < (s->M + 255) / 256; ++b) S->h_fblocks[fi++] = make_int2(q, b);
  }
  S->ntargets = nt;
  S->nvb = nv;
  S->nfb = nf;
  hipStream_t st = S->stream;
  return hipMemcpyAsync(S->d_sus, S->h_sus, ns * sizeof(rf::SusDev), hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(S->d_targets, S->h_targets, nt * sizeof(int2), hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(S->d_vblocks, S->h_vblocks, nv * sizeof(int2), hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(S->d_fblocks, S->h_fblocks, nf * sizeof(int2), hipMemcpyHostToDevice, st) == hipSuccess;
}

// Lockstep iterations per round: a round is replayed as one graph and ends
// with the round's only device->host copy (the suspensions still working).
constexpr int kRound = 4;

// The kernels of one lockstep iteration over the admitted set and its control
// step (fixed shapes within a tick: the Gram-Schmidt launches cover all m
// basis vectors and the workgroups past a suspension's K exit).
bool iteration_launches(rf_server* S) {
  const int ns = static_cast<int>(S->active.size());
  hipStream_t st = S->stream;
  if (rf::launch_iter_prep_batch(S->nvb, S->d_sus, S->d_it, S->d_vblocks, st) != hipSuccess) return false;
  if (rf::launch_gemv_theta_batch(S->N, S->ntargets, S->d_sus, S->d_it, S->d_targets, S->pc, st) != hipSuccess)
    return false;
  // <V_j, w> partials, then two fused passes (orthogonalise, next partials)
  if (rf::launch_dots_partial_batch(S->nvb, S->m, S->d_sus, S->d_it, S->d_vblocks, S->d_part, st) != hipSuccess)
    return false;
  for (int ph = 0; ph < 2; ++ph)
    if (rf::launch_orth_dots_batch(S->nvb, S->kdim, ph, S->d_sus, S->d_it, S->d_vblocks, S->d_part, st) !=
        hipSuccess)
      return false;
  return rf::launch_gmres_control_batch(ns, S->m, S->kdim, S->nvb, S->cfg.base.gmres_tolerance,
                                        S->cfg.base.gmres_max_iterations, S->cfg.base.confirm_residual, S->d_sus,
                                        S->d_it, S->d_ctl, S->d_part, S->d_act, st) == hipSuccess;
}

bool round_launches(rf_server* S) {
  for (int r = 0; r < kRound; ++r)
    if (!iteration_launches(S)) return false;
  return true;
}

void release_iteration(rf_server* S) {
  if (S->gexec) (void)hipGraphExecDestroy(S->gexec);
  S->gexec = nullptr;
}

// Captures round_launches as a graph.  Its launch parameters are the table
// sizes and device addresses (the tables' contents are written before a
// round or by the control steps inside it), so the graph is kept across
// ticks until the admitted set changes them.  Kernels only: no memset or
// copy nodes.
bool ensure_iteration_graph(rf_server* S) {
  const std::array<int64_t, 10> key = {static_cast<int64_t>(S->active.size()),
                                       S->nvb,
                                       S->ntargets,
                                       reinterpret_cast<int64_t>(S->d_sus),
                                       reinterpret_cast<int64_t>(S->d_it),
                                       reinterpret_cast<int64_t>(S->d_targets),
                                       reinterpret_cast<int64_t>(S->d_vblocks),
                                       reinterpret_cast<int64_t>(S->d_part),
                                       reinterpret_cast<int64_t>(S->d_ctl),
                                       reinterpret_cast<int64_t>(S->d_act)};
  if (S->gexec && key == S->gkey) return true;
  release_iteration(S);
  hipStream_t st = S->stream;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  const bool ok = round_launches(S);
  if (hipStreamEndCapture(st, &g) != hipSuccess || !ok) {
    if (g) (void)hipGraphDestroy(g);
    return false;
  }
  const hipError_t e = hipGraphInstantiate(&S->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    S->gexec = nullptr;
    return false;
  }
  S->gkey = key;
  return true;
}

// One round of lockstep iterations (graph replay or eager), then one copy of
// the control states and work flags; *any = whether a suspension still has
// work afterwards.
bool lockstep_round(rf_server* S, bool* any) {
  const int ns = static_cast<int>(S->active.size());
  hipStream_t st = S->stream;
  if (S->gexec) {
    if (hipGraphLaunch(S->gexec, st) != hipSuccess) return false;
  } else if (!round_launches(S)) {
    return false;
  }
  if (hipMemcpyAsync(S->h_act, S->d_act, ns * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(S->h_ctl, S->d_ctl, ns * sizeof(rf::SusCtl), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return false;
  *any = false;
  for (int q = 0; q < ns; ++q) *any = *any || S->h_act[q] != 0;
  return true;
}

// GMRES(m) for every admitted suspension (the same algorithm as the
// single-simulation solver: true residual at each restart, classical
// Gram-Schmidt twice, Givens rotations, the solve ending on a converged
// cycle's estimate unless confirm_residual), its control on the device: the
// assembly has written the first residual b - A x0 (b when cold) into V_0 and
// the norms' per-target partials; a start step of the control takes them
// (x = 0 when cold) and sets up the first Arnoldi step; the host replays
// rounds of iterations until no suspension has work left.
bool solve_lockstep(rf_server* S) {
  const int ns = static_cast<int>(S->active.size());
  const int m = S->m;
  const int maxit = S->cfg.base.gmres_max_iterations;
  const bool warm_ok = S->cfg.base.initial_guess != RF_GUESS_ZERO;
  const int64_t per = static_cast<int64_t>(m + 1) * m + 4 * m + 1;
  for (int q = 0; q < ns; ++q) {
    Suspension* s = S->active[q];
    rf::SusCtl& C = S->h_ctl[q];
    C = rf::SusCtl{};
    double* mem = S->d_ctlmem + q * per;
    C.H = mem;
    C.cs = C.H + static_cast<int64_t>(m + 1) * m;
    C.sn = C.cs + m;
    C.g = C.sn + m;
    C.y = C.g + m + 1;
    C.phase = 3;
    C.cold = !(warm_ok && s->have_solution);
    rf::SusIter& c = S->h_it[q];
    c = rf::SusIter{};
    c.ycoef = C.y;
  }
  hipStream_t st = S->stream;
  // the start step: the control reads the norms of b and of the first
  // residual, which the assembly left in V_0, and sets up the first Arnoldi step
  if (hipMemcpyAsync(S->d_ctl, S->h_ctl, ns * sizeof(rf::SusCtl), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(S->d_it, S->h_it, ns * sizeof(rf::SusIter), hipMemcpyHostToDevice, st) != hipSuccess ||
      rf::launch_gmres_control_batch(ns, S->m, S->kdim, S->nvb, S->cfg.base.gmres_tolerance, maxit,
                                     S->cfg.base.confirm_residual, S->d_sus, S->d_it, S->d_ctl, S->d_part, S->d_act,
                                     st) != hipSuccess)
    return false;
  if (!S->cfg.eager && !ensure_iteration_graph(S)) return false;
  // Every suspension stops by itself: at most maxit Arnoldi steps, and a true
  // residual after each restart cycle, which holds at least one Arnoldi step
  // (cycles can end early, on a zero ||w|| or, with confirm_residual, on an
  // estimate the true residual does not confirm), so at most 2 maxit + 1
  // iterations, one to apply the last update and the start step.  The bound
  // only guards the host loop.
  const int64_t max_rounds = (2 * static_cast<int64_t>(maxit) + 4) / kRound + 2;
  bool any = true;
  for (int64_t r = 0; any; ++r) {
    if (r >= max_rounds) return false;
    if (!lockstep_round(S, &any)) return false;
  }
  for (int q = 0; q < ns; ++q) {  // h_ctl holds the last round's control states
    Suspension* s = S->active[q];
    const rf::SusCtl& C = S->h_ctl[q];
    s->have_solution = true;
    s->req->info.gmres_iterations += C.it;
    if (C.nonfinite)
      s->status = RF_ERR_NONFINITE;
    else if (!C.converged && s->status == RF_OK)
      s->status = RF_WARN_NOT_CONVERGED;
  }
  return true;
}

// One timestep of every admitted suspension.
bool tick(rf_server* S) {
  hipStream_t st = S->stream;
  if (!build_tables(S)) return false;
  if (rf::launch_assemble_theta_batch(S->N, S->NQI, S->analytical, S->ntargets, S->d_sus, S->d_targets, S->pc, st) !=
      hipSuccess)
    return false;
  if (!solve_lockstep(S)) return false;
  if (rf::launch_velocities_stored_batch(S->N, S->ntargets, S->d_sus, S->d_targets, S->pc, st) != hipSuccess)
    return false;
  if (rf::launch_update_batch(S->nfb, S->d_sus, S->d_fblocks, S->cfg.base.timestep, st) != hipSuccess) return false;
  if (hipStreamSynchronize(st) != hipSuccess) return false;
  for (auto* s : S->active) {  // rf_advance
    const int old_prev = s->prev;
    s->prev = s->cur;
    s->cur = s->next;
    s->next = old_prev;
    std::swap(s->vcur, s->vprev);
    s->have_history = true;
    ++s->steps_done;
  }
  return true;
}

// Copies the final state, frees the suspension, publishes the result.
void retire(rf_server* S, Suspension* s, rf_status status) {
  Request* r = s->req;
  const size_t f3 = 3 * static_cast<size_t>(s->M);
  r->pos_out.assign(f3, 0.0);
  r->ori_out.assign(f3, 0.0);
  if (status == RF_OK || status == RF_WARN_NOT_CONVERGED) {
    if (hipMemcpy(r->pos_out.data(), s->pos[s->cur], f3 * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(r->ori_out.data(), s->ori[s->cur], f3 * 8, hipMemcpyDeviceToHost) != hipSuccess) {
      (void)hipGetLastError();
      status = RF_ERR_DEVICE;
    }
    for (size_t i = 0; i < f3 && status != RF_ERR_DEVICE; ++i)
      if (!std::isfinite(r->pos_out[i]) || !std::isfinite(r->ori_out[i])) status = RF_ERR_NONFINITE;
  }
  (void)hipFree(s->base);
  std::lock_guard<std::mutex> lk(S->mu);
  S->used -= s->bytes;
  --S->n_active;
  r->info.status = status;
  r->info.finished = S->now();
  r->done = true;
  ++S->completed;
  delete s;
  S->cv_done.notify_all();
}

void worker_main(rf_server* S) {
  (void)hipSetDevice(S->device);
  while (true) {
    std::vector<Request*> admitted;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (S->active.empty())
        S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
      if (S->stop) break;
      while (!S->queue.empty() &&
             static_cast<int>(S->active.size() + admitted.size()) < S->cfg.max_active &&
             S->used + S->queue.front()->footprint <= S->budget) {
        Request* r = S->queue.front();
        S->queue.pop_front();
        S->used += r->footprint;
        S->peak = std::max(S->peak, S->used);
        r->info.admitted = S->now();
        ++S->n_active;
        admitted.push_back(r);
      }
    }
    for (Request* r : admitted) {
      Suspension* s = nullptr;
      const rf_status st = admit(S, r, &s);
      if (st == RF_OK) {
        S->active.push_back(s);
      } else {
        std::lock_guard<std::mutex> lk(S->mu);
        S->used -= r->footprint;
        --S->n_active;
        r->info.status = st;
        r->info.finished = S->now();
        r->done = true;
        ++S->completed;
        S->cv_done.notify_all();
      }
    }
    if (S->active.empty()) continue;
    const auto t_start = Clock::now();
    const bool ok = tick(S);
    const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t_start).count();
    if (!ok) {  // let queued work finish before the suspensions' memory is freed
      (void)hipStreamSynchronize(S->stream);
      (void)hipGetLastError();
      release_iteration(S);
    }
    std::vector<Suspension*> keep;
    int64_t fsteps = 0;
    for (auto* s : S->active) {
      if (!ok) {
        (void)hipGetLastError();
        retire(S, s, RF_ERR_DEVICE);
        continue;
      }
      s->req->step_ms.push_back(ms);
      fsteps += s->M;
      if (s->status != RF_OK && s->status != RF_WARN_NOT_CONVERGED) {
        retire(S, s, s->status);
      } else if (s->steps_done >= s->req->steps) {
        retire(S, s, s->status);
      } else {
        keep.push_back(s);
      }
    }
    S->active.swap(keep);
    std::lock_guard<std::mutex> lk(S->mu);
    ++S->ticks;
    S->fiber_steps += fsteps;
    S->busy += ms * 1e-3;
  }
  for (auto* s : S->active) retire(S, s, RF_ERR_DEVICE);  // stopped before finishing
  S->active.clear();
}

void free_tables(rf_server* S) {
  for (void* p : {static_cast<void*>(S->d_sus), static_cast<void*>(S->d_it), static_cast<void*>(S->d_targets),
                  static_cast<void*>(S->d_vblocks), static_cast<void*>(S->d_fblocks), static_cast<void*>(S->d_part),
                  static_cast<void*>(S->d_ctl), static_cast<void*>(S->d_ctlmem),
                  static_cast<void*>(S->d_act)})
    if (p) (void)hipFree(p);
  for (void* p : {static_cast<void*>(S->h_sus), static_cast<void*>(S->h_it), static_cast<void*>(S->h_targets),
                  static_cast<void*>(S->h_vblocks), static_cast<void*>(S->h_fblocks),
                  static_cast<void*>(S->h_ctl), static_cast<void*>(S->h_act)})
    if (p) (void)hipHostFree(p);
}

}  // namespace

extern "C" {

rf_status rf_server_create(const rf_server_config* cfg, rf_server** out) {
  if (!cfg || !out) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  rf_config base = cfg->base;
  base.number_of_fibers = 1;  // per request
  rf_status st = rf::check_config(&base);
  if (st != RF_OK) return st;
  if (base.solver != RF_SOLVER_GMRES)
    return rf::set_error(RF_ERR_UNSUPPORTED, "the server steps its suspensions with GMRES");
  const bool analytical = base.integration == RF_INTEGRATION_ANALYTICAL;
  const bool vel_analytical = base.velocity_integration == RF_VELOCITY_FOLLOW
                                  ? analytical
                                  : base.velocity_integration == RF_VELOCITY_ANALYTICAL;
  if (vel_analytical != analytical)
    return rf::set_error(RF_ERR_UNSUPPORTED, "the server integrates the velocities like the system");
  if (cfg->max_active < 1) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "max_active must be >= 1");
  if (cfg->memory_budget < 0) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "memory_budget must be >= 0");
  auto* S = new (std::nothrow) rf_server();
  if (!S) return rf::set_error(RF_ERR_OUT_OF_MEMORY, "host allocation failed");
  S->cfg = *cfg;
  S->cfg.base = base;
  S->N = base.number_of_terms_in_force_expansion;
  S->NQI = base.number_of_quadrature_intervals;
  S->m = base.gmres_restart;
  S->kdim = S->m + 1;
  S->analytical = analytical;
  rf::physical_constants(base, &S->pc);
  S->t0 = Clock::now();
  if (hipGetDevice(&S->device) != hipSuccess ||
      hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete S;
    return rf::set_error(RF_ERR_DEVICE, "no HIP device or stream");
  }
  S->budget = cfg->memory_budget;
  if (S->budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipStreamDestroy(S->stream);
      delete S;
      return rf::set_error(RF_ERR_DEVICE, "hipMemGetInfo failed");
    }
    S->budget = static_cast<int64_t>(0.8 * static_cast<double>(free_b));
  }
  try {
    S->worker = std::thread(worker_main, S);
  } catch (...) {
    (void)hipStreamDestroy(S->stream);
    delete S;
    return rf::set_error(RF_ERR_OUT_OF_MEMORY, "cannot start the worker thread");
  }
  *out = S;
  return RF_OK;
}

void rf_server_destroy(rf_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  (void)hipStreamSynchronize(S->stream);
  release_iteration(S);
  free_tables(S);
  for (auto& kv : S->requests) delete kv.second;
  (void)hipStreamDestroy(S->stream);
  delete S;
}

rf_status rf_server_submit(rf_server* S, int32_t fibers, const double* positions, const double* orientations,
                           int32_t steps, int64_t* id) {
  if (!S || !positions || !orientations || !id) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (fibers < 1 || steps < 1) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "fibers and steps must be >= 1");
  const size_t f3 = 3 * static_cast<size_t>(fibers);
  for (size_t i = 0; i < f3; ++i)
    if (!std::isfinite(positions[i]) || !std::isfinite(orientations[i]))
      return rf::set_error(RF_ERR_INVALID_ARGUMENT, "non-finite initial state");
  auto* r = new (std::nothrow) Request();
  if (!r) return rf::set_error(RF_ERR_OUT_OF_MEMORY, "host allocation failed");
  r->M = fibers;
  r->steps = steps;
  r->footprint = footprint(S, fibers, nullptr);
  if (r->footprint > S->budget) {
    delete r;
    return rf::set_error(RF_ERR_OUT_OF_MEMORY, "the suspension's footprint exceeds the server's memory budget");
  }
  r->pos0.assign(positions, positions + f3);
  r->ori0.assign(orientations, orientations + f3);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    r->info.id = r->id;
    r->info.fibers = fibers;
    r->info.steps = steps;
    r->info.footprint = r->footprint;
    r->info.submitted = S->now();
    S->requests[r->id] = r;
    S->queue.push_back(r);
  }
  S->cv_work.notify_all();
  *id = r->id;
  return RF_OK;
}

rf_status rf_server_wait(rf_server* S, int64_t id, double timeout_s, int32_t* done) {
  if (!S) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::unique_lock<std::mutex> lk(S->mu);
  auto itr = S->requests.find(id);
  if (itr == S->requests.end()) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "unknown request");
  Request* r = itr->second;
  if (timeout_s < 0)
    S->cv_done.wait(lk, [&] { return r->done; });
  else
    S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->done; });
  if (done) *done = r->done ? 1 : 0;
  return RF_OK;
}

rf_status rf_server_collect(rf_server* S, int64_t id, double* positions, double* orientations, double* step_ms,
                            rf_request_info* info) {
  if (!S) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "NULL argument");
  Request* r = nullptr;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto itr = S->requests.find(id);
    if (itr == S->requests.end()) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "unknown request");
    r = itr->second;
    if (!r->done) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "request has not finished");
    S->requests.erase(itr);
  }
  const size_t f3 = 3 * static_cast<size_t>(r->M);
  if (positions && r->pos_out.size() == f3) std::memcpy(positions, r->pos_out.data(), f3 * sizeof(double));
  if (orientations && r->ori_out.size() == f3) std::memcpy(orientations, r->ori_out.data(), f3 * sizeof(double));
  if (step_ms)
    for (int k = 0; k < r->steps; ++k) step_ms[k] = k < static_cast<int>(r->step_ms.size()) ? r->step_ms[k] : 0.0;
  if (info) *info = r->info;
  delete r;
  return RF_OK;
}

rf_status rf_server_get_stats(rf_server* S, rf_server_stats* st) {
  if (!S || !st) return rf::set_error(RF_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  st->ticks = S->ticks;
  st->fiber_steps = S->fiber_steps;
  st->completed = S->completed;
  st->queued = static_cast<int64_t>(S->queue.size());
  st->active = S->n_active;
  st->busy_seconds = S->busy;
  st->peak_bytes = S->peak;
  st->memory_budget = S->budget;
  return RF_OK;
}

double rf_server_clock(rf_server* S) { return S ? S->now() : 0.0; }

}  // extern "C"
