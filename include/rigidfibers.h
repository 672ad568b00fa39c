/*
 * rigidfibers.h -- C ABI of librigidfibers.so, the MI355X (gfx950) rigid-fiber
 * Stokes-flow timestep.
 *
 * The reference has no library ABI: its timestep lives in the C++ class
 * Simulation (src/simulation.h:38-98) behind a process boundary
 * (runner/cuda_runner.py:154-163, `fibers <layout>`).  Every entry point below
 * replaces one member or free function of that class, cited beside it, so a
 * host (the Python runner here, or a ctypes/cgo/JNI binding) can drive the
 * same assemble -> solve -> update velocities -> update fibers sequence.
 *
 * Conventions
 *  - All functions return rf_status (0 = RF_OK).  rf_last_error() gives the
 *    message of the last failure on the calling thread.
 *  - Fiber vectors are host arrays of 3*M doubles (x0,y0,z0,x1,...).  The
 *    system has n = 3*N*M unknowns; A is returned column-major, row
 *    3N*i + 3k + a, column 3N*j + 3k + b (src/kernels/assemble_system_2D.cu:144-160).
 *  - The caller owns every host array; the library copies what it needs.
 *    The library owns all device memory and frees it in rf_destroy.
 *  - One host thread per rf_simulation; each simulation runs on its own HIP
 *    stream on the device that was current when rf_create was called.
 */
#ifndef RIGIDFIBERS_H_
#define RIGIDFIBERS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RF_API __attribute__((visibility("default")))

typedef enum rf_status {
  RF_OK = 0,
  RF_ERR_INVALID_ARGUMENT = 1,
  RF_ERR_UNSUPPORTED = 2,
  RF_ERR_OUT_OF_MEMORY = 3,
  RF_ERR_DEVICE = 4,
  RF_ERR_NONFINITE = 5,    /* NaN/Inf in the solution or velocities (e.g. coincident fibers) */
  RF_ERR_SINGULAR = 6,     /* zero pivot in the direct solver */
  RF_WARN_NOT_CONVERGED = 100 /* step completed, iterative solver hit its iteration limit */
} rf_status;

/* --magma (direct LU), --gmres, --bicgstab of fibers_runner.py:101-104, plus
 * the mixed-precision forms. */
typedef enum rf_solver {
  RF_SOLVER_GMRES = 0,
  RF_SOLVER_BICGSTAB = 1,
  RF_SOLVER_DIRECT = 2,            /* FP32 LU with partial pivoting (magma_sgesv_gpu / sgesv) */
  RF_SOLVER_DIRECT_MIXED_BF16 = 3, /* BF16-MFMA LU + FP64 iterative refinement */
  RF_SOLVER_DIRECT_MIXED_FP16 = 4, /* FP16-MFMA LU + FP64 iterative refinement */
  RF_SOLVER_GMRES_IR = 5           /* GMRES on the FP16-stored theta system inside an iterative refinement
                                      whose residual uses the FP32 theta system in FP64 */
} rf_solver;

/* --numerical / --analytical (fibers_runner.py:105-108). */
typedef enum rf_integration { RF_INTEGRATION_NUMERICAL = 0, RF_INTEGRATION_ANALYTICAL = 1 } rf_integration;

/* Inner integral of the velocity update.  The shipped Fortran and CUDA
 * (fortran/update_velocities.incl, src/kernels/update_velocities_2D.cu) always
 * use the quadrature; the reference's golden vectors for --analytical
 * (tests/reference/100_analytic_*) integrate it exactly.  FOLLOW uses the
 * system's integration. */
typedef enum rf_velocity_integration {
  RF_VELOCITY_FOLLOW = 0,
  RF_VELOCITY_NUMERICAL = 1,
  RF_VELOCITY_ANALYTICAL = 2
} rf_velocity_integration;

/* Initial guess of the iterative solvers.  The shipped Fortran and ViennaCL
 * start from zero (fortran/solve_system.incl:16-17); the reference's golden
 * vectors for --gmres (tests/reference/100_*_gmres) start each step from the
 * previous step's solution, which also saves iterations.  EXTRAPOLATED starts
 * from 2 x_n - x_{n-1} once two solutions exist (the previous one before
 * that): along a trajectory its error is O(dt^2) against O(dt). */
typedef enum rf_initial_guess {
  RF_GUESS_ZERO = 0,
  RF_GUESS_PREVIOUS = 1,
  RF_GUESS_EXTRAPOLATED = 2
} rf_initial_guess;

/* Storage of A on the device.  DENSE: n x n FP32 column-major, the
 * reference's layout (src/simulation.cu:64-110).  THETA: each 3x3 block kept
 * as the 3 + 6(N-1) numbers that determine a block column (Q0 and the
 * symmetric projections Theta, fortran/assemble_matrix.incl:59-137), 40% fewer
 * bytes for N = 5; the matrix-vector product rebuilds the blocks on the fly.
 * AUTO: THETA.  The direct solvers expand the theta form into their factor
 * storage and refine with the theta matrix-vector product; with DENSE they
 * copy the dense matrix and refine with the dense product.  Dumps
 * (rf_get_system) always return the dense matrix. */
typedef enum rf_matrix_format { RF_MATRIX_AUTO = 0, RF_MATRIX_DENSE = 1, RF_MATRIX_THETA = 2 } rf_matrix_format;

/* How the velocity update forms its sums over source fibers
 * (fortran/update_velocities.incl:29-86).  Both are projections of the inner
 * integrals the assembly already computes, so with STORED (the default) a
 * theta-form assembly whose integration the velocities share also keeps 12
 * FP32 numbers per (target, source, source mode), and rf_update_velocities
 * becomes one streaming pass over them (valid until the state changes).
 * PAIRWISE recomputes the kernel integrals for every fiber pair.  STORED
 * falls back to PAIRWISE for the dense format, when the velocity integration
 * differs from the system's, or after rf_set_state without a new assembly. */
typedef enum rf_velocity_operator { RF_VELOCITY_STORED = 0, RF_VELOCITY_PAIRWISE = 1 } rf_velocity_operator;

/* The keys of the reference's configuration file (tests/config.ini:18-37,
 * read by runner/cuda_runner.py:18-47), as run-time values. */
typedef struct rf_config {
  int32_t number_of_fibers;                           /* M */
  int32_t number_of_terms_in_force_expansion;         /* N, 1..8 */
  int32_t number_of_quadrature_intervals;             /* 4, 8 or 16 */
  int32_t number_of_quadrature_points_per_interval;   /* 3 (the only rule the reference implements) */
  double timestep;
  double slenderness;
  double external_force[3];                           /* F; the reference carries F/2 = (0,0,-0.5) */
  int32_t solver;                                     /* rf_solver */
  int32_t integration;                                /* rf_integration */
  int32_t velocity_integration;                       /* rf_velocity_integration */
  int32_t gmres_restart;
  int32_t gmres_max_iterations;
  double gmres_tolerance;                             /* relative residual ||b - Ax|| / ||b|| */
  int32_t bicgstab_max_iterations;
  double bicgstab_tolerance;
  double refinement_tolerance;                        /* mixed direct solvers: FP64 residual target */
  int32_t refinement_max_iterations;
  int32_t initial_guess;                              /* rf_initial_guess */
  int32_t matrix_format;                              /* rf_matrix_format */
  int32_t velocity_operator;                          /* rf_velocity_operator */
  double inner_tolerance;                             /* GMRES-IR: relative residual of each inner solve (0: 1e-4) */
  int32_t reserved0;                                  /* unused: 0 */
  int32_t separate_residual;                          /* iterative solvers on the theta form: 0 the first residual
                                                         b - A x0 of a warm start comes from the assembly kernel
                                                         (one matvec fewer per solve), 1 from a matvec */
  int32_t confirm_residual;                           /* GMRES: 0 stops on the restart cycle's own residual
                                                         estimate, as lib/viennacl/linalg/gmres.hpp:330-335
                                                         does; 1 confirms it with the true residual b - A x
                                                         (one matvec more per solve), as fortran/sgmres.f does
                                                         with icntl(8) = 1.  BiCGStab: 0 reports a converged
                                                         solve's recurrence residual, as
                                                         lib/viennacl/linalg/bicgstab.hpp:146-160 does; 1 the
                                                         true residual of the returned x */
  int32_t reserved[1];
} rf_config;

typedef struct rf_solver_stats {
  int32_t iterations;          /* Krylov iterations (matvecs) or refinement sweeps of the last solve */
  int32_t converged;           /* 1 if the tolerance was met */
  double relative_residual;    /* ||b - A x|| / ||b|| of the returned x (FP64); a converged GMRES or
                                  BiCGStab solve without confirm_residual: the restart cycle's
                                  estimate of it, or BiCGStab's recurrence residual */
} rf_solver_stats;

typedef struct rf_timings {
  /* milliseconds of the last step, HIP events on the simulation's stream;
   * names of src/performance.cc:110-134 */
  double assemble_system;
  double solve_system;
  double update_velocities;
  double update_fibers;
  double total;
  double gemv;             /* summed duration of the matrix-vector kernels of the solve */
  int32_t gemv_launches;   /* how many there were */
  int32_t gemm_launches;   /* direct solvers: trailing-update GEMMs of the LU factorisation (built-in MFMA kernel) */
  double gemm;             /* their summed duration */
  double gemm_flop;        /* their algorithmic flops, 2 m n K each */
  double gemv_half;        /* GMRES-IR: summed duration of the FP16 matrix-vector kernels */
  int32_t gemv_half_launches;
  int32_t direct_fallback; /* direct solvers: 1 when the refinement did not converge and GMRES on the dense
                              system finished the solve (iterations then count both) */
  double assemble_kernel;  /* the assembly kernel alone (assemble_system less its partial-sum kernels) */
  int32_t solver_restarts; /* BiCGStab: restarts from the true residual in the last solve (a degenerate search
                              direction, or every 200 iterations, as lib/viennacl/linalg/bicgstab.hpp:150-152) */
  int32_t reserved1;
} rf_timings;

typedef struct rf_simulation rf_simulation;

/* Fills *cfg with the values of the reference's tests/config.ini. */
RF_API rf_status rf_config_init(rf_config* cfg);

/* Simulation::Simulation(Configuration) (src/simulation.cu:44-57): allocates
 * device memory (initializeGPUMemory, :64-110), uploads the initial state
 * (writeFiberStateToDevice, :112-118) and the constants
 * (precomputeLegendrePolynomials, :153-284). */
RF_API rf_status rf_create(const rf_config* cfg, const double* positions, const double* orientations,
                           rf_simulation** out);

/* Simulation::~Simulation (src/simulation.cu:59-62), which in the reference
 * never frees device memory. */
RF_API void rf_destroy(rf_simulation* sim);

/* Simulation::step(size_t) (src/simulation.cu:286-329): assemble, solve,
 * update velocities, update fibers (forward Euler when current_timestep == 0),
 * then rotate the state buffers. */
RF_API rf_status rf_step(rf_simulation* sim, int64_t current_timestep);

/* The four phases of Simulation::step, for VALIDATE-style dumps between them. */
RF_API rf_status rf_assemble_system(rf_simulation* sim);   /* assembleSystem, :331-399 */
RF_API rf_status rf_solve_system(rf_simulation* sim);      /* solveSystem, :401-445 */
RF_API rf_status rf_update_velocities(rf_simulation* sim); /* updateVelocities, :447-489 */
RF_API rf_status rf_update_fibers(rf_simulation* sim, int32_t first_timestep); /* updateFibers, :491-528 */
RF_API rf_status rf_advance(rf_simulation* sim);           /* the buffer swaps, :307-311 */

/* Current positions/orientations (what saveFibers writes, :530-566); either
 * pointer may be NULL. */
RF_API rf_status rf_get_state(rf_simulation* sim, double* positions, double* orientations);
/* Positions/orientations produced by the last rf_update_fibers (dumpFibers,
 * :606-646, and the Fortran _POS/_ORIENT dumps, fortran/fibers.f95:287-303). */
RF_API rf_status rf_get_next_state(rf_simulation* sim, double* positions, double* orientations);
/* Velocities of the last rf_update_velocities (dumpVelocities, :770-822). */
RF_API rf_status rf_get_velocities(rf_simulation* sim, double* translational, double* rotational);
/* Replaces the state.  With history != 0 the previous positions,
 * orientations and velocities are set too, so the next step may use the
 * multistep update (a resume that the reference's state files cannot do). */
RF_API rf_status rf_set_state(rf_simulation* sim, const double* positions, const double* orientations,
                              int32_t history, const double* previous_positions,
                              const double* previous_orientations, const double* previous_translational,
                              const double* previous_rotational);
/* System of the last assembly (dumpLinearSystem, :648-745) and solution
 * (dumpSolutionSystem, :747-768).  a_matrix: n*n floats column-major, or NULL;
 * b_vector, x_vector: n doubles each, or NULL. */
RF_API rf_status rf_get_system(rf_simulation* sim, float* a_matrix, double* b_vector, double* x_vector);
/* rf_get_system without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Where
 * rf_get_system expands a theta-form matrix through a temporary n*n device
 * buffer, this one expands it a range of source fibers at a time through a
 * 16 MiB staging buffer that the library holds in static device memory (one
 * per GPU, kept until the process ends); RF_ERR_UNSUPPORTED when one source
 * fiber's 3N columns exceed it (3N*n*4 bytes > 16 MiB, n > 279 620 for
 * N = 5).  Calls that overlap on one GPU share that buffer and race.  Same
 * arguments and results as rf_get_system. */
RF_API rf_status rf_get_system_noalloc(rf_simulation* sim, float* a_matrix, double* b_vector, double* x_vector);
/* Number of unknowns n = 3*N*M. */
RF_API int64_t rf_system_size(const rf_simulation* sim);

RF_API rf_status rf_get_solver_stats(rf_simulation* sim, rf_solver_stats* stats);
RF_API rf_status rf_get_timings(rf_simulation* sim, rf_timings* timings);
/* Device bytes held by the simulation (admission accounting). */
RF_API int64_t rf_device_bytes(const rf_simulation* sim);

RF_API const char* rf_last_error(void);
RF_API const char* rf_version(void);

/* ------------------------------------------------------------------------
 * Serving: many independent suspensions stepped together on one GPU.
 *
 * No counterpart in the reference (one suspension per process,
 * src/main.cu:83-99); each suspension is the reference's Simulation with its
 * own fiber count, state and number of steps.  A worker thread inside the
 * server admits queued requests when a slot and device memory are free
 * (footprint = the suspension's theta-form system, velocity operator, Krylov
 * basis and state), and advances every admitted suspension one timestep per
 * tick: one launch covers the assembly of all of them, one the matrix-vector
 * products of a lockstep GMRES iteration, one the velocities, one the update.
 * Every per-suspension reduction has a fixed order, so a suspension's result
 * is bitwise the same alone and in any mix.  The shared settings come from
 * rf_server_config.base (GMRES is the solver; its tolerance, restart and
 * iteration limit apply to each suspension separately). */
typedef struct rf_server rf_server;

typedef struct rf_server_config {
  rf_config base;            /* physics and GMRES settings shared by all suspensions */
  int64_t memory_budget;     /* device bytes for admitted suspensions; 0 = 80% of the free memory */
  int32_t max_active;        /* suspensions stepped together, >= 1 */
  int32_t eager;             /* 0: each tick captures its GMRES iteration's launches once as a HIP graph
                                and replays it; 1: launch them one by one (same results) */
  int32_t reserved[4];
} rf_server_config;

typedef struct rf_request_info {
  int64_t id;
  int32_t fibers, steps;
  int32_t status;            /* RF_OK, RF_WARN_NOT_CONVERGED (some step hit the iteration limit) or an error */
  int32_t gmres_iterations;  /* summed over the steps */
  double submitted, admitted, finished;  /* seconds on the server's clock */
  int64_t footprint;         /* device bytes while admitted */
} rf_request_info;

typedef struct rf_server_stats {
  int64_t ticks, fiber_steps, completed, queued, active;
  double busy_seconds;       /* wall time spent in ticks */
  int64_t peak_bytes;        /* device bytes held by admitted suspensions, high-water mark */
  int64_t memory_budget;
} rf_server_stats;

/* Starts the worker thread on the current device. */
RF_API rf_status rf_server_create(const rf_server_config* cfg, rf_server** out);
/* Stops the worker after the current tick and frees everything. */
RF_API void rf_server_destroy(rf_server* srv);
/* Queues a suspension (positions/orientations copied); *id identifies it.
 * RF_ERR_OUT_OF_MEMORY if its footprint alone exceeds the budget. */
RF_API rf_status rf_server_submit(rf_server* srv, int32_t fibers, const double* positions,
                                  const double* orientations, int32_t steps, int64_t* id);
/* Blocks until the request has finished (timeout_s < 0: no limit); *done = 1
 * if it has. */
RF_API rf_status rf_server_wait(rf_server* srv, int64_t id, double timeout_s, int32_t* done);
/* Final state (3*fibers each, may be NULL), per-step latencies in ms (steps
 * values, may be NULL: the wall time of the tick that advanced the step) and
 * the request's record; then forgets the request. */
RF_API rf_status rf_server_collect(rf_server* srv, int64_t id, double* positions, double* orientations,
                                   double* step_ms, rf_request_info* info);
RF_API rf_status rf_server_get_stats(rf_server* srv, rf_server_stats* stats);
/* Seconds on the server's clock (the times in rf_request_info). */
RF_API double rf_server_clock(rf_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* RIGIDFIBERS_H_ */
