/*
 * zeus_c.h -- C ABI of the MI355X (gfx950) Zeus multistart optimizer
 * (libzeus_amd.so).  Plain C: fixed-width integers, doubles and pointers; no
 * C++ types and no exceptions cross this line.
 *
 * The reference (fnal-numerics/zeus) has no FFI: its boundary is the C++
 * template zeus::Zeus (include/zeus.cuh:341-375) instantiated in the caller's
 * translation unit.  This ABI holds everything that does not depend on a
 * user functor: the whole PSO -> BFGS -> argmin pipeline for the built-in
 * objectives, its PSO and argmin phases on their own, device memory, the
 * counter-based random stream and the test-only fault hooks.  The C++
 * front end (zeus_amd/csrc/zeus/zeus.hpp) keeps the reference's template
 * API for user functors and calls into this ABI for the functor-free parts.
 *
 * Error convention: every call returns zeus_status_t: 0 ok, 3 device
 * allocation failure, 4 launch or runtime error, negative = API misuse
 * (see ZEUS_E_*).  Per-start status codes (0..5) are reported in
 * zeus_result.status.  Calls are synchronous with respect to the host and
 * one call per device may be in flight at a time.
 */
#ifndef ZEUS_C_H
#define ZEUS_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t zeus_status_t;

/* return codes */
#define ZEUS_OK 0
#define ZEUS_E_ALLOC 3
#define ZEUS_E_LAUNCH 4
#define ZEUS_E_ARG (-10)        /* invalid argument value                */
#define ZEUS_E_UNSUPPORTED (-11) /* objective/dimension not instantiated */
#define ZEUS_E_NODEVICE (-12)   /* no usable gfx950 device               */
#define ZEUS_E_IO (-13)         /* a file could not be written           */
#define ZEUS_E_SCRATCH (-14)    /* request exceeds the fixed scratch of an
                                   allocation-free entry point            */

/* per-start / per-call status codes (reference README.md:27-35) */
#define ZEUS_UNSET (-1)
#define ZEUS_SURRENDERED 0
#define ZEUS_CONVERGED 1
#define ZEUS_STOPPED 2
#define ZEUS_MALLOC 3
#define ZEUS_KERNEL 4
#define ZEUS_NONFINITE 5

/* built-in objectives (reference examples/{rosenbrock,rastrigin,ackley,
   gaussian}.hpp, examples/simulation/fitting/LogLikelihood.h) */
#define ZEUS_OBJ_ROSENBROCK 1 /* params: none                              */
#define ZEUS_OBJ_RASTRIGIN 2  /* params: none                              */
#define ZEUS_OBJ_ACKLEY 3     /* params: none                              */
#define ZEUS_OBJ_GAUSSIAN 4   /* params: C, dim x dim row-major            */
#define ZEUS_OBJ_LOGLIK 5     /* params: x[n] k[n] lgamma(k+1)[n]; dim 4   */
#define ZEUS_OBJ_MLP 6        /* params: x[m] y[m]; dim 321                */
#define ZEUS_OBJ_SPHERE 7     /* params: none                              */

/* Generators (reference PRNGType, include/traits.hpp:67-71), one stream of
   uniforms in (0, 1]: XORWOW (Marsaglia 2003; a state per (start,
   coordinate) seeded from (seed, stream, start, coordinate)), PHILOX
   (counter-based Philox4x32-10 keyed by (seed, stream), one stream per
   start), SOBOL (dimension 1 of Sobol's sequence read from index start on;
   seed unused).  None reproduces cuRAND's streams. */
#define ZEUS_PRNG_XORWOW 0
#define ZEUS_PRNG_PHILOX 1
#define ZEUS_PRNG_SOBOL 2

#define ZEUS_GRAD_ADJOINT 0
#define ZEUS_GRAD_DUAL 1

#define ZEUS_FORM_AUTO 0   /* autotuner's choice among all forms          */
#define ZEUS_FORM_THREAD 1 /* one lane per start (reference sequential)   */
#define ZEUS_FORM_GROUP 2  /* cooperative forms (reference parallel): one
                              workgroup per start, or -- Gaussian, adjoint
                              gradient, unless group_threads is set -- the
                              tile form, or -- separable objectives of
                              small dim, unless group_threads is set --
                              the team form; the autotuner picks          */
#define ZEUS_FORM_TILE 3   /* one workgroup per tile of 16 starts, the
                              objective evaluated for the tile on FP64 MFMA
                              (Gaussian with dim >= 16, adjoint gradient)  */
#define ZEUS_FORM_TEAM 4   /* teams of group_threads = 4, 8 or 16 lanes per
                              start, 64 / team starts per wavefront
                              (Rosenbrock, Rastrigin, Ackley, sphere with
                              small dim); part of ZEUS_FORM_GROUP and AUTO  */

/* Trajectory recording of one problem (reference zeus.cuh:62-103,189-239;
   recording in bfgs_sequential.cuh:100-109 and bfgs_parallel.cuh:155-184).
   Caller-owned DEVICE buffers, filled by the call: coords NaN, fval and
   grad NaN, status -1, alpha 0, then for iteration `it` < steps of start
   `i` < starts (problem-local):
     coords[(it * dim + d) * starts + i]  the point at the top of iteration it
     fval[it * starts + i]                its value (best so far)
     grad[it * starts + i]                its gradient norm
     status[it * starts + i]              -1, or the final status at the
                                          iteration it was decided (0 at
                                          MAX_ITER - 1 for surrendered)
     alpha[it * starts + i]               the line-search step (NULL: none;
                                          the reference records it only in
                                          its parallel form)
   steps and starts bound the buffers (the reference keeps all MAX_ITER x N).
   A traced problem runs the per-thread or per-workgroup form (not the tile
   form) and its call is never replayed from a graph. */
typedef struct zeus_trajectory {
  double* coords;  /* steps x dim x starts */
  double* fval;    /* steps x starts       */
  double* grad;    /* steps x starts       */
  int8_t* status;  /* steps x starts       */
  double* alpha;   /* steps x starts, or NULL */
  int32_t steps;   /* iterations kept, >= 0 */
  int32_t starts;  /* starts kept (the first ones), >= 0 */
} zeus_trajectory;

/* One optimization problem: the arguments of zeus::Zeus for a built-in
   objective.  Replaces the template parameters/arguments of
   include/zeus.cuh:341-358 (f, lower, upper, N, MAX_ITER, PSO_ITER,
   requiredConverged, fun_name, tolerance, seed). */
typedef struct zeus_problem {
  int32_t objective;   /* ZEUS_OBJ_*                                       */
  int32_t dim;         /* D                                                */
  double lower, upper; /* box [lower, upper]^D                             */
  int64_t n_starts;    /* N                                                */
  int32_t max_iter;    /* MAX_ITER                                         */
  int32_t pso_iter;    /* PSO_ITER (< 0: no PSO, BFGS draws its starts)    */
  int32_t required;    /* requiredConverged (<= 0: never stop early)       */
  uint32_t seed;       /* Philox key word 0                                */
  uint32_t stream;     /* Philox key word 1 (distinct streams per problem) */
  int32_t prng;        /* ZEUS_PRNG_* (0: XORWOW, the reference default)   */
  double tol;          /* gradient-norm tolerance                          */
  const double* params;/* device pointer, objective data (see ZEUS_OBJ_*)  */
  int64_t n_params;
  const double* x0;    /* optional device start points N x D (skips PSO)   */
  const char* name;    /* fun_name: selects the known minimum for the
                          "actual" count (reference src/util.cpp:11-41)    */
  const zeus_trajectory* trajectory; /* optional (NULL: not recorded)      */
} zeus_problem;

typedef struct zeus_options {
  int32_t gradient;      /* ZEUS_GRAD_*                                    */
  int32_t form;          /* ZEUS_FORM_*                                    */
  int32_t group_threads; /* workgroup size of the group form (0 = tuned)   */
  int32_t nzerosteps;    /* accepted for API parity; inert (the line search
                            never returns a zero step)                     */
  int32_t use_graph;     /* capture the launch sequence once per batch
                            shape and replay it as a hipGraph (once the
                            autotuner has settled the launch shapes; the
                            first call of a shape runs uncaptured)         */
  int32_t flags;         /* ZEUS_FLAG_* bits                               */
} zeus_options;

/* zeus_options.flags: time gradient evaluations and H updates per start in
   the per-thread form and report them in zeus_result.ad / .bfgs (reference
   bfgs_sequential.cuh:236-382; the reference's parallel form reports zeros,
   and so do the workgroup forms here). */
#define ZEUS_FLAG_METRICS 1

/* Timing summary of one kind of work over a problem's starts (reference
   Metrics, include/util.hpp:16-25; computed as report_metrics_and_cleanup,
   bfgs_sequential.cuh:240-382, from a constant-rate device clock). */
typedef struct zeus_metrics {
  double ms_per_call;           /* mean time per call                      */
  double calls_per_thread_mean; /* mean calls per start                    */
  double fraction_of_kernel;    /* mean time per start / kernel time       */
  double block95;               /* 95th percentile over 256-start blocks of
                                   (summed block time / kernel time)       */
  double serialized;            /* summed time over starts / kernel time   */
  double fraction_of_thread;    /* mean over starts of time / start's total */
} zeus_metrics;

/* Per-problem outcome: the fields of zeus::Result<D> (include/util.hpp:36-54)
   except the coordinates, which go to the best_x argument. */
typedef struct zeus_result {
  int32_t idx;          /* best start (argmin of fval, NaN skipped)        */
  int32_t status;
  int32_t iter;
  int32_t nonfinite;    /* number of status-5 starts                       */
  double fval;
  double gradient_norm;
  int32_t claimed, actual, surrendered, stopped; /* util.hpp:27-32        */
  double ms_pso;        /* device time of the PSO phase of this problem's
                           class (problems of one objective and dim; 0
                           when replayed from a graph)                     */
  double ms_opt;        /* device time of its class's BFGS phase (of the
                           whole replayed graph when use_graph)            */
  double ms_total;      /* host wall time of the call                      */
  int64_t sum_iter;     /* sum over starts of BFGS iterations              */
  int64_t sum_ls_evals; /* sum over starts of line-search evaluations      */
  int64_t sum_grads;    /* sum over starts of gradient evaluations         */
  zeus_metrics ad;      /* gradient evaluations (ZEUS_FLAG_METRICS)        */
  zeus_metrics bfgs;    /* H updates (ZEUS_FLAG_METRICS)                   */
} zeus_result;

/* Optional per-start host outputs, concatenated over problems in order;
   any pointer may be NULL.  x holds sum_i n_starts_i * dim_i doubles. */
typedef struct zeus_start_outputs {
  int32_t* status;
  int32_t* iter;
  double* fval;
  double* gnorm;
  double* x;
  int32_t* ls_evals;
  int32_t* grads;
} zeus_start_outputs;

/* The whole pipeline (PSO -> BFGS -> argmin) for a batch of problems, in one
   stream-ordered launch sequence (captured as a hipGraph with use_graph).
   Replaces impl::ZeusInternal (include/zeus.cuh:37-242) for built-in
   objectives.  results[n_probs]; best_x: sum_i dim_i host doubles.
   stream: a hipStream_t, or NULL for the library's own stream. */
zeus_status_t zeus_run_batch(const zeus_problem* probs, int32_t n_probs,
                             const zeus_options* opt, zeus_result* results,
                             double* best_x, const zeus_start_outputs* starts,
                             void* stream);

/* PSO phase alone (reference pso::launch, include/pso.cuh:194-345): device
   outputs pbest_x (N x D), pbest_val (N), gbest_x (D), gbest_val (1). */
zeus_status_t zeus_pso_run(const zeus_problem* prob, double* pbest_x,
                           double* pbest_val, double* gbest_x,
                           double* gbest_val, void* stream);

/* zeus_pso_run without any allocation: no counterpart in the reference.  For
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
   kernels and results (bit for bit) as zeus_pso_run; every buffer beyond
   the caller's outputs (positions, velocities, per-workgroup partials, the
   Gaussian's prepared matrix, XORWOW states, the control block) lives in a
   fixed 32 MiB device array of the library, one per GPU, held until the
   process ends.  A problem whose buffers exceed it returns ZEUS_E_SCRATCH
   (up to about 0.8e6 start-coordinates with XORWOW draws and 2e6 with
   PHILOX / SOBOL: 4096 starts x 150 dimensions fit under every generator).  The array keeps the last call's
   contents; calls are serialized by the library and each returns after its
   stream has drained, so two calls in one process never overlap (calls that
   overlapped on one GPU would race on the array). */
zeus_status_t zeus_pso_run_noalloc(const zeus_problem* prob, double* pbest_x,
                                   double* pbest_val, double* gbest_x,
                                   double* gbest_val, void* stream);

/* The swarm of one problem, caller-owned device buffers: positions and
   velocities (N x D, start-major), pBest coordinates and values (N x D, N),
   gBest coordinates and value (D, 1), and for the XORWOW generator the
   (start, coordinate) states (6 x N x D uint32; NULL otherwise). */
typedef struct zeus_swarm {
  double* x;
  double* v;
  double* pbest_x;
  double* pbest_val;
  double* gbest_x;
  double* gbest_val;
  uint32_t* rng;
} zeus_swarm;

/* PSO phases (reference pso::initKernel and pso::iterKernel,
   include/pso.cuh:17-155): zeus_pso_init draws positions and velocities,
   sets pBest = X and computes gBest; zeus_pso_step runs iteration iter
   (>= 1) with the swarm's gBest frozen, then updates pBest (strict
   improvement) and gBest.  init followed by steps 1 .. T gives the same
   swarm as zeus_pso_run with pso_iter = T. */
zeus_status_t zeus_pso_init(const zeus_problem* prob, const zeus_swarm* swarm,
                            void* stream);
zeus_status_t zeus_pso_step(const zeus_problem* prob, int32_t iter,
                            const zeus_swarm* swarm, void* stream);

/* BFGS update of a batch of n x n matrices (device, row-major) with device
   vectors s, y (batch x n each): reference util::bfgsUpdate
   (include/utils.cuh:245-287), skipped when |s.y| < 1e-14, in the O(n^2)
   rank-2 form H' = H + s a^T + a s^T (a = (beta/2) s - rho H y, beta =
   rho^2 y.Hy + rho, rho = 1/s.y).  n <= 1024. */
zeus_status_t zeus_bfgs_update(double* H, const double* s, const double* y,
                               int64_t batch, int32_t n, void* stream);

/* Objective values (and gradients when grad != NULL) at n points (device
   arrays, x: n x dim), lane per point; gradient ZEUS_GRAD_*. */
zeus_status_t zeus_eval(const zeus_problem* prob, const double* x, int64_t n,
                        double* f, double* grad, int32_t gradient,
                        void* stream);

/* Deterministic argmin of n device doubles (NaN skipped, ties to the lowest
   index; all-NaN -> 0).  Replaces bfgs::launchReduction's CUB ArgMin
   (include/bfgs_common.cuh:280-319).  *idx is a host int64. */
zeus_status_t zeus_argmin(const double* values, int64_t n, int64_t* idx,
                          void* stream);

/* Counter-based Philox4x32-10 draw in (0,1] for (seed, stream, start, draw):
   host evaluation, and a device fill of n consecutive draws of one start
   (replaces the cuRAND states of src/utils.cu:17-50). */
double zeus_philox_uniform(uint32_t seed, uint32_t stream, uint64_t start,
                           uint64_t draw);
zeus_status_t zeus_philox_fill(double* out, int64_t n, uint32_t seed,
                               uint32_t stream, uint64_t start,
                               uint64_t draw0, void* stream_handle);

/* Device fill of draws 0..n-1 of one stream under generator prng
   (ZEUS_PRNG_*), in the order the optimizer consumes them: XORWOW, the
   stream of (start, coord); PHILOX and SOBOL, the draw indices of start
   (coord unused). */
zeus_status_t zeus_stream_fill(double* out, int64_t n, int32_t prng, uint32_t seed,
                               uint32_t stream, uint64_t start, uint32_t coord,
                               void* stream_handle);

/* Device memory (replaces CudaBuffer's cudaMalloc/cudaMemcpy,
   include/cuda_buffer.hpp:14-209). */
zeus_status_t zeus_alloc(void** ptr, uint64_t bytes);
zeus_status_t zeus_free(void* ptr);
zeus_status_t zeus_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
zeus_status_t zeus_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
zeus_status_t zeus_memcpy_d2d(void* dst, const void* src, uint64_t bytes);
zeus_status_t zeus_synchronize(void);

/* Write trajectory HOST arrays (the layout of zeus_trajectory, with
   starts = n) to `filename`: reference util::writeTrajectoryData
   (src/utils.cu:126-286).  ".tsv": one row per (start, step) with columns
   traj step fval grad status alpha x0 .. x{dim-1}, values in %.17e, NaN as
   "NaN".  ".nc": NetCDF4 is not built in -> ZEUS_E_UNSUPPORTED (the
   reference's answer when built without NetCDF).  Any other extension ->
   ZEUS_E_ARG.  alpha may be NULL (written as 0).  I/O failure -> ZEUS_E_IO. */
zeus_status_t zeus_write_trajectory(const char* filename, const double* coords,
                                    const double* fval, const double* grad,
                                    const int8_t* status, const double* alpha,
                                    int32_t n, int32_t steps, int32_t dim);

/* Metrics from per-start host arrays: ticks[n] and calls[n] of one kind of
   work, total[n] the starts' whole times (NULL: fraction_of_thread = 0),
   starts grouped in blocks of threads_per_block for block95, ms_kernel the
   kernel time, tick_hz the clock rate. */
zeus_status_t zeus_metrics_from_ticks(const uint64_t* ticks, const int32_t* calls,
                                      const uint64_t* total, int64_t n,
                                      int32_t threads_per_block, double ms_kernel,
                                      double tick_hz, zeus_metrics* out);

/* Distance of x from the known minimum implied by fun_name (reference
   util::calculateEuclideanError, src/util.cpp:11-41): "rosenbrock" -> all
   ones, "goldstein_price" -> (0, -1) (NaN unless dim == 2), otherwise the
   origin. */
double zeus_euclidean_error(const char* fun_name, const double* x, int32_t dim);

/* Append one row for a call to the results table `path` (NULL:
   "zeus_<dim>d_results.tsv" in the working directory), writing the header
   first when the file is empty: reference util::appendResultsToTsv
   (include/util.hpp:78-149), same columns and number formats.  best holds
   the Result fields (and metrics), coords its dim coordinates. */
zeus_status_t zeus_append_results_tsv(const char* path, int32_t dim, int64_t n,
                                      const char* fun_name, double ms_init,
                                      double ms_pso, double ms_opt, double ms_rand,
                                      int32_t max_iter, int32_t pso_iter,
                                      double error, int32_t run,
                                      const zeus_result* best,
                                      const double* coords);

/* Name of a status code ("surrendered", "converged", ...). */
const char* zeus_status_string(int32_t status);

/* Test-only fault injection: fail the allocation number `fail_alloc_at`
   (counted from this call, 1-based; 0 disables) and/or report the next
   `fail_launches` kernel launches as failed.  Forces status 3 / 4 paths
   without exhausting or crashing the device. */
zeus_status_t zeus_set_fault_hooks(int32_t fail_alloc_at, int32_t fail_launches);

/* Form and workgroup size the last zeus_run_batch used for (objective, dim,
   gradient) -- the autotuner's choice unless the options fixed them -- for
   inspection: *form (ZEUS_FORM_*), *threads. */
zeus_status_t zeus_tuned_config(int32_t objective, int32_t dim,
                                int32_t gradient, int32_t* form,
                                int32_t* threads);

/* Synthetic inputs of the benchmark configurations, generated on the host
   from Philox streams keyed by `seed` (documented in zeus_c.hip):
   Poisson-sampled binned dijet spectra (x, k, lgamma(k+1) for n_bins bins
   of m in [1000, 7000] GeV, x = m/13000, one experiment), MLP training data
   (x ~ U[-1,1], y = sin(3x) + N(0, 0.1^2)) and SPD matrices
   C = A A^T / D + I with A ~ N(0,1). */
zeus_status_t zeus_make_dijet(uint32_t seed, uint32_t experiment, int32_t n_bins,
                              double* params_out /* 3 * n_bins */);
zeus_status_t zeus_make_mlp_data(uint32_t seed, int32_t m,
                                 double* params_out /* 2 * m */);
zeus_status_t zeus_make_spd(uint32_t seed, uint32_t stream, int32_t dim,
                            double* c_out /* dim * dim */);

#ifdef __cplusplus
}
#endif
#endif
