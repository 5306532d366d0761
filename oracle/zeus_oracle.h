/*
 * zeus_oracle.h -- CPU restatement of the reference multistart PSO -> BFGS
 * method (fnal-numerics/zeus).  TEST INFRASTRUCTURE ONLY: this library is the
 * checker that the HIP product path is compared against.  Only the test
 * suite, the smoke check and the benchmark's CPU-baseline leg load it; the
 * product never links, imports or calls it.
 *
 * Parity status per function (how each one is pinned):
 *   orc_philox4x32_10   Random123 known-answer vectors (tests/golden).
 *   orc_uniform         closed form of the 53-bit conversion + Philox KATs.
 *   orc_digamma         reference test KATs (test/dual_tests.cu:511-571) and
 *                       scipy.special.digamma vectors (tests/golden).
 *   dual arithmetic     reference test KATs (test/dual_tests.cu:192-509).
 *   objectives          reference KATs (test/fun_tests.cu:75-128), closed
 *                       forms, central finite differences, dual vs adjoint.
 *   line search         KAT alpha = 0.5 (test/bfgs_tests.cu:121-151).
 *   bfgs update         KAT (test/bfgs_tests.cu:228-264); literal O(D^4) vs
 *                       rank-2 O(D^2) brute force; secant condition H'y = s.
 *   per-start BFGS      reference launch KATs (test/bfgs_tests.cu:266-453,
 *                       554-671); quadratics reach 0, Rosenbrock reaches 1.
 *   PSO                 reference KATs (test/pso_tests.cu:27-277) and the
 *                       invariants pBestVal non-increasing, gBest = min pBest.
 *   argmin              brute force.
 * Nothing here is "parity unpinned".
 *
 * Arithmetic contract: built with -ffp-contract=off; the fused multiply-adds
 * the method uses are written out as fma() calls.  The per-thread HIP form
 * performs the same IEEE operations in the same order, so for objectives made
 * of + - * / only, oracle and GPU agree bit for bit.
 */
#ifndef ZEUS_ORACLE_H
#define ZEUS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Objective identifiers (same numbering as the product's C ABI). */
enum {
  ORC_ROSENBROCK = 1, /* examples/rosenbrock.hpp:6-41            */
  ORC_RASTRIGIN = 2,  /* examples/rastrigin.hpp:6-38             */
  ORC_ACKLEY = 3,     /* examples/ackley.hpp:10-50               */
  ORC_GAUSSIAN = 4,   /* examples/gaussian.hpp:19-33, params = C */
  ORC_LOGLIK = 5,     /* simulation/fitting/LogLikelihood.h:29-58 */
  ORC_MLP = 6,        /* 1-16-16-1 tanh least squares (no reference twin) */
  ORC_SPHERE = 7      /* sum x_i^2 (test/bfgs_tests.cu GoodObjective / Quad) */
};

enum { ORC_GRAD_ADJOINT = 0, ORC_GRAD_DUAL = 1 };
/* generators (include/traits.hpp:67-71) */
enum { ORC_PRNG_XORWOW = 0, ORC_PRNG_PHILOX = 1, ORC_PRNG_SOBOL = 2 };
enum { ORC_UPDATE_RANK2 = 0, ORC_UPDATE_LITERAL = 1 };

/* ---- counter-based random numbers ------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
/* Draw number `draw` of stream (seed, stream, start) as a double in (0,1]. */
double orc_uniform(uint32_t seed, uint32_t stream, uint64_t start,
                   uint64_t draw);

/* ---- dual numbers ----------------------------------------------------- */
typedef struct {
  double re, du;
} orc_dual;

double orc_digamma(double x);
/* op codes for orc_dual_apply (unary: a only; binary: a op b) */
enum {
  ORC_D_ADD, ORC_D_SUB, ORC_D_MUL, ORC_D_DIV, ORC_D_NEG, ORC_D_ABS,
  ORC_D_SIN, ORC_D_COS, ORC_D_EXP, ORC_D_SQRT, ORC_D_ATAN2, ORC_D_LOG,
  ORC_D_POW, ORC_D_POWD, ORC_D_LGAMMA, ORC_D_TANH
};
orc_dual orc_dual_apply(int op, orc_dual a, orc_dual b, double e);

/* ---- objectives ------------------------------------------------------- */
typedef struct {
  int kind;
  int dim;
  const double* params;
  int64_t n_params;
} orc_objective;

double orc_value(const orc_objective* f, const double* x);
orc_dual orc_value_dual(const orc_objective* f, const orc_dual* x);
void orc_grad_dual(const orc_objective* f, const double* x, double* g);
double orc_value_grad_adjoint(const orc_objective* f, const double* x,
                              double* g);

/* ---- BFGS building blocks (reference include/utils.cuh) --------------- */
double orc_dot(const double* a, const double* b, int n);
double orc_norm(const double* g, int n);
void orc_search_direction(double* p, const double* H, const double* g, int n);
void orc_update_literal(double* H, const double* s, const double* y,
                        double sTy, int n);
void orc_update_rank2(double* H, const double* s, const double* y, double sTy,
                      int n);
/* returns alpha; *n_evals receives the number of objective evaluations and
   *f_acc the value at the returned step when it was evaluated (else NaN). */
double orc_line_search(const orc_objective* f, double f0, const double* x,
                       const double* p, const double* g, int n, int* n_evals,
                       double* f_acc);

/* ---- multistart driver ------------------------------------------------ */
typedef struct {
  int objective;
  int dim;
  double lower, upper;
  int64_t n_starts;
  int required;     /* requiredConverged (<= 0 or > n_starts: never stop) */
  int max_iter;     /* MAX_ITER                                          */
  int pso_iter;     /* PSO_ITER (< 0: BFGS draws its own starts)         */
  uint32_t seed;    /* Philox key word 0                                 */
  uint32_t stream;  /* Philox key word 1                                 */
  double tol;
  const double* params;
  int64_t n_params;
  const double* x0; /* optional n_starts x dim start points (skips PSO)  */
  int prng;         /* ORC_PRNG_*                                        */
} orc_problem;

/* Generators: XORWOW state seeding (per start and coordinate) and step,
   SOBOL dimension-1 draw, and n consecutive uniforms of one stream (XORWOW:
   of (start, coord); PHILOX / SOBOL: of start) -- returns 0, -1 on bad args. */
void orc_xorwow_seed(uint32_t seed, uint32_t stream, uint64_t start, uint32_t coord,
                     uint32_t st[6]);
uint32_t orc_xorwow_next(uint32_t st[6]);
double orc_sobol_uniform(uint64_t start, uint64_t draw);
int orc_stream_uniforms(int prng, uint32_t seed, uint32_t stream, uint64_t start,
                        uint32_t coord, int64_t n, double* out);

/* Trajectory recording (include/zeus.cuh:62-103, bfgs_sequential.cuh:
   100-109, bfgs_parallel.cuh:155-184, bfgs_common.cuh:108-238): for step
   it < steps of start i < starts, coords[(it*dim + d)*starts + i],
   fval/grad/status/alpha[it*starts + i].  orc_run fills NaN / -1 / 0 first.
   alpha may be NULL. */
typedef struct {
  double* coords;
  double* fval;
  double* grad;
  int8_t* status;
  double* alpha;
  int32_t steps, starts;
} orc_traj;

typedef struct {
  int grad_mode;   /* ORC_GRAD_*                            */
  int update_mode; /* ORC_UPDATE_*                          */
  int n_threads;   /* OpenMP threads, 0 = runtime default   */
  int early_stop;  /* 0: ignore required (no stop flag)     */
  const orc_traj* traj; /* NULL: no recording               */
} orc_options;

typedef struct {
  int32_t idx, status, iter;
  double fval, gnorm;
  int32_t claimed, stopped, surrendered;
} orc_best;

/* Per-start outputs are optional (NULL to skip).  x_out is n_starts x dim.
   evals_out counts objective evaluations, grads_out gradient evaluations.
   Returns 0 on success, negative on bad arguments. */
int orc_run(const orc_problem* prob, const orc_options* opt, int32_t* status,
            int32_t* iter, double* fval, double* gnorm, double* x_out,
            int64_t* evals_out, int64_t* grads_out, double* best_x,
            orc_best* best);

/* Synchronous PSO alone: writes pBestX (n x dim), pBestVal (n), gBest. */
int orc_pso(const orc_problem* prob, double* pbest_x, double* pbest_val,
            double* gbest_x, double* gbest_val);

/* The same, also writing the final positions x and velocities v (n x dim
   each; NULL to skip). */
int orc_pso_swarm(const orc_problem* prob, double* x, double* v,
                  double* pbest_x, double* pbest_val, double* gbest_x,
                  double* gbest_val);

/* One PSO velocity/position step with explicit coefficients and uniforms
   (used by the reference-KAT tests): r has 2*n*dim entries (r1,r2 per d). */
void orc_pso_step(int n, int dim, double w, double c1, double c2,
                  const double* r, double* X, double* V, const double* pbx,
                  const double* gbx);

/* Deterministic argmin: NaN entries skipped, lowest index wins ties. */
int64_t orc_argmin(const double* v, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
