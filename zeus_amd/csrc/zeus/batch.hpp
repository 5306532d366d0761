// batch.hpp -- device-side description of a batch of independent problems.
//
// Every launch works on a *class* of problems that share the objective type
// and the dimension D (a compile-time constant of the kernel); problems of a
// class differ in box, start count, tolerances, seed and parameter block.
// Starts of a class are numbered globally and contiguously:
// problem p owns starts [start0, start0 + n_starts).
#pragma once

#include "config.hpp"

namespace zeus::dev {

// Trajectory buffers of one problem (what reference include/zeus.cuh:62-103
// allocates and bfgs_sequential.cuh:100-109 fills): coordinates in
// [step][dim][start] order, fval (best value so far), grad (its gradient
// norm), status and alpha (line-search step) in [step][start] order.  Only
// the first `starts` starts and the first `steps` iterations are kept, so
// the buffers can be bounded (the reference keeps all N x MAX_ITER).
// fval == nullptr: nothing is recorded; alpha may be null on its own.
struct TrajOut {
  double* coords;
  double* fval;
  double* grad;
  int8_t* status;
  double* alpha;
  int32_t steps;
  int32_t starts;
};

struct ProblemDesc {
  int32_t objective;
  int32_t dim;
  double lower, upper;
  int64_t start0;   // first global start of this problem
  int64_t n_starts;
  int32_t required; // requiredConverged (<= 0: never stop early)
  int32_t max_iter; // MAX_ITER
  int32_t pso_iter; // PSO_ITER (< 0: no PSO, starts drawn in the BFGS)
  int32_t block0;   // first PSO workgroup of this problem
  int32_t min_kind; // known minimum for the "actual" count: 0 origin,
                    // 1 all ones (rosenbrock), 2 (0, -1) (goldstein_price)
  int32_t prng;     // generator (prng.hpp: 0 XORWOW, 1 PHILOX, 2 SOBOL)
  uint32_t seed;    // Philox key word 0
  uint32_t stream;  // Philox key word 1
  double tol;
  const double* params; // device
  int64_t n_params;
  const double* x0;     // optional device start points (n_starts x dim)
  const double* aux;    // library-prepared objective data (Gaussian: Cs)
  const double* aux_frag; // Gaussian: Cs in MFMA A-operand order (cs_product.hpp)
  TrajOut traj;           // trajectory recording (traj.fval == nullptr: off)
};

// Per-problem control block; zeroed before each call (reset_control).
struct ControlHead {
  int32_t stop_flag;       // set once `required` starts have converged
  int32_t converged;       // number of status-1 starts so far
  uint32_t pad0;
  int32_t best_idx;        // argmin start (local index), summary kernel
  double gbest_val;        // PSO global best value
  int32_t claimed, stopped, surrendered, nonfinite, actual, pad1;
  int64_t sum_iter, sum_ls_evals, sum_grads;
  int32_t best_status, best_iter; // the best start's Result fields
  double best_fval, best_gnorm;
};
// The tile form's converged count, split over kConvLines counters 256 bytes
// apart (workgroup b adds to counter b % kConvLines, one per die under the
// round-robin dispatch): the tiles of the chip add their counts in the same
// rounds, and device-scope adds to one address queue behind each other.
constexpr int kConvLines = 8;
constexpr int kConvStride = 64; // 32-bit words between counters
struct Control : ControlHead {
  int32_t conv_lines[kConvLines * kConvStride];
};

// Zero a control block (the counters' padding words are never used).
ZEUS_D ZEUS_FI void
reset_control(Control* c)
{
  ControlHead z{};
  z.best_idx = -1;
  z.gbest_val = __builtin_huge_val();
  *static_cast<ControlHead*>(c) = z;
#pragma unroll
  for (int k = 0; k < kConvLines; ++k)
    c->conv_lines[k * kConvStride] = 0;
}

// Per-start outputs (structure of arrays over global start index).
struct StartOut {
  int32_t* status;
  int32_t* iter;
  double* fval;
  double* gnorm;
  double* x;        // n x D
  int32_t* ls_evals; // objective evaluations made by line searches
  int32_t* grads;    // gradient evaluations
  int8_t* near;      // 1: converged within 0.5 (Euclidean) of the known minimum
  // Optional per-start timing of the per-thread form (reference
  // bfgs_sequential.cuh:46-233): constant-rate clock ticks spent in gradient
  // evaluations and in H updates, their call counts, and the start's total.
  // ad_ticks == nullptr: not collected.
  uint64_t* ad_ticks;
  int32_t* ad_calls;
  uint64_t* bfgs_ticks;
  int32_t* bfgs_calls;
  uint64_t* total_ticks;
};

// Coordinate d of the known minimum selected by ProblemDesc::min_kind
// (reference src/util.cpp:11-41: rosenbrock -> ones, goldstein_price ->
// (0, -1), anything else -> origin).
ZEUS_HD ZEUS_FI double
known_min(int kind, int d)
{
  return kind == 1 ? 1.0 : (kind == 2 && d == 1 ? -1.0 : 0.0);
}

struct ClassArgs {
  const ProblemDesc* probs;
  int32_t n_probs;
  int32_t dim;
  int64_t total_starts;
  Control* ctl;           // n_probs entries
  const double* starts;   // PSO pBestX (total x D) or nullptr
  StartOut out;
  double* h_slab;         // group form: global H storage when not resident
  int64_t h_slab_stride;  // doubles per workgroup slab
};

// Problem owning global start s (problems are sorted by start0).
ZEUS_HD ZEUS_FI int
find_problem(const ProblemDesc* probs, int n, int64_t s)
{
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].start0 <= s)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Agent-scope atomics on the control block (vector memory instructions).
ZEUS_D ZEUS_FI int
load_flag(const int32_t* p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
ZEUS_D ZEUS_FI void
store_flag(int32_t* p, int v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
ZEUS_D ZEUS_FI int
fetch_add(int32_t* p, int v)
{
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Trajectory slot of problem-local start sl at iteration iter, or -1 when
// that step is not kept.
ZEUS_HD ZEUS_FI int64_t
traj_slot(const TrajOut& t, int64_t sl, int iter)
{
  if (t.fval == nullptr || sl >= t.starts || iter < 0 || iter >= t.steps)
    return -1;
  return int64_t(iter) * t.starts + sl;
}

// The status of a finished start, at the iteration the reference writes it
// (bfgs_common.cuh:108-238): the breaking iteration for stopped (2),
// non-finite (5) and converged (1); MAX_ITER - 1 for surrendered (0).
ZEUS_D ZEUS_FI void
traj_final_status(const TrajOut& t, int64_t sl, int status, int iter, int max_iter)
{
  const int at = status == kStatusSurrendered ? max_iter - 1 : iter;
  const int64_t o = traj_slot(t, sl, at);
  if (o >= 0)
    t.status[o] = int8_t(status);
}

// Count a converged start and raise the stop flag when it is the
// `required`-th one (reference include/bfgs_common.cuh:170-208).
ZEUS_D ZEUS_FI void
count_converged(Control* c, int required)
{
  const int old = fetch_add(&c->converged, 1);
  if (old + 1 == required)
    store_flag(&c->stop_flag, 1);
}

} // namespace zeus::dev
