// zeus_c.hip -- implementation of the C ABI (include/zeus_c.h).
//
// Host orchestration of the multistart pipeline for the built-in
// objectives: problems are grouped into classes (same objective and
// dimension), each class runs PSO -> BFGS -> summary on one stream with no
// host synchronisation in between, and only the per-problem summaries and
// best coordinates come back to the host.  Restates the control flow of the
// reference's impl::ZeusInternal (include/zeus.cuh:37-242),
// pso::launch (include/pso.cuh:194-345) and bfgs::*::launch
// (bfgs_sequential.cuh:384-589, bfgs_parallel.cuh:301-472).
#include "zeus_c.h"

#include "runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace zeus::rt {

const ClassOps*
find_class(int objective, int dim)
{
  switch (objective) {
    case obj::kRosenbrock: return ops_rosenbrock(dim);
    case obj::kRastrigin: return ops_rastrigin(dim);
    case obj::kAckley: return ops_ackley(dim);
    case obj::kGaussian: return ops_gaussian(dim);
    case obj::kLogLikelihood: return ops_loglik(dim);
    case obj::kMlp: return ops_mlp(dim);
    case obj::kSphere: return ops_sphere(dim);
    default: return nullptr;
  }
}

namespace {
int g_fail_alloc_at = 0; // 1-based allocation number to fail, 0 = off
int g_alloc_count = 0;
int g_fail_launches = 0;
} // namespace

bool
fault_take_launch()
{
  if (g_fail_launches > 0) {
    --g_fail_launches;
    return true;
  }
  return false;
}

} // namespace zeus::rt

namespace {

using namespace zeus;
using zeus::dev::ClassArgs;
using zeus::dev::Control;
using zeus::dev::ProblemDesc;
using zeus::dev::PsoArgs;
using zeus::dev::StartOut;
using zeus::rt::ClassOps;

std::mutex g_mu;

// Autotuner cache key: objective, dimension, gradient method and the class's
// start count rounded down to a power of two.
struct TuneKey {
  int objective, dim, gradient, bucket;
  bool operator<(const TuneKey& o) const
  {
    return std::tie(objective, dim, gradient, bucket) <
           std::tie(o.objective, o.dim, o.gradient, o.bucket);
  }
};
struct TuneVal {
  int form, threads;
};

// timing events (see ensure_device's note on the system-scope fence)
constexpr unsigned kTimingEventFlags = hipEventDisableSystemFence;

struct DeviceState {
  bool init = false;
  int dev = 0;
  int n_cu = 0;
  double tick_hz = 1e8; // rate of the device's constant clock (wall_clock64)
  hipStream_t stream = nullptr;
  hipEvent_t ev[6] = {};
  // classes of a settled batch run side by side on these streams (default
  // priority), forked from and joined back into the caller's stream
  static constexpr int kSide = 4;
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSide] = {};
  std::vector<hipEvent_t> cls_ev; // per class: PSO start, PSO end, BFGS end
  void* arena = nullptr;
  size_t arena_bytes = 0;
  void* pinned = nullptr; // host mirror of the arena's upload/download regions
  size_t pinned_bytes = 0;
  std::map<TuneKey, TuneVal> tuned; // autotuner results (only tuned choices)
  std::map<std::tuple<int, int, int>, TuneVal> last; // last shape used per class and gradient
  struct Graph {
    std::vector<uint64_t> key; // batch shape, arena and stream it was captured for
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
};
DeviceState g_ds;

hipError_t
dev_malloc(void** p, size_t bytes)
{
  ++rt::g_alloc_count;
  if (rt::g_fail_alloc_at > 0 && rt::g_alloc_count == rt::g_fail_alloc_at) {
    rt::g_fail_alloc_at = 0;
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  return hipMalloc(p, bytes);
}

zeus_status_t
to_status(hipError_t e)
{
  if (e == hipSuccess)
    return ZEUS_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return ZEUS_E_ALLOC;
  return ZEUS_E_LAUNCH;
}

zeus_status_t
ensure_device()
{
  if (g_ds.init)
    return ZEUS_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return ZEUS_E_NODEVICE;
  if (hipGetDevice(&g_ds.dev) != hipSuccess)
    return ZEUS_E_NODEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, g_ds.dev) != hipSuccess)
    return ZEUS_E_NODEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return ZEUS_E_NODEVICE; // code objects are built for gfx950 only
  g_ds.n_cu = prop.multiProcessorCount;
  int wall_khz = 0;
  if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, g_ds.dev) ==
        hipSuccess &&
      wall_khz > 0)
    g_ds.tick_hz = double(wall_khz) * 1e3;

  if (hipStreamCreateWithFlags(&g_ds.stream, hipStreamNonBlocking) != hipSuccess)
    return ZEUS_E_LAUNCH;
  // Events order work on this device and time it; none is used to make
  // memory visible to the host (the results come back through a stream-
  // ordered copy and a stream synchronisation), so they are recorded without
  // the system-scope fence, whose cache writeback and invalidation otherwise
  // idles the GPU for several microseconds at every record between kernels.
  for (auto& e : g_ds.ev)
    if (hipEventCreateWithFlags(&e, kTimingEventFlags) != hipSuccess)
      return ZEUS_E_LAUNCH;
  for (auto& sd : g_ds.side)
    if (hipStreamCreateWithFlags(&sd, hipStreamNonBlocking) != hipSuccess)
      return ZEUS_E_LAUNCH;
  if (hipEventCreateWithFlags(&g_ds.fork, hipEventDisableTiming | hipEventDisableSystemFence) !=
      hipSuccess)
    return ZEUS_E_LAUNCH;
  for (auto& e : g_ds.join)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence) !=
        hipSuccess)
      return ZEUS_E_LAUNCH;
  g_ds.init = true;
  return ZEUS_OK;
}

// Grow-only device arena; calls are host-synchronous, so a resize never
// races with work in flight.
zeus_status_t
arena_reserve(size_t bytes)
{
  if (bytes <= g_ds.arena_bytes)
    return ZEUS_OK;
  if (g_ds.arena) {
    (void)hipFree(g_ds.arena);
    g_ds.arena = nullptr;
    g_ds.arena_bytes = 0;
  }
  for (auto& ge : g_ds.graphs) // captured graphs point into the old arena
    (void)hipGraphExecDestroy(ge.exec);
  g_ds.graphs.clear();
  const size_t want = bytes + bytes / 4 + (1u << 20);
  void* p = nullptr;
  const hipError_t e = dev_malloc(&p, want);
  if (e != hipSuccess)
    return ZEUS_E_ALLOC;
  g_ds.arena = p;
  g_ds.arena_bytes = want;
  return ZEUS_OK;
}

// Grow-only pinned host buffer (a fresh allocation: pinned memory is
// allocated, never registered after the fact).
zeus_status_t
pinned_reserve(size_t bytes)
{
  if (bytes <= g_ds.pinned_bytes)
    return ZEUS_OK;
  if (g_ds.pinned) {
    (void)hipHostFree(g_ds.pinned);
    g_ds.pinned = nullptr;
    g_ds.pinned_bytes = 0;
  }
  const size_t want = bytes + bytes / 4 + 4096;
  void* p = nullptr;
  if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess)
    return ZEUS_E_ALLOC;
  g_ds.pinned = p;
  g_ds.pinned_bytes = want;
  return ZEUS_OK;
}

// Two-pass carving: with base == nullptr only measures.
struct Carver {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t n)
  {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += n * sizeof(T);
    return p;
  }
};

int
min_kind_of(const char* name)
{
  if (!name)
    return 0;
  if (std::strcmp(name, "rosenbrock") == 0)
    return 1;
  if (std::strcmp(name, "goldstein_price") == 0)
    return 2;
  return 0;
}

zeus_status_t
validate(const zeus_problem& p)
{
  if (p.dim <= 0 || p.n_starts < 0 || p.max_iter < 0 || p.n_starts > (int64_t(1) << 40))
    return ZEUS_E_ARG;
  switch (p.objective) {
    case ZEUS_OBJ_ROSENBROCK:
    case ZEUS_OBJ_RASTRIGIN:
    case ZEUS_OBJ_ACKLEY:
    case ZEUS_OBJ_SPHERE: break;
    case ZEUS_OBJ_GAUSSIAN:
      if (!p.params || p.n_params != int64_t(p.dim) * p.dim)
        return ZEUS_E_ARG;
      break;
    case ZEUS_OBJ_LOGLIK:
      if (p.dim != 4 || !p.params || p.n_params <= 0 || p.n_params % 3 != 0)
        return ZEUS_E_ARG;
      break;
    case ZEUS_OBJ_MLP:
      if (p.dim != 321 || !p.params || p.n_params <= 0 || p.n_params % 2 != 0)
        return ZEUS_E_ARG;
      break;
    default: return ZEUS_E_ARG;
  }
  if (p.prng != ZEUS_PRNG_XORWOW && p.prng != ZEUS_PRNG_PHILOX && p.prng != ZEUS_PRNG_SOBOL)
    return ZEUS_E_ARG;
  if (const zeus_trajectory* t = p.trajectory) {
    if (t->steps < 0 || t->starts < 0)
      return ZEUS_E_ARG;
    if (int64_t(t->steps) * t->starts > 0 && (!t->coords || !t->fval || !t->grad || !t->status))
      return ZEUS_E_ARG;
  }
  if (!rt::find_class(p.objective, p.dim))
    return ZEUS_E_UNSUPPORTED;
  return ZEUS_OK;
}

// Device view of a problem's trajectory request (fval == nullptr: off).
dev::TrajOut
traj_of(const zeus_problem& p)
{
  dev::TrajOut t{};
  const zeus_trajectory* u = p.trajectory;
  if (!u || int64_t(u->steps) * u->starts == 0)
    return t;
  t.coords = u->coords;
  t.fval = u->fval;
  t.grad = u->grad;
  t.status = u->status;
  t.alpha = u->alpha;
  t.steps = u->steps;
  t.starts = u->starts;
  return t;
}

void
default_result(zeus_result& r)
{
  std::memset(&r, 0, sizeof(r));
  r.idx = -1;
  r.status = ZEUS_UNSET;
  r.fval = 37.0; // reference Result<D> defaults (util.hpp:36-54)
}

// Launch shapes the autotuner may choose between for a class: the
// per-thread form (where offered) and every offered workgroup size of the
// per-workgroup form, unless the caller fixed the form or the size.
std::vector<TuneVal>
candidates(const ClassOps* ops, const zeus_options& opt)
{
  std::vector<TuneVal> v;
  const bool tile_ok = ops->tile_ok && opt.gradient == ZEUS_GRAD_ADJOINT;
  if (opt.form == ZEUS_FORM_TILE && tile_ok)
    return {{ZEUS_FORM_TILE, dev::kTileThreads}};
  const bool want_thread = opt.form == ZEUS_FORM_THREAD || opt.form == ZEUS_FORM_AUTO;
  const bool want_group =
    opt.form == ZEUS_FORM_GROUP || opt.form == ZEUS_FORM_AUTO || !ops->thread_ok;
  if (want_thread && ops->thread_ok)
    v.push_back({ZEUS_FORM_THREAD, 0});
  if (want_group && tile_ok && opt.group_threads <= 0)
    v.push_back({ZEUS_FORM_TILE, dev::kTileThreads});
  if ((want_group || opt.form == ZEUS_FORM_TEAM) && opt.group_threads <= 0)
    for (int k = 0; k < ops->n_team; ++k)
      v.push_back({ZEUS_FORM_TEAM, ops->team_ts[k]});
  if (opt.form == ZEUS_FORM_TEAM && opt.group_threads > 0)
    for (int k = 0; k < ops->n_team; ++k)
      if (ops->team_ts[k] == opt.group_threads)
        v.push_back({ZEUS_FORM_TEAM, ops->team_ts[k]});
  if (want_group || v.empty())
    for (int k = 0; k < ops->n_group; ++k)
      if (opt.group_threads <= 0 || ops->group_nt[k] == opt.group_threads)
        v.push_back({ZEUS_FORM_GROUP, ops->group_nt[k]});
  if (v.empty()) // requested size not offered: fall back to the largest
    v.push_back({ZEUS_FORM_GROUP, ops->group_nt[ops->n_group - 1]});
  return v;
}

int
size_bucket(int64_t n)
{
  int b = 0;
  while ((int64_t(1) << (b + 1)) <= n)
    ++b;
  return b;
}

struct ClassPlan {
  const ClassOps* ops = nullptr;
  std::vector<int> probs; // indices into the caller's problem array
  int64_t total = 0;
  int pso_blocks = 0;
  int max_pso_iter = -1;
  TuneVal cfg{};
  std::vector<TuneVal> cands; // > 1 entry: time them and keep the fastest
  TuneKey key{};
  int grid = 0;
  int64_t slab = 0; // doubles
  // device carve
  ProblemDesc* d_desc = nullptr;
  int32_t* d_block_problem = nullptr;
  double *X = nullptr, *V = nullptr, *pbx = nullptr, *pbv = nullptr, *gbx = nullptr;
  double* part_val = nullptr;  // PSO partials, [2][pso_blocks] (iteration parity)
  int64_t* part_idx = nullptr;
  double* part_x = nullptr;    // [2][pso_blocks][D]
  double* gstate = nullptr;    // per problem [2][1 + D]
  StartOut out{};
  double* aux = nullptr; // Gaussian Cs, one D x D block per problem
  double* h_slab = nullptr;
  int ctl0 = 0;          // first control block of the class
  bool traced = false;   // some problem records a trajectory
  bool prepared = false; // the objective was prepared by the call's first kernel
  bool metrics = false;  // per-start timing requested (ZEUS_FLAG_METRICS)
  bool xorwow = false;   // some problem runs PSO on XORWOW streams
  int keep_rng = 0;      // PSO phases: the caller keeps the XORWOW states
  uint32_t* xw = nullptr; // their (start, coordinate) states: 6 x total x D words
  // host
  std::vector<ProblemDesc> desc;
  std::vector<int32_t> block_problem;
};

// Host-written inputs of a class (uploaded in one copy with the others).
void
carve_upload(Carver& cv, ClassPlan& c)
{
  c.d_desc = cv.take<ProblemDesc>(c.probs.size());
  c.d_block_problem = cv.take<int32_t>(size_t(std::max(c.pso_blocks, 1)));
}

void
carve_class(Carver& cv, ClassPlan& c)
{
  const int D = c.ops->dim;
  const size_t n = size_t(c.total), np = c.probs.size(), nb = size_t(std::max(c.pso_blocks, 1));
  const bool pso = c.max_pso_iter >= 0;
  c.X = cv.take<double>(pso ? n * D : 1);
  c.V = cv.take<double>(pso ? n * D : 1);
  c.pbx = cv.take<double>(pso ? n * D : 1);
  c.pbv = cv.take<double>(pso ? n : 1);
  c.gbx = cv.take<double>(np * D);
  c.part_val = cv.take<double>(2 * nb);
  c.part_idx = cv.take<int64_t>(2 * nb);
  c.part_x = cv.take<double>(pso ? 2 * nb * D : 1);
  c.gstate = cv.take<double>(pso ? np * 2 * (D + 1) : 1);
  c.out.status = cv.take<int32_t>(n);
  c.out.iter = cv.take<int32_t>(n);
  c.out.ls_evals = cv.take<int32_t>(n);
  c.out.grads = cv.take<int32_t>(n);
  c.out.near = cv.take<int8_t>(n);
  c.out.fval = cv.take<double>(n);
  c.out.gnorm = cv.take<double>(n);
  c.out.x = cv.take<double>(n * D);
  c.aux = cv.take<double>(c.ops->objective == obj::kGaussian ? np * dev::gaussian_aux_doubles(D)
                                                              : 1);
  c.h_slab = cv.take<double>(size_t(std::max<int64_t>(c.slab, 1)));
  if (c.xorwow)
    c.xw = cv.take<uint32_t>(6 * n * D);
  if (c.metrics) {
    c.out.ad_ticks = cv.take<uint64_t>(n);
    c.out.bfgs_ticks = cv.take<uint64_t>(n);
    c.out.total_ticks = cv.take<uint64_t>(n);
    c.out.ad_calls = cv.take<int32_t>(n);
    c.out.bfgs_calls = cv.take<int32_t>(n);
  } else {
    c.out.ad_ticks = c.out.bfgs_ticks = c.out.total_ticks = nullptr;
    c.out.ad_calls = c.out.bfgs_calls = nullptr;
  }
}

PsoArgs
pso_args(const ClassPlan& c, Control* ctl)
{
  return PsoArgs{c.d_desc,   int32_t(c.probs.size()),
                 0,          c.pso_blocks,
                 ctl,        c.X,
                 c.V,        c.pbx,
                 c.pbv,      c.gbx,
                 c.part_val, c.part_idx,
                 c.part_x,   c.gstate,
                 c.d_block_problem, c.xw,
                 c.total * c.ops->dim, c.keep_rng};
}

// Objective preparation and PSO iterations t0 .. max_pso_iter of a class.
hipError_t
launch_class(ClassPlan& c, Control* ctl_all, int gradient, hipStream_t st, int t0 = 0)
{
  const int D = c.ops->dim;
  const int np = int(c.probs.size());
  Control* ctl = ctl_all + c.ctl0;
  hipError_t e = hipSuccess;
  if (c.ops->objective == obj::kGaussian && !c.prepared) {
    const int64_t na = dev::gaussian_aux_doubles(D);
    for (int k = 0; k < np; ++k) {
      double* cs = c.aux + k * na;
      dev::symmetrize_kernel<<<int((na + 255) / 256), 256, 0, st>>>(c.d_desc, k, cs,
                                                                    cs + int64_t(D) * D, D);
    }
    if ((e = hipGetLastError()) != hipSuccess)
      return e;
  }
  if (c.max_pso_iter >= 0 && c.pso_blocks > 0) {
    PsoArgs pa = pso_args(c, ctl);
    for (int it = t0; it <= c.max_pso_iter; ++it) {
      pa.iter = it;
      if ((e = c.ops->pso(pa, c.pso_blocks, st)) != hipSuccess)
        return e;
    }
  }
  return hipSuccess;
}

// Persistent grid of the group and tile forms: as many workgroups as can
// be resident (at most one per start / per tile).
int
group_grid(const ClassPlan& c, const TuneVal& v, int gradient)
{
  if (v.form == ZEUS_FORM_TILE) {
    const int occ = c.ops->tile_blocks_per_cu();
    const int g = int(std::min<int64_t>(c.pso_blocks, int64_t(g_ds.n_cu) * occ));
    return std::max(g, 1);
  }
  if (v.form == ZEUS_FORM_TEAM) {
    const int occ = c.ops->team_blocks_per_cu(gradient, v.threads);
    const int64_t teams = dev::kTeamThreads / v.threads;
    const int g = int(std::min<int64_t>((c.total + teams - 1) / teams, int64_t(g_ds.n_cu) * occ));
    return std::max(g, 1);
  }
  const int occ = c.ops->group_blocks_per_cu(gradient, v.threads);
  const int g = int(std::min<int64_t>(c.total, int64_t(g_ds.n_cu) * occ));
  return std::max(g, 1);
}

int64_t
slab_doubles(const ClassPlan& c, const TuneVal& v, int gradient)
{
  if (v.form == ZEUS_FORM_THREAD)
    return c.ops->thread_slab_per_start * c.total;
  if (v.form == ZEUS_FORM_TILE)
    return c.ops->tile_slot_doubles * dev::kTileStarts * group_grid(c, v, gradient);
  if (v.form == ZEUS_FORM_TEAM)
    return 0; // H in registers
  return c.ops->group_slab_per_block(v.threads) * group_grid(c, v, gradient);
}

hipError_t
launch_bfgs_cfg(ClassPlan& c, const TuneVal& v, Control* ctl_all, int gradient, hipStream_t st)
{
  ClassArgs a{c.d_desc, int32_t(c.probs.size()), c.ops->dim, c.total, ctl_all + c.ctl0,
              nullptr, c.out, c.h_slab, 0};
  if (c.total == 0)
    return hipSuccess;
  if (v.form == ZEUS_FORM_THREAD)
    return c.ops->bfgs_thread(a, gradient, st);
  if (v.form == ZEUS_FORM_TILE) {
    const dev::TileArgs ta{c.d_desc, c.d_block_problem, c.pso_blocks, ctl_all + c.ctl0, c.out,
                           c.h_slab};
    return c.ops->bfgs_tile(ta, group_grid(c, v, gradient), st);
  }
  if (v.form == ZEUS_FORM_TEAM)
    return c.ops->bfgs_team(a, gradient, v.threads, group_grid(c, v, gradient), st);
  a.h_slab_stride = c.ops->group_slab_per_block(v.threads);
  return c.ops->bfgs_group(a, gradient, v.threads, group_grid(c, v, gradient), st);
}

// BFGS phase of a class.  With several candidate shapes (first call for this
// key), each runs on the real inputs from the same PSO result and control
// state; the fastest is cached and its results are the ones kept.  Each
// candidate is timed twice and keeps its faster time, and the first one is
// run once more before it is timed: a call made while the GPU's clocks are
// still coming up (the first work of a process) would otherwise time its
// first candidate slow and can keep a shape several times slower for every
// later call.
hipError_t
launch_bfgs(ClassPlan& c, Control* ctl_all, int gradient, hipStream_t st)
{
  if (c.cands.size() <= 1)
    return launch_bfgs_cfg(c, c.cfg, ctl_all, gradient, st);
  const int np = int(c.probs.size());
  float best_ms = 0.f;
  size_t best = 0;
  hipError_t e = hipSuccess;
  for (size_t k = 0; k < c.cands.size(); ++k) {
    float ms = 0.f;
    for (int rep = k == 0 ? -1 : 0; rep < 2; ++rep) {
      dev::init_controls_kernel<<<(np + 255) / 256, 256, 0, st>>>(ctl_all + c.ctl0, np, nullptr,
                                                                  0);
      (void)hipEventRecord(g_ds.ev[4], st);
      if ((e = launch_bfgs_cfg(c, c.cands[k], ctl_all, gradient, st)) != hipSuccess)
        return e;
      (void)hipEventRecord(g_ds.ev[5], st);
      if ((e = hipEventSynchronize(g_ds.ev[5])) != hipSuccess)
        return e;
      float t = 0.f;
      (void)hipEventElapsedTime(&t, g_ds.ev[4], g_ds.ev[5]);
      if (rep == 0 || (rep > 0 && t < ms))
        ms = t;
    }
    if (k == 0 || ms < best_ms) {
      best_ms = ms;
      best = k;
    }
  }
  c.cfg = c.cands[best];
  g_ds.tuned[c.key] = c.cfg;
  if (best + 1 != c.cands.size()) {
    dev::init_controls_kernel<<<(np + 255) / 256, 256, 0, st>>>(ctl_all + c.ctl0, np, nullptr, 0);
    e = launch_bfgs_cfg(c, c.cfg, ctl_all, gradient, st);
  }
  return e;
}

} // namespace

extern "C" {

zeus_status_t
zeus_run_batch(const zeus_problem* probs, int32_t n_probs, const zeus_options* opt_in,
               zeus_result* results, double* best_x, const zeus_start_outputs* so,
               void* stream)
{
  const auto t_start = std::chrono::steady_clock::now();
  if (n_probs < 0 || (n_probs > 0 && (!probs || !results)))
    return ZEUS_E_ARG;
  zeus_options opt{};
  if (opt_in)
    opt = *opt_in;
  if (opt.gradient != ZEUS_GRAD_ADJOINT && opt.gradient != ZEUS_GRAD_DUAL)
    return ZEUS_E_ARG;
  int64_t total_all = 0, dims_all = 0;
  for (int i = 0; i < n_probs; ++i) {
    const zeus_status_t v = validate(probs[i]);
    if (v != ZEUS_OK)
      return v;
    total_all += probs[i].n_starts;
    dims_all += probs[i].dim;
  }
  for (int i = 0; i < n_probs; ++i)
    default_result(results[i]);
  if (best_x)
    std::fill(best_x, best_x + dims_all, 0.0);
  if (total_all == 0)
    return ZEUS_OK; // N == 0 -> Result{} (zeus.cuh:58-60)

  std::lock_guard<std::mutex> lock(g_mu);
  zeus_status_t rc = ensure_device();
  if (rc != ZEUS_OK)
    return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g_ds.stream;

  // ---- plan: classes in order of first appearance ------------------------
  std::vector<ClassPlan> classes;
  std::vector<int> cls_of(n_probs, -1), local_of(n_probs, -1);
  for (int i = 0; i < n_probs; ++i) {
    const ClassOps* ops = rt::find_class(probs[i].objective, probs[i].dim);
    int c = -1;
    for (size_t k = 0; k < classes.size(); ++k)
      if (classes[k].ops == ops)
        c = int(k);
    if (c < 0) {
      classes.emplace_back();
      classes.back().ops = ops;
      c = int(classes.size()) - 1;
    }
    cls_of[i] = c;
    local_of[i] = int(classes[c].probs.size());
    classes[c].probs.push_back(i);
  }
  int ctl_total = 0;
  for (auto& c : classes) {
    c.ctl0 = ctl_total;
    ctl_total += int(c.probs.size());
    int64_t s0 = 0;
    int b0 = 0;
    for (int pi : c.probs) {
      const zeus_problem& p = probs[pi];
      ProblemDesc d{};
      d.objective = p.objective;
      d.dim = p.dim;
      d.lower = p.lower;
      d.upper = p.upper;
      d.start0 = s0;
      d.n_starts = p.n_starts;
      d.required = p.required;
      d.max_iter = p.max_iter;
      d.pso_iter = (p.x0 != nullptr) ? -1 : p.pso_iter;
      d.block0 = b0;
      d.min_kind = min_kind_of(p.name);
      d.seed = p.seed;
      d.stream = p.stream;
      d.tol = p.tol;
      d.params = p.params;
      d.n_params = p.n_params;
      d.x0 = p.x0;
      d.prng = p.prng;
      c.xorwow = c.xorwow || (d.prng == ZEUS_PRNG_XORWOW && d.pso_iter > 0);
      d.traj = traj_of(p);
      c.traced = c.traced || d.traj.fval != nullptr;
      c.desc.push_back(d);
      const int nb = int((p.n_starts + c.ops->pso_ppb - 1) / c.ops->pso_ppb);
      for (int b = 0; b < nb; ++b)
        c.block_problem.push_back(int32_t(c.desc.size() - 1));
      b0 += nb;
      s0 += p.n_starts;
      c.max_pso_iter = std::max(c.max_pso_iter, d.pso_iter);
    }
    c.total = s0;
    c.pso_blocks = b0;
    c.metrics = (opt.flags & ZEUS_FLAG_METRICS) != 0;
    c.key = TuneKey{c.ops->objective, c.ops->dim, opt.gradient, size_bucket(c.total)};
    c.cands = candidates(c.ops, opt);
    if (c.traced) {
      // recording runs one shape, not the tile or team form, and is not
      // timed against others (each trial would rewrite the trajectory)
      if (opt.form == ZEUS_FORM_TILE || opt.form == ZEUS_FORM_TEAM)
        return ZEUS_E_ARG;
      std::vector<TuneVal> keep;
      for (const auto& v : c.cands)
        if (v.form != ZEUS_FORM_TILE && v.form != ZEUS_FORM_TEAM)
          keep.push_back(v);
      if (keep.empty())
        keep.push_back({ZEUS_FORM_GROUP, c.ops->group_nt[c.ops->n_group - 1]});
      c.cands = keep;
    }
    if (c.cands.size() > 1) {
      auto it = g_ds.tuned.find(c.key);
      if (it != g_ds.tuned.end())
        for (const auto& v : c.cands)
          if (v.form == it->second.form && v.threads == it->second.threads)
            c.cands = {v};
    }
    if (c.traced && c.cands.size() > 1)
      c.cands.resize(1);
    c.cfg = c.cands.front();
    c.slab = 0;
    for (const auto& v : c.cands)
      c.slab = std::max(c.slab, slab_doubles(c, v, opt.gradient));
  }

  // ---- device memory ------------------------------------------------------
  // Arena layout: [inputs written by the host | results read back | the
  // rest].  The first two regions are mirrored in pinned host memory, so a
  // call makes one upload copy and one download copy.
  size_t up_end = 0, dn_begin = 0, dn_end = 0;
  auto layout = [&](Carver& cv, Control** ctl, double** bx, int64_t** bxo) {
    for (auto& c : classes)
      carve_upload(cv, c);
    *bxo = cv.take<int64_t>(size_t(ctl_total));
    up_end = cv.off;
    *ctl = cv.take<Control>(size_t(ctl_total));
    dn_begin = reinterpret_cast<char*>(*ctl) - cv.base;
    *bx = cv.take<double>(size_t(dims_all));
    dn_end = cv.off;
    for (auto& c : classes)
      carve_class(cv, c);
  };
  Control* d_ctl;
  double* d_bx;
  int64_t* d_bxo;
  {
    Carver m{nullptr};
    layout(m, &d_ctl, &d_bx, &d_bxo);
    if ((rc = arena_reserve(m.off)) != ZEUS_OK || (rc = pinned_reserve(dn_end)) != ZEUS_OK) {
      for (int i = 0; i < n_probs; ++i)
        results[i].status = ZEUS_MALLOC;
      return rc;
    }
    Carver cv{static_cast<char*>(g_ds.arena)};
    layout(cv, &d_ctl, &d_bx, &d_bxo);
  }
  // host mirror of the first two regions, at the same offsets
  char* const pin = static_cast<char*>(g_ds.pinned);
  auto mirror = [&](const void* dev_ptr) {
    return pin + (static_cast<const char*>(dev_ptr) - static_cast<char*>(g_ds.arena));
  };
  // problems that run PSO start their BFGS from pBestX
  for (auto& c : classes)
    for (auto& d : c.desc) {
      if (d.x0 == nullptr && d.pso_iter >= 0)
        d.x0 = c.pbx + d.start0 * d.dim;
      if (c.ops->objective == obj::kGaussian) {
        d.aux = c.aux + (&d - c.desc.data()) * dev::gaussian_aux_doubles(d.dim);
        d.aux_frag = d.aux + int64_t(d.dim) * d.dim;
      }
    }
  // best-x offsets per control block (class order)
  std::vector<int64_t> bxo(ctl_total);
  std::vector<int64_t> bx_user(n_probs);
  {
    int64_t o = 0;
    for (int i = 0; i < n_probs; ++i) {
      bx_user[i] = o;
      o += probs[i].dim;
    }
    for (int i = 0; i < n_probs; ++i)
      bxo[classes[cls_of[i]].ctl0 + local_of[i]] = bx_user[i];
  }

  hipError_t e = hipSuccess;
  auto fail = [&](hipError_t err) {
    (void)hipStreamSynchronize(st);
    const zeus_status_t s = to_status(err);
    for (int i = 0; i < n_probs; ++i)
      results[i].status = (s == ZEUS_E_ALLOC) ? ZEUS_MALLOC : ZEUS_KERNEL;
    return s;
  };
  for (auto& c : classes) {
    std::memcpy(mirror(c.d_desc), c.desc.data(), sizeof(ProblemDesc) * c.desc.size());
    std::memcpy(mirror(c.d_block_problem), c.block_problem.data(),
                sizeof(int32_t) * c.block_problem.size());
  }
  std::memcpy(mirror(d_bxo), bxo.data(), sizeof(int64_t) * bxo.size());
  if ((e = hipMemcpyAsync(g_ds.arena, pin, up_end, hipMemcpyHostToDevice, st)) != hipSuccess)
    return fail(e);

  // ---- launch sequence ----------------------------------------------------
  // PSO -> BFGS -> summary, stream-ordered, no host synchronisation.  With
  // use_graph (and every class's launch shape settled), the sequence is
  // captured once per batch shape and replayed as a hipGraph; the problem
  // descriptors are copied before it, outside the graph.  Fills are kernels
  // (no memset nodes), so replays start from cleared control blocks.
  // Classes run side by side once every class's shape is settled (no
  // autotuner timing in this call): class k on side stream slot[k], the
  // heaviest classes (starts x dimension) spread first.  A class of surrendering
  // starts can keep a few wavefronts busy for a long time; the other classes
  // fill the rest of the GPU meanwhile.
  bool concurrent = classes.size() > 1;
  for (const auto& c : classes)
    concurrent = concurrent && c.cands.size() == 1;
  std::vector<int> slot(classes.size(), 0);
  if (concurrent) {
    std::vector<size_t> order(classes.size());
    for (size_t k = 0; k < order.size(); ++k)
      order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
      return classes[x].total * classes[x].ops->dim > classes[y].total * classes[y].ops->dim;
    });
    for (size_t r = 0; r < order.size(); ++r)
      slot[order[r]] = int(r % DeviceState::kSide);
  }
  // per class: PSO start, PSO end, BFGS end (classes are timed one by one)
  while (g_ds.cls_ev.size() < 3 * classes.size()) {
    hipEvent_t ev;
    if (hipEventCreateWithFlags(&ev, kTimingEventFlags) != hipSuccess)
      return fail(hipErrorOutOfMemory);
    g_ds.cls_ev.push_back(ev);
  }
  // One class one after the other: its own events also time the whole call
  // (each event record between kernels costs the GPU a few microseconds).
  bool any_traj = false;
  for (const auto& c : classes)
    for (const auto& d : c.desc)
      any_traj = any_traj || d.traj.fval != nullptr;
  const bool one = classes.size() == 1;
  const hipEvent_t ev_begin = one ? g_ds.cls_ev[0] : g_ds.ev[0];
  const hipEvent_t ev_pso = one && !any_traj ? g_ds.cls_ev[1] : g_ds.ev[1];
  const hipEvent_t ev_end = one ? g_ds.cls_ev[2] : g_ds.ev[2];
  auto sequence = [&](bool timed) -> hipError_t {
    hipError_t err = hipSuccess;
    // the controls and the first class's gBest coordinates in one launch
    // (grid-stride over the coordinates), the other classes' after it
    for (size_t i = 0; i < classes.size(); ++i) {
      ClassPlan& c = classes[i];
      const int64_t ng = int64_t(c.probs.size()) * c.ops->dim;
      // a first class of one Gaussian problem: its preparation (Cs and the
      // fragments) joins this first launch
      c.prepared = i == 0 && c.ops->objective == obj::kGaussian && c.probs.size() == 1;
      if (c.prepared) {
        const int D = c.ops->dim;
        const int64_t na = dev::gaussian_aux_doubles(D);
        const int64_t n = std::max<int64_t>(ctl_total, na);
        dev::init_symmetrize_kernel<<<int((n + 255) / 256), 256, 0, st>>>(
          d_ctl, ctl_total, c.gbx, ng, c.d_desc, 0, c.aux, c.aux + int64_t(D) * D, D);
      } else if (i == 0)
        dev::init_controls_kernel<<<(ctl_total + 255) / 256, 256, 0, st>>>(d_ctl, ctl_total,
                                                                           c.gbx, ng);
      else
        dev::init_controls_kernel<<<int((ng + 255) / 256), 256, 0, st>>>(nullptr, 0, c.gbx, ng);
    }
    if (classes.empty())
      dev::init_controls_kernel<<<(ctl_total + 255) / 256, 256, 0, st>>>(d_ctl, ctl_total,
                                                                         nullptr, 0);
    if ((err = hipGetLastError()) != hipSuccess)
      return err;
    if (concurrent) {
      if (timed)
        (void)hipEventRecord(g_ds.ev[0], st);
      if ((err = hipEventRecord(g_ds.fork, st)) != hipSuccess)
        return err;
      bool used[DeviceState::kSide] = {};
      for (size_t k = 0; k < classes.size(); ++k) {
        auto& c = classes[k];
        hipStream_t cs = g_ds.side[slot[k]];
        if (!used[slot[k]] && (err = hipStreamWaitEvent(cs, g_ds.fork, 0)) != hipSuccess)
          return err;
        used[slot[k]] = true;
        if (timed)
          (void)hipEventRecord(g_ds.cls_ev[3 * k], cs);
        if ((err = launch_class(c, d_ctl, opt.gradient, cs)) != hipSuccess)
          return err;
        for (const auto& d : c.desc)
          if (d.traj.fval) {
            const int64_t n = int64_t(d.traj.steps) * d.traj.starts * d.dim;
            const int g = int(std::min<int64_t>((n + 255) / 256, 4096));
            dev::traj_fill_kernel<<<g, 256, 0, cs>>>(d.traj, d.dim);
          }
        if (timed)
          (void)hipEventRecord(g_ds.cls_ev[3 * k + 1], cs);
        if ((err = launch_bfgs(c, d_ctl, opt.gradient, cs)) != hipSuccess)
          return err;
        g_ds.last[{c.ops->objective, c.ops->dim, opt.gradient}] = c.cfg;
        if (timed)
          (void)hipEventRecord(g_ds.cls_ev[3 * k + 2], cs);
      }
      for (int j = 0; j < DeviceState::kSide; ++j)
        if (used[j]) {
          if ((err = hipEventRecord(g_ds.join[j], g_ds.side[j])) != hipSuccess)
            return err;
          if ((err = hipStreamWaitEvent(st, g_ds.join[j], 0)) != hipSuccess)
            return err;
        }
      if (timed)
        (void)hipEventRecord(g_ds.ev[2], st);
      for (auto& c : classes) {
        dev::summary_kernel<<<int(c.probs.size()), dev::kSummaryThreads, 0, st>>>(
          c.d_desc, d_ctl + c.ctl0, c.out, d_bx, d_bxo + c.ctl0);
      }
      return hipGetLastError();
    }
    if (timed && ev_begin == g_ds.ev[0])
      (void)hipEventRecord(g_ds.ev[0], st);
    for (size_t k = 0; k < classes.size(); ++k) {
      if (timed)
        (void)hipEventRecord(g_ds.cls_ev[3 * k], st);
      if ((err = launch_class(classes[k], d_ctl, opt.gradient, st)) != hipSuccess)
        return err;
      if (timed)
        (void)hipEventRecord(g_ds.cls_ev[3 * k + 1], st);
    }
    // trajectory buffers: NaN / -1 / 0 before the BFGS writes its steps
    for (const auto& c : classes)
      for (const auto& d : c.desc)
        if (d.traj.fval) {
          const int64_t n = int64_t(d.traj.steps) * d.traj.starts * d.dim;
          const int g = int(std::min<int64_t>((n + 255) / 256, 4096));
          dev::traj_fill_kernel<<<g, 256, 0, st>>>(d.traj, d.dim);
        }
    if ((err = hipGetLastError()) != hipSuccess)
      return err;
    if (timed && ev_pso == g_ds.ev[1])
      (void)hipEventRecord(g_ds.ev[1], st);
    for (size_t k = 0; k < classes.size(); ++k) {
      auto& c = classes[k];
      if ((err = launch_bfgs(c, d_ctl, opt.gradient, st)) != hipSuccess)
        return err;
      if (timed)
        (void)hipEventRecord(g_ds.cls_ev[3 * k + 2], st);
      g_ds.last[{c.ops->objective, c.ops->dim, opt.gradient}] = c.cfg;
    }
    if (timed && ev_end == g_ds.ev[2])
      (void)hipEventRecord(g_ds.ev[2], st);
    for (auto& c : classes) {
      dev::summary_kernel<<<int(c.probs.size()), dev::kSummaryThreads, 0, st>>>(
        c.d_desc, d_ctl + c.ctl0, c.out, d_bx, d_bxo + c.ctl0);
    }
    return hipGetLastError();
  };
  bool settled = true;
  for (const auto& c : classes)
    settled = settled && c.cands.size() == 1 && !c.traced;
  bool graphed = false;
  if (opt.use_graph && settled) {
    std::vector<uint64_t> key = {uint64_t(reinterpret_cast<uintptr_t>(g_ds.arena)),
                                 uint64_t(g_ds.arena_bytes), uint64_t(ctl_total),
                                 uint64_t(dims_all), uint64_t(opt.gradient),
                                 uint64_t(reinterpret_cast<uintptr_t>(st))};
    for (const auto& c : classes) {
      key.push_back(uint64_t(reinterpret_cast<uintptr_t>(c.ops)));
      key.push_back(uint64_t(c.total));
      key.push_back(uint64_t(c.pso_blocks));
      key.push_back(uint64_t(int64_t(c.max_pso_iter)));
      key.push_back(uint64_t(c.probs.size()));
      key.push_back(uint64_t(c.cfg.form) << 32 | uint32_t(c.cfg.threads));
      key.push_back(uint64_t(c.xorwow));
      key.push_back(uint64_t(c.metrics));
    }
    key.push_back(uint64_t(concurrent));
    hipGraphExec_t exec = nullptr;
    for (const auto& ge : g_ds.graphs)
      if (ge.key == key)
        exec = ge.exec;
    if (!exec) {
      hipGraph_t graph = nullptr;
      if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return fail(e);
      const hipError_t le = sequence(false);
      e = hipStreamEndCapture(st, &graph);
      if (le != hipSuccess || e != hipSuccess) {
        if (graph)
          (void)hipGraphDestroy(graph);
        return fail(le != hipSuccess ? le : e);
      }
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess)
        return fail(e);
      g_ds.graphs.push_back({key, exec});
    }
    (void)hipEventRecord(g_ds.ev[0], st);
    if ((e = hipGraphLaunch(exec, st)) != hipSuccess)
      return fail(e);
    (void)hipEventRecord(g_ds.ev[2], st);
    graphed = true;
  } else if ((e = sequence(true)) != hipSuccess) {
    return fail(e);
  }

  if ((e = hipMemcpyAsync(pin + dn_begin, static_cast<char*>(g_ds.arena) + dn_begin,
                          dn_end - dn_begin, hipMemcpyDeviceToHost, st)) != hipSuccess)
    return fail(e);
  if ((e = hipStreamSynchronize(st)) != hipSuccess)
    return fail(e);
  const Control* h_ctl = reinterpret_cast<const Control*>(mirror(d_ctl));
  const double* h_bx = reinterpret_cast<const double*>(mirror(d_bx));

  // device times per class: its PSO phase (objective preparation included)
  // and its BFGS phase; a replayed graph reports its whole device time
  std::vector<float> cls_pso(classes.size(), 0.f), cls_opt(classes.size(), 0.f);
  float ms_pso = 0.f, ms_opt = 0.f;
  if (graphed) { // one replay: its device time (PSO, BFGS and summary) in ms_opt
    (void)hipEventElapsedTime(&ms_opt, g_ds.ev[0], g_ds.ev[2]);
    std::fill(cls_opt.begin(), cls_opt.end(), ms_opt);
  } else {
    for (size_t k = 0; k < classes.size(); ++k) {
      // side by side, a class's BFGS follows its own PSO; one after the other,
      // it follows the previous class's BFGS (or the end of every PSO phase)
      const hipEvent_t b0 = concurrent ? g_ds.cls_ev[3 * k + 1]
                            : k == 0   ? ev_pso
                                       : g_ds.cls_ev[3 * (k - 1) + 2];
      (void)hipEventElapsedTime(&cls_pso[k], g_ds.cls_ev[3 * k], g_ds.cls_ev[3 * k + 1]);
      (void)hipEventElapsedTime(&cls_opt[k], b0, g_ds.cls_ev[3 * k + 2]);
    }
    if (concurrent) { // the longest PSO and BFGS phases
      for (size_t k = 0; k < classes.size(); ++k) {
        ms_pso = std::max(ms_pso, cls_pso[k]);
        ms_opt = std::max(ms_opt, cls_opt[k]);
      }
    } else {
      (void)hipEventElapsedTime(&ms_pso, ev_begin, ev_pso);
      (void)hipEventElapsedTime(&ms_opt, ev_pso, ev_end);
    }
  }

  // per-start outputs, scattered back into problem order
  if (so) {
    int64_t off_s = 0, off_x = 0;
    std::vector<int64_t> offs(n_probs), offx(n_probs);
    for (int i = 0; i < n_probs; ++i) {
      offs[i] = off_s;
      offx[i] = off_x;
      off_s += probs[i].n_starts;
      off_x += probs[i].n_starts * probs[i].dim;
    }
    for (auto& c : classes) {
      const int D = c.ops->dim;
      for (size_t k = 0; k < c.probs.size(); ++k) {
        const int pi = c.probs[k];
        const int64_t s0 = c.desc[k].start0, n = c.desc[k].n_starts;
        if (n == 0)
          continue;
        auto cp = [&](void* dst, const void* src, size_t bytes) {
          if (dst && e == hipSuccess)
            e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
        };
        cp(so->status ? so->status + offs[pi] : nullptr, c.out.status + s0, 4 * n);
        cp(so->iter ? so->iter + offs[pi] : nullptr, c.out.iter + s0, 4 * n);
        cp(so->ls_evals ? so->ls_evals + offs[pi] : nullptr, c.out.ls_evals + s0, 4 * n);
        cp(so->grads ? so->grads + offs[pi] : nullptr, c.out.grads + s0, 4 * n);
        cp(so->fval ? so->fval + offs[pi] : nullptr, c.out.fval + s0, 8 * n);
        cp(so->gnorm ? so->gnorm + offs[pi] : nullptr, c.out.gnorm + s0, 8 * n);
        cp(so->x ? so->x + offx[pi] : nullptr, c.out.x + s0 * D, 8 * n * D);
      }
    }
    if (e != hipSuccess)
      return fail(e);
  }

  // per-start timing of the per-thread form -> metrics per problem
  std::vector<zeus_metrics> m_ad(n_probs), m_bfgs(n_probs);
  for (auto& c : classes) {
    if (!c.metrics || c.cfg.form != ZEUS_FORM_THREAD)
      continue;
    for (size_t k = 0; k < c.probs.size(); ++k) {
      const int64_t s0 = c.desc[k].start0, n = c.desc[k].n_starts;
      if (n == 0)
        continue;
      std::vector<uint64_t> at(n), bt(n), tt(n);
      std::vector<int32_t> ac(n), bc(n);
      auto cp = [&](void* dst, const void* src, size_t bytes) {
        if (e == hipSuccess)
          e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
      };
      cp(at.data(), c.out.ad_ticks + s0, 8 * n);
      cp(bt.data(), c.out.bfgs_ticks + s0, 8 * n);
      cp(tt.data(), c.out.total_ticks + s0, 8 * n);
      cp(ac.data(), c.out.ad_calls + s0, 4 * n);
      cp(bc.data(), c.out.bfgs_calls + s0, 4 * n);
      if (e != hipSuccess)
        return fail(e);
      const int pi = c.probs[k];
      (void)zeus_metrics_from_ticks(at.data(), ac.data(), tt.data(), n, 256, ms_opt,
                                    g_ds.tick_hz, &m_ad[pi]);
      (void)zeus_metrics_from_ticks(bt.data(), bc.data(), tt.data(), n, 256, ms_opt,
                                    g_ds.tick_hz, &m_bfgs[pi]);
    }
  }

  // ---- results ------------------------------------------------------------
  // best Result per problem: fetch its status/iter/fval/gnorm
  const double ms_total =
    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
      .count();
  for (int i = 0; i < n_probs; ++i) {
    const ClassPlan& c = classes[cls_of[i]];
    const Control& k = h_ctl[c.ctl0 + local_of[i]];
    zeus_result& r = results[i];
    if (probs[i].n_starts == 0)
      continue;
    r.idx = k.best_idx;
    r.status = k.best_status;
    r.iter = k.best_iter;
    r.fval = k.best_fval;
    r.gradient_norm = k.best_gnorm;
    r.claimed = k.claimed;
    r.actual = k.actual;
    r.stopped = k.stopped;
    r.surrendered = k.surrendered;
    r.nonfinite = k.nonfinite;
    r.sum_iter = k.sum_iter;
    r.sum_ls_evals = k.sum_ls_evals;
    r.sum_grads = k.sum_grads;
    r.ms_pso = cls_pso[cls_of[i]];
    r.ms_opt = cls_opt[cls_of[i]];
    r.ms_total = ms_total;
    r.ad = m_ad[i];
    r.bfgs = m_bfgs[i];
    if (best_x)
      std::memcpy(best_x + bx_user[i], h_bx + bx_user[i], sizeof(double) * probs[i].dim);
  }
  return ZEUS_OK;
}

} // extern "C"

// ===========================================================================
// Phase entry points, memory, random numbers, hooks.
// ===========================================================================
namespace {

// Draws 0..n-1 of one stream, in order, by one lane.
__global__ void
stream_fill_kernel(double* out, int64_t n, int prng, uint32_t seed, uint32_t stream,
                   uint64_t start, uint32_t coord)
{
  rng::Xorwow st = rng::xorwow_seed(seed, stream, start, coord);
  for (int64_t k = 0; k < n; ++k)
    out[k] = prng == rng::kXorwow  ? rng::xorwow_uniform(st)
             : prng == rng::kSobol ? rng::sobol_uniform(start, uint64_t(k))
                                   : rng::uniform(seed, stream, start, uint64_t(k));
}

__global__ void
argmin_kernel(const double* v, int64_t n, int64_t* out)
{
  __shared__ double sv[4];
  __shared__ int64_t si[4];
  double bv = __builtin_nan("");
  int64_t bi = int64_t(1) << 62;
  for (int64_t k = threadIdx.x; k < n; k += 256)
    if (dev::better(v[k], k, bv, bi)) {
      bv = v[k];
      bi = k;
    }
  dev::block_argmin<256>(bv, bi, sv, si);
  if (threadIdx.x == 0)
    *out = n > 0 ? (bi < n ? bi : 0) : -1;
}

__global__ void
philox_fill_kernel(double* out, int64_t n, uint32_t seed, uint32_t stream, uint64_t start,
                   uint64_t draw0)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    out[i] = rng::uniform(seed, stream, start, draw0 + uint64_t(i));
}

// Rank-2 BFGS update of a batch of n x n matrices H[b] (row-major) with
// s[b], y[b] (reference utils.cuh:245-287 in its O(n^2) form: skipped when
// |s.y| < 1e-14; H' = H + s a^T + a s^T, a = (beta/2) s - rho H y, beta =
// rho^2 y.Hy + rho), one workgroup per matrix, sums in the oracle's order.
__global__ void
bfgs_update_kernel(double* H, const double* s, const double* y, int n)
{
  extern __shared__ double sm[];
  __shared__ double red[2];
  const int64_t b = blockIdx.x;
  double* h = H + b * int64_t(n) * n;
  const double* sb = s + b * n;
  const double* yb = y + b * n;
  double* u = sm;     // H y
  double* a = sm + n; // the update vector
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j)
      acc = ::fma(h[int64_t(i) * n + j], yb[j], acc);
    u[i] = acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) { // left-to-right sums, as the oracle's dot products
    double sty = 0.0, yu = 0.0;
    for (int j = 0; j < n; ++j)
      sty = ::fma(sb[j], yb[j], sty);
    for (int j = 0; j < n; ++j)
      yu = ::fma(yb[j], u[j], yu);
    red[0] = sty;
    red[1] = yu;
  }
  __syncthreads();
  const double sty = red[0];
  if (::fabs(sty) < 1e-14)
    return; // uniform over the workgroup
  const double rho = 1.0 / sty;
  const double hb = 0.5 * ::fma(rho * rho, red[1], rho);
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    a[i] = ::fma(hb, sb[i], -(rho * u[i]));
  __syncthreads();
  for (int64_t e = threadIdx.x; e < int64_t(n) * n; e += blockDim.x) {
    const int i = int(e / n), j = int(e % n);
    h[e] = h[e] + (sb[i] * a[j] + a[i] * sb[j]);
  }
}

// PSO iterations t0 .. t1 of one problem on the given buffers (any of X, V,
// xw may be null: then the arena holds them).  With `seed`, iteration t0
// (>= 1) starts from the caller's gBest (gbx, gbv) instead of the partials
// of iteration t0 - 1.  Ends with the final fold: gbx and *gbv hold the
// gBest after iteration t1.
zeus_status_t
pso_phase(const zeus_problem* prob, int t0, int t1, double* X, double* V, double* pbx,
          double* pbv, double* gbx, double* gbv, uint32_t* xw, bool seed, hipStream_t st)
{
  ClassPlan c;
  c.ops = rt::find_class(prob->objective, prob->dim);
  c.probs.push_back(0);
  ProblemDesc d{};
  d.objective = prob->objective;
  d.dim = prob->dim;
  d.lower = prob->lower;
  d.upper = prob->upper;
  d.n_starts = prob->n_starts;
  d.pso_iter = t1;
  d.seed = prob->seed;
  d.stream = prob->stream;
  d.prng = prob->prng;
  d.params = prob->params;
  d.n_params = prob->n_params;
  c.desc.push_back(d);
  c.xorwow = d.prng == ZEUS_PRNG_XORWOW && t1 > 0 && xw == nullptr;
  c.pso_blocks = int((prob->n_starts + c.ops->pso_ppb - 1) / c.ops->pso_ppb);
  c.block_problem.assign(c.pso_blocks, 0);
  c.total = prob->n_starts;
  c.max_pso_iter = t1;
  Control* d_ctl;
  auto layout = [&](Carver& cv) {
    carve_upload(cv, c);
    carve_class(cv, c);
    d_ctl = cv.take<Control>(1);
  };
  zeus_status_t rc;
  {
    Carver m{nullptr};
    layout(m);
    if ((rc = arena_reserve(m.off)) != ZEUS_OK)
      return rc;
    Carver cv{static_cast<char*>(g_ds.arena)};
    layout(cv);
  }
  if (X)
    c.X = X;
  if (V)
    c.V = V;
  if (xw) {
    c.xw = xw;
    c.keep_rng = 1;
  }
  c.pbx = pbx;
  c.pbv = pbv;
  if (gbx)
    c.gbx = gbx;
  if (c.ops->objective == obj::kGaussian) {
    c.desc[0].aux = c.aux;
    c.desc[0].aux_frag = c.aux + int64_t(prob->dim) * prob->dim;
  }
  hipError_t e = hipMemcpyAsync(c.d_desc, c.desc.data(), sizeof(ProblemDesc),
                                hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c.d_block_problem, c.block_problem.data(), 4 * c.block_problem.size(),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    dev::init_controls_kernel<<<1, 256, 0, st>>>(d_ctl, 1, seed ? nullptr : c.gbx,
                                                 seed ? 0 : prob->dim);
    e = hipGetLastError();
  }
  if (e == hipSuccess && seed) { // the fold of iteration t0 keeps the caller's gBest
    dev::pso_seed_kernel<<<1, 256, 0, st>>>(pso_args(c, d_ctl), t0 - 1, gbv, prob->dim);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = launch_class(c, d_ctl, 0, st, t0);
  if (e == hipSuccess)
    e = c.ops->pso_final(pso_args(c, d_ctl), 1, st);
  if (e == hipSuccess && gbv)
    e = hipMemcpyAsync(gbv, &d_ctl->gbest_val, sizeof(double), hipMemcpyDeviceToDevice, st);
  const hipError_t e2 = hipStreamSynchronize(st);
  return to_status(e != hipSuccess ? e : e2);
}

// ---- fixed device scratch of the allocation-free PSO entry point ----------
// One array per GPU the library runs on, held until the process ends.  Its
// name is a host-side stand-in: host code uses only the address that
// hipGetSymbolAddress returns for the current device (scratch_base).
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch array (resolved once per device).
zeus_status_t
scratch_base(void** out)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return ZEUS_E_NODEVICE;
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return ZEUS_E_LAUNCH;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ZEUS_OK;
}

// Writes a problem's descriptor and its all-zero workgroup -> problem map
// into device memory (no host staging copy).
__global__ void
noalloc_setup_kernel(ProblemDesc d, ProblemDesc* out, int32_t* block_problem, int n_blocks)
{
  const int i = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
  if (i == 0)
    *out = d;
  if (i < n_blocks)
    block_problem[i] = 0;
}

// zeus_pso_run on the scratch array: the same kernels in the same order on
// the same inputs, with every buffer beyond the caller's outputs carved from
// `base` (kScratchBytes) instead of the library's arena.
zeus_status_t
pso_run_noalloc(const zeus_problem* prob, double* pbx, double* pbv, double* gbx_out,
                double* gbv, void* base, hipStream_t st)
{
  const ClassOps* ops = rt::find_class(prob->objective, prob->dim);
  const int D = prob->dim;
  const int64_t n = prob->n_starts;
  const int nb = int((n + ops->pso_ppb - 1) / ops->pso_ppb);
  const int T = prob->pso_iter;
  const bool gauss = ops->objective == obj::kGaussian;
  const bool xorwow = prob->prng == ZEUS_PRNG_XORWOW && T > 0;

  Carver cv{static_cast<char*>(base)};
  ProblemDesc* d_desc = cv.take<ProblemDesc>(1);
  int32_t* d_bp = cv.take<int32_t>(size_t(nb));
  double* X = cv.take<double>(size_t(n) * D);
  double* V = cv.take<double>(size_t(n) * D);
  double* gbx = gbx_out ? gbx_out : cv.take<double>(size_t(D));
  double* part_val = cv.take<double>(2 * size_t(nb));
  int64_t* part_idx = cv.take<int64_t>(2 * size_t(nb));
  double* part_x = cv.take<double>(2 * size_t(nb) * D);
  double* gstate = cv.take<double>(2 * size_t(D + 1));
  double* aux = gauss ? cv.take<double>(size_t(dev::gaussian_aux_doubles(D))) : nullptr;
  uint32_t* xw = xorwow ? cv.take<uint32_t>(6 * size_t(n) * D) : nullptr;
  Control* ctl = cv.take<Control>(1);
  if (cv.off > kScratchBytes)
    return ZEUS_E_SCRATCH;

  ProblemDesc d{};
  d.objective = prob->objective;
  d.dim = D;
  d.lower = prob->lower;
  d.upper = prob->upper;
  d.n_starts = n;
  d.pso_iter = T;
  d.seed = prob->seed;
  d.stream = prob->stream;
  d.prng = prob->prng;
  d.params = prob->params;
  d.n_params = prob->n_params;
  if (gauss) {
    d.aux = aux;
    d.aux_frag = aux + int64_t(D) * D;
  }
  noalloc_setup_kernel<<<(std::max(nb, 1) + 255) / 256, 256, 0, st>>>(d, d_desc, d_bp, nb);
  dev::init_controls_kernel<<<1, 256, 0, st>>>(ctl, 1, gbx, D);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && gauss) {
    const int64_t na = dev::gaussian_aux_doubles(D);
    dev::symmetrize_kernel<<<int((na + 255) / 256), 256, 0, st>>>(d_desc, 0, aux,
                                                                  aux + int64_t(D) * D, D);
    e = hipGetLastError();
  }
  PsoArgs pa{d_desc, 1,    0,        nb,       ctl,    X,    V,         pbx, pbv,
             gbx,    part_val, part_idx, part_x, gstate, d_bp, xw, n * D, 0};
  for (int it = 0; e == hipSuccess && it <= T; ++it) {
    pa.iter = it;
    e = ops->pso(pa, nb, st);
  }
  if (e == hipSuccess)
    e = ops->pso_final(pa, 1, st);
  if (e == hipSuccess && gbv)
    e = hipMemcpyAsync(gbv, &ctl->gbest_val, sizeof(double), hipMemcpyDeviceToDevice, st);
  const hipError_t e2 = hipStreamSynchronize(st);
  return to_status(e != hipSuccess ? e : e2);
}

} // namespace

extern "C" {

zeus_status_t
zeus_pso_run_noalloc(const zeus_problem* prob, double* pbest_x, double* pbest_val,
                     double* gbest_x, double* gbest_val, void* stream)
{
  if (!prob || !pbest_x || !pbest_val)
    return ZEUS_E_ARG;
  zeus_status_t rc = validate(*prob);
  if (rc != ZEUS_OK)
    return rc;
  if (prob->pso_iter < 0)
    return ZEUS_E_ARG;
  if (prob->n_starts == 0)
    return ZEUS_OK;
  std::lock_guard<std::mutex> lock(g_mu);
  if ((rc = ensure_device()) != ZEUS_OK)
    return rc;
  void* base = nullptr;
  if ((rc = scratch_base(&base)) != ZEUS_OK)
    return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g_ds.stream;
  return pso_run_noalloc(prob, pbest_x, pbest_val, gbest_x, gbest_val, base, st);
}

zeus_status_t
zeus_pso_run(const zeus_problem* prob, double* pbest_x, double* pbest_val, double* gbest_x,
             double* gbest_val, void* stream)
{
  if (!prob || !pbest_x || !pbest_val)
    return ZEUS_E_ARG;
  zeus_status_t rc = validate(*prob);
  if (rc != ZEUS_OK)
    return rc;
  if (prob->pso_iter < 0)
    return ZEUS_E_ARG;
  if (prob->n_starts == 0)
    return ZEUS_OK;
  std::lock_guard<std::mutex> lock(g_mu);
  if ((rc = ensure_device()) != ZEUS_OK)
    return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g_ds.stream;
  return pso_phase(prob, 0, prob->pso_iter, nullptr, nullptr, pbest_x, pbest_val, gbest_x,
                   gbest_val, nullptr, false, st);
}

zeus_status_t
zeus_pso_init(const zeus_problem* prob, const zeus_swarm* sw, void* stream)
{
  if (!prob || !sw || !sw->x || !sw->v || !sw->pbest_x || !sw->pbest_val || !sw->gbest_x ||
      !sw->gbest_val || (prob->prng == ZEUS_PRNG_XORWOW && !sw->rng))
    return ZEUS_E_ARG;
  zeus_status_t rc = validate(*prob);
  if (rc != ZEUS_OK)
    return rc;
  if (prob->n_starts == 0)
    return ZEUS_OK;
  std::lock_guard<std::mutex> lock(g_mu);
  if ((rc = ensure_device()) != ZEUS_OK)
    return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g_ds.stream;
  // iteration 0 seeds the XORWOW states; they are kept for the steps
  return pso_phase(prob, 0, 0, sw->x, sw->v, sw->pbest_x, sw->pbest_val, sw->gbest_x,
                   sw->gbest_val, sw->rng, false, st);
}

zeus_status_t
zeus_pso_step(const zeus_problem* prob, int32_t iter, const zeus_swarm* sw, void* stream)
{
  if (!prob || !sw || iter < 1 || !sw->x || !sw->v || !sw->pbest_x || !sw->pbest_val ||
      !sw->gbest_x || !sw->gbest_val || (prob->prng == ZEUS_PRNG_XORWOW && !sw->rng))
    return ZEUS_E_ARG;
  zeus_status_t rc = validate(*prob);
  if (rc != ZEUS_OK)
    return rc;
  if (prob->n_starts == 0)
    return ZEUS_OK;
  std::lock_guard<std::mutex> lock(g_mu);
  if ((rc = ensure_device()) != ZEUS_OK)
    return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g_ds.stream;
  return pso_phase(prob, iter, iter, sw->x, sw->v, sw->pbest_x, sw->pbest_val, sw->gbest_x,
                   sw->gbest_val, sw->rng, true, st);
}

zeus_status_t
zeus_bfgs_update(double* H, const double* s, const double* y, int64_t batch, int32_t n,
                 void* stream)
{
  if (batch < 0 || n <= 0 || n > 1024 || (batch > 0 && (!H || !s || !y)))
    return ZEUS_E_ARG;
  if (batch == 0)
    return ZEUS_OK;
  std::lock_guard<std::mutex> lock(g_mu);
  zeus_status_t rc = ensure_device();
  if (rc != ZEUS_OK)
    return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g_ds.stream;
  if (batch > (int64_t(1) << 31) - 1)
    return ZEUS_E_ARG;
  bfgs_update_kernel<<<int(batch), 256, size_t(2) * n * sizeof(double), st>>>(H, s, y, n);
  hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(st);
  return to_status(e != hipSuccess ? e : e2);
}

zeus_status_t
zeus_eval(const zeus_problem* prob, const double* x, int64_t n, double* f, double* grad,
          int32_t gradient, void* stream)
{
  if (!prob || n < 0 || (n > 0 && (!x || !f)))
    return ZEUS_E_ARG;
  if (gradient != ZEUS_GRAD_ADJOINT && gradient != ZEUS_GRAD_DUAL)
    return ZEUS_E_ARG;
  zeus_status_t rc = validate(*prob);
  if (rc != ZEUS_OK)
    return rc;
  std::lock_guard<std::mutex> lock(g_mu);
  if ((rc = ensure_device()) != ZEUS_OK)
    return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g_ds.stream;
  ProblemDesc d{};
  d.objective = prob->objective;
  d.dim = prob->dim;
  d.params = prob->params;
  d.n_params = prob->n_params;
  const ClassOps* ops = rt::find_class(prob->objective, prob->dim);
  hipError_t e = ops->eval(d, x, n, f, grad, gradient, st);
  const hipError_t e2 = hipStreamSynchronize(st);
  return to_status(e != hipSuccess ? e : e2);
}

zeus_status_t
zeus_argmin(const double* values, int64_t n, int64_t* idx, void* stream)
{
  if (!idx || n < 0 || (n > 0 && !values))
    return ZEUS_E_ARG;
  if (n == 0) {
    *idx = -1;
    return ZEUS_OK;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  zeus_status_t rc = ensure_device();
  if (rc != ZEUS_OK)
    return rc;
  if ((rc = arena_reserve(256)) != ZEUS_OK)
    return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g_ds.stream;
  int64_t* d_out = static_cast<int64_t*>(g_ds.arena);
  if (rt::fault_take_launch())
    return ZEUS_E_LAUNCH;
  argmin_kernel<<<1, 256, 0, st>>>(values, n, d_out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(idx, d_out, sizeof(int64_t), hipMemcpyDeviceToHost, st);
  const hipError_t e2 = hipStreamSynchronize(st);
  return to_status(e != hipSuccess ? e : e2);
}

double
zeus_philox_uniform(uint32_t seed, uint32_t stream, uint64_t start, uint64_t draw)
{
  return rng::uniform(seed, stream, start, draw);
}

zeus_status_t
zeus_philox_fill(double* out, int64_t n, uint32_t seed, uint32_t stream, uint64_t start,
                 uint64_t draw0, void* stream_handle)
{
  if (n < 0 || (n > 0 && !out))
    return ZEUS_E_ARG;
  if (n == 0)
    return ZEUS_OK;
  std::lock_guard<std::mutex> lock(g_mu);
  zeus_status_t rc = ensure_device();
  if (rc != ZEUS_OK)
    return rc;
  hipStream_t st = stream_handle ? static_cast<hipStream_t>(stream_handle) : g_ds.stream;
  philox_fill_kernel<<<int((n + 255) / 256), 256, 0, st>>>(out, n, seed, stream, start, draw0);
  hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(st);
  return to_status(e != hipSuccess ? e : e2);
}

zeus_status_t
zeus_stream_fill(double* out, int64_t n, int32_t prng, uint32_t seed, uint32_t stream,
                 uint64_t start, uint32_t coord, void* stream_handle)
{
  if (n < 0 || (n > 0 && !out) ||
      (prng != ZEUS_PRNG_XORWOW && prng != ZEUS_PRNG_PHILOX && prng != ZEUS_PRNG_SOBOL))
    return ZEUS_E_ARG;
  if (n == 0)
    return ZEUS_OK;
  std::lock_guard<std::mutex> lock(g_mu);
  zeus_status_t rc = ensure_device();
  if (rc != ZEUS_OK)
    return rc;
  hipStream_t st = stream_handle ? static_cast<hipStream_t>(stream_handle) : g_ds.stream;
  stream_fill_kernel<<<1, 1, 0, st>>>(out, n, prng, seed, stream, start, coord);
  hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(st);
  return to_status(e != hipSuccess ? e : e2);
}

zeus_status_t
zeus_alloc(void** ptr, uint64_t bytes)
{
  if (!ptr)
    return ZEUS_E_ARG;
  *ptr = nullptr;
  if (bytes == 0)
    return ZEUS_OK; // size 0 -> nullptr (cuda_buffer.hpp:253-260)
  std::lock_guard<std::mutex> lock(g_mu);
  zeus_status_t rc = ensure_device();
  if (rc != ZEUS_OK)
    return rc;
  return dev_malloc(ptr, bytes) == hipSuccess ? ZEUS_OK : ZEUS_E_ALLOC;
}

zeus_status_t
zeus_free(void* ptr)
{
  if (!ptr)
    return ZEUS_OK;
  return hipFree(ptr) == hipSuccess ? ZEUS_OK : ZEUS_E_LAUNCH;
}

zeus_status_t
zeus_memcpy_h2d(void* dst, const void* src, uint64_t bytes)
{
  if (bytes == 0)
    return ZEUS_OK;
  if (!dst || !src)
    return ZEUS_E_ARG;
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? ZEUS_OK
                                                                          : ZEUS_E_LAUNCH;
}

zeus_status_t
zeus_memcpy_d2h(void* dst, const void* src, uint64_t bytes)
{
  if (bytes == 0)
    return ZEUS_OK;
  if (!dst || !src)
    return ZEUS_E_ARG;
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? ZEUS_OK
                                                                          : ZEUS_E_LAUNCH;
}

zeus_status_t
zeus_memcpy_d2d(void* dst, const void* src, uint64_t bytes)
{
  if (bytes == 0)
    return ZEUS_OK;
  if (!dst || !src)
    return ZEUS_E_ARG;
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? ZEUS_OK
                                                                            : ZEUS_E_LAUNCH;
}

zeus_status_t
zeus_synchronize(void)
{
  return hipDeviceSynchronize() == hipSuccess ? ZEUS_OK : ZEUS_E_LAUNCH;
}

const char*
zeus_status_string(int32_t status)
{
  switch (status) {
    case ZEUS_UNSET: return "unset";
    case ZEUS_SURRENDERED: return "surrendered";
    case ZEUS_CONVERGED: return "converged";
    case ZEUS_STOPPED: return "stopped early";
    case ZEUS_MALLOC: return "device allocation failure";
    case ZEUS_KERNEL: return "kernel launch or runtime error";
    case ZEUS_NONFINITE: return "non-finite value";
    case ZEUS_E_ARG: return "invalid argument";
    case ZEUS_E_UNSUPPORTED: return "objective/dimension not instantiated";
    case ZEUS_E_NODEVICE: return "no gfx950 device";
    case ZEUS_E_IO: return "file could not be written";
    case ZEUS_E_SCRATCH: return "request exceeds the fixed scratch of the allocation-free entry point";
    default: return "unknown";
  }
}

zeus_status_t
zeus_set_fault_hooks(int32_t fail_alloc_at, int32_t fail_launches)
{
  if (fail_alloc_at < 0 || fail_launches < 0)
    return ZEUS_E_ARG;
  std::lock_guard<std::mutex> lock(g_mu);
  rt::g_alloc_count = 0;
  rt::g_fail_alloc_at = fail_alloc_at;
  rt::g_fail_launches = fail_launches;
  // a forced allocation failure must reach the allocator: drop the arena
  if (fail_alloc_at > 0 && g_ds.arena) {
    (void)hipFree(g_ds.arena);
    g_ds.arena = nullptr;
    g_ds.arena_bytes = 0;
  }
  return ZEUS_OK;
}

zeus_status_t
zeus_tuned_config(int32_t objective, int32_t dim, int32_t gradient, int32_t* form,
                  int32_t* threads)
{
  if (!form || !threads)
    return ZEUS_E_ARG;
  std::lock_guard<std::mutex> lock(g_mu);
  const auto it = g_ds.last.find({objective, dim, gradient});
  if (it == g_ds.last.end())
    return ZEUS_E_ARG;
  *form = it->second.form;
  *threads = it->second.threads;
  return ZEUS_OK;
}

} // extern "C"
