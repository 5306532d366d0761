// zeus.hpp -- C++ front end: the reference's zeus::Zeus API for user functors.
//
//   auto r = zeus::Zeus(f, lower, upper, N, MAX_ITER, PSO_ITER,
//                       requiredConverged, fun_name, tolerance, seed, run,
//                       parallel = true, prng = PRNGType::XORWOW,
//                       trajectory_file = {}, nzerosteps = 0,
//                       gradient = Gradient::Adjoint, trajectory_steps = -1,
//                       trajectory_starts = -1);
//
// Same signature, defaults, Result<D> fields and status codes as the
// reference (include/zeus.cuh:341-375).  F is any reference-style callable
// (operator() over std::array<T, D> for T = double and dual::DualNumber);
// when F also has `double value_and_gradient(const double*, double*) const`
// the BFGS uses that adjoint instead of D dual passes (override with the
// trailing `gradient` argument).  Compile the calling file with hipcc for
// gfx950 and link libzeus_amd.so: the PSO, BFGS and summary kernel
// templates are instantiated here for F; device memory and copies go
// through the library's C ABI (zeus_c.h).  F's bytes are uploaded once per
// call and read by the kernels through a pointer, so a Matrix member is
// never deep-copied per launch.
#pragma once

#include "zeus_c.h"

#include "batch.hpp"
#include "bfgs_group.hpp"
#include "bfgs_thread.hpp"
#include "builtins.hpp"
#include "config.hpp"
#include "dual.hpp"
#include "exception.hpp"
#include "hip_buffer.hpp"
#include "matrix.hpp"
#include "objectives.hpp"
#include "prng.hpp"
#include "pso.hpp"
#include "report.hpp"
#include "result.hpp"
#include "summary.hpp"
#include "traits.hpp"
#include "util_types.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace zeus {

// Every call appends its row to zeus_<D>d_results.tsv in the working
// directory, as the reference does (include/zeus.cuh:172-183); a process
// can switch that off here.
inline bool&
results_table_enabled()
{
  static bool on = true;
  return on;
}
inline void
set_results_table(bool on)
{
  results_table_enabled() = on;
}

namespace detail {

// Metrics of one kind of work from per-start tick arrays, printed the way
// reference report_metrics_and_cleanup does (bfgs_sequential.cuh:240-382).
inline Metrics
report_metrics(const char* label, const std::vector<uint64_t>& ticks,
               const std::vector<int32_t>& calls, const std::vector<uint64_t>& total,
               float ms_opt, double tick_hz)
{
  zeus_metrics m{};
  const int64_t n = int64_t(ticks.size());
  (void)zeus_metrics_from_ticks(ticks.data(), calls.data(), total.data(), n, 256, ms_opt,
                                tick_hz, &m);
  unsigned long long sum = 0;
  for (auto t : ticks)
    sum += t;
  const double avg_ticks = n > 0 ? double(sum) / double(n) : 0.0;
  std::printf("[%s] Kernel wall time: %.3f ms\n", label, ms_opt);
  std::printf("[%s] Avg cycles/thread: %.3f ; avg ms/thread: %.6f\n", label, avg_ticks,
              avg_ticks / tick_hz * 1e3);
  std::printf("[%s] Calls/thread \u2248 %.2f; ms/call \u2248 %.6f\n", label,
              m.calls_per_thread_mean, m.ms_per_call);
  std::printf("[%s] Heuristic fraction: %.3f%%\n", label, 100.0 * m.fraction_of_kernel);
  std::printf("[%s] Serialized sum: %.3f ms; fraction vs kernel: %.3f%%\n", label,
              double(sum) / tick_hz * 1e3, 100.0 * m.serialized);
  std::printf("[%s] Block-level fraction p95: %.3f%%\n", label, 100.0 * m.block95);
  Metrics out;
  out.label = label;
  out.ms_per_call = m.ms_per_call;
  out.calls_per_thread_mean = m.calls_per_thread_mean;
  out.fraction_of_kernel = m.fraction_of_kernel;
  out.block95 = m.block95;
  out.serialized = m.serialized;
  out.fraction_of_thread = m.fraction_of_thread;
  return out;
}

inline void
hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess)
    throw HipError(e, what);
}

inline void
abi_check(int rc, const char* what)
{
  if (rc != ZEUS_OK)
    throw HipError(hip_error_of(rc), what);
}

// Workgroup size of the per-workgroup form for user functors: the objective
// is evaluated by one lane and its dual passes spread over the lanes.
template <int D>
constexpr int
front_group_threads()
{
  return D <= 64 ? 64 : (D <= 256 ? 256 : 512);
}

inline int
min_kind_of(std::string_view name)
{
  if (name == "rosenbrock")
    return 1;
  if (name == "goldstein_price")
    return 2;
  return 0;
}

// Trajectory request of one call: file, bounds, and whether the line-search
// steps are kept (the reference records them only in its parallel form).
struct TrajSpec {
  std::string file;
  int steps = 0;
  int starts = 0;
  bool alpha = false;
};

template <class F, int D, Gradient G>
Result<D>
run(const F& f, double lower, double upper, std::size_t N, int MAX_ITER, int PSO_ITER,
    int required, std::string_view fun_name, double tol, int seed, bool parallel,
    const TrajSpec* tr = nullptr, int run_id = 0, int prng = rng::kXorwow)
{
  Result<D> res;
  if (N == 0)
    return res; // reference zeus.cuh:58-60
  if (MAX_ITER < 0)
    throw std::invalid_argument("MAX_ITER must be >= 0");
  try {
    using namespace dev;
    const bool pso = PSO_ITER >= 0;
    const int64_t n = int64_t(N);
    const int nblk = int((n + 255) / 256);

    HipBuffer<unsigned char> fbuf(sizeof(F));
    abi_check(zeus_memcpy_h2d(fbuf.data(), &f, sizeof(F)), "functor upload");
    const UserProvider<F> prov{reinterpret_cast<const F*>(fbuf.data())};

    HipBuffer<double> X(pso ? n * D : 0), V(pso ? n * D : 0), pbx(pso ? n * D : 0),
      pbv(pso ? n : 0), gbx(D), part_val(2 * nblk), part_x(pso ? 2 * int64_t(nblk) * D : 0),
      gstate(2 * (D + 1));
    HipBuffer<int64_t> part_idx(2 * nblk), bxo(1);
    HipBuffer<int32_t> bprob(nblk);
    std::vector<int32_t> zeros(nblk, 0);
    abi_check(zeus_memcpy_h2d(bprob.data(), zeros.data(), 4 * nblk), "upload");
    const int64_t zero64 = 0;
    abi_check(zeus_memcpy_h2d(bxo.data(), &zero64, 8), "upload");

    ProblemDesc P{};
    P.dim = D;
    P.lower = lower;
    P.upper = upper;
    P.start0 = 0;
    P.n_starts = n;
    P.required = required;
    P.max_iter = MAX_ITER;
    P.pso_iter = PSO_ITER;
    P.block0 = 0;
    P.min_kind = min_kind_of(fun_name);
    P.seed = uint32_t(seed);
    P.stream = 0;
    P.tol = tol;
    P.x0 = pso ? pbx.data() : nullptr;
    // XORWOW: one state per (start, coordinate) for the PSO iterations
    P.prng = prng;
    const bool xorwow = prng == rng::kXorwow && PSO_ITER > 0;
    HipBuffer<uint32_t> xw(xorwow ? 6 * n * D : 0);
    const int64_t t_steps = tr ? tr->steps : 0, t_starts = tr ? tr->starts : 0;
    const int64_t t_n = t_steps * t_starts;
    HipBuffer<double> t_coords(t_n * D), t_fval(t_n), t_grad(t_n),
      t_alpha(tr && tr->alpha ? t_n : 0);
    HipBuffer<int8_t> t_status(t_n);
    if (t_n > 0) {
      P.traj = TrajOut{t_coords.data(), t_fval.data(), t_grad.data(), t_status.data(),
                       tr->alpha ? t_alpha.data() : nullptr, int32_t(t_steps),
                       int32_t(t_starts)};
    }
    HipBuffer<ProblemDesc> dP(1);
    abi_check(zeus_memcpy_h2d(dP.data(), &P, sizeof(P)), "upload");
    HipBuffer<Control> ctl(1);

    HipBuffer<int32_t> st(n), it(n), ls(n), gr(n);
    HipBuffer<double> fv(n), gn(n), xo(n * D), bx(D);
    HipBuffer<int8_t> near(n);
    StartOut out{st.data(), it.data(), fv.data(), gn.data(), xo.data(), ls.data(), gr.data(),
                 near.data()};
    // the sequential form times its gradients and updates (the reference's
    // parallel form reports zero metrics)
    const int64_t nm = parallel ? 0 : n;
    HipBuffer<uint64_t> m_at(nm), m_bt(nm), m_tt(nm);
    HipBuffer<int32_t> m_ac(nm), m_bc(nm);
    if (!parallel) {
      out.ad_ticks = m_at.data();
      out.bfgs_ticks = m_bt.data();
      out.total_ticks = m_tt.data();
      out.ad_calls = m_ac.data();
      out.bfgs_calls = m_bc.data();
    }

    constexpr int NT = front_group_threads<D>();
    constexpr bool kHReg = RowChunk<D, NT>::kPerThread <= 50;
    int grid = 1;
    int64_t slab = 0, stride = 0;
    if (parallel) {
      int dev = 0, n_cu = 1, occ = 1;
      hip_check(hipGetDevice(&dev), "hipGetDevice");
      hip_check(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev),
                "hipDeviceGetAttribute");
      hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &occ,
                  reinterpret_cast<const void*>(
                    &bfgs_group_kernel<F, D, NT, G, kHReg, UserProvider<F>>),
                  NT, 0),
                "occupancy");
      grid = int(std::max<int64_t>(1, std::min<int64_t>(n, int64_t(n_cu) * std::max(occ, 1))));
      if (!kHReg) {
        stride = int64_t(RowChunk<D, NT>::kPerThread) * NT;
        slab = stride * grid;
      }
    } else if (D > kThreadRegMaxD) {
      slab = n * (int64_t(D) * (D + 1) / 2);
    }
    HipBuffer<double> hslab(slab);

    hipStream_t s = nullptr;
    hipEvent_t ep, e0, e1;
    // timing only: recorded without the system-scope fence (the results come
    // back through synchronous copies)
    hip_check(hipEventCreateWithFlags(&ep, hipEventDisableSystemFence), "event");
    hip_check(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence), "event");
    hip_check(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence), "event");
    init_controls_kernel<<<1, 256, 0, s>>>(ctl.data(), 1, gbx.data(), D);
    hip_check(hipEventRecord(ep, s), "event");
    if (pso) {
      PsoArgs pa{dP.data(),       1,           0,           nblk,           ctl.data(),
                 X.data(),        V.data(),    pbx.data(),  pbv.data(),     gbx.data(),
                 part_val.data(), part_idx.data(), part_x.data(), gstate.data(), bprob.data(),
                 xorwow ? xw.data() : nullptr, n * D};
      for (int t = 0; t <= PSO_ITER; ++t) {
        pa.iter = t;
        pso_lane_kernel<F, D, UserProvider<F>><<<nblk, 256, 0, s>>>(pa, prov);
      }
    }
    hip_check(hipGetLastError(), "PSO launch");
    if (t_n > 0) {
      traj_fill_kernel<<<int(std::min<int64_t>((t_n * D + 255) / 256, 4096)), 256, 0, s>>>(
        P.traj, D);
      hip_check(hipGetLastError(), "trajectory fill");
    }
    hip_check(hipEventRecord(e0, s), "event");
    ClassArgs a{dP.data(), 1, D, n, ctl.data(), nullptr, out, hslab.data(), stride};
    if (parallel)
      bfgs_group_kernel<F, D, NT, G, kHReg, UserProvider<F>><<<grid, NT, 0, s>>>(a, prov);
    else
      bfgs_thread_kernel<F, D, G, UserProvider<F>><<<nblk, 256, 0, s>>>(a, prov);
    hip_check(hipGetLastError(), "BFGS launch");
    hip_check(hipEventRecord(e1, s), "event");
    summary_kernel<<<1, kSummaryThreads, 0, s>>>(dP.data(), ctl.data(), out, bx.data(), bxo.data());
    hip_check(hipGetLastError(), "summary launch");
    hip_check(hipStreamSynchronize(s), "BFGS runtime");
    float ms = 0.f, ms_pso = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventElapsedTime(&ms_pso, ep, e0);
    (void)hipEventDestroy(ep);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);

    Control k{};
    abi_check(zeus_memcpy_d2h(&k, ctl.data(), sizeof(k)), "result copy");
    abi_check(zeus_memcpy_d2h(res.coordinates, bx.data(), sizeof(double) * D), "result copy");
    res.idx = k.best_idx;
    res.status = k.best_status;
    res.iter = k.best_iter;
    res.fval = k.best_fval;
    res.gradientNorm = k.best_gnorm;
    res.c.claimed = k.claimed;
    res.c.actual = k.actual;
    res.c.stopped = k.stopped;
    res.c.surrendered = k.surrendered;
    res.ms_opt = ms;
    if (!parallel) {
      std::vector<uint64_t> at(n), bt(n), tt(n);
      std::vector<int32_t> ac(n), bc(n);
      abi_check(zeus_memcpy_d2h(at.data(), m_at.data(), 8 * n), "metrics copy");
      abi_check(zeus_memcpy_d2h(bt.data(), m_bt.data(), 8 * n), "metrics copy");
      abi_check(zeus_memcpy_d2h(tt.data(), m_tt.data(), 8 * n), "metrics copy");
      abi_check(zeus_memcpy_d2h(ac.data(), m_ac.data(), 4 * n), "metrics copy");
      abi_check(zeus_memcpy_d2h(bc.data(), m_bc.data(), 4 * n), "metrics copy");
      int dev = 0, khz = 0;
      hip_check(hipGetDevice(&dev), "hipGetDevice");
      double hz = 1e8;
      if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess &&
          khz > 0)
        hz = double(khz) * 1e3;
      res.ad = report_metrics("AD", at, ac, tt, ms, hz);
      res.bfgs = report_metrics("BFGS", bt, bc, tt, ms, hz);
    }
    if (results_table_enabled()) { // zeus.cuh:172-183
      const double err = util::calculateEuclideanError(std::string(fun_name), res.coordinates, D);
      util::appendResultsToTsv<std::size_t(D)>(D, int(n), std::string(fun_name), 0.f, ms_pso,
                                               ms, 0.f, MAX_ITER, PSO_ITER, err, run_id, res);
    }
    if (t_n > 0) { // zeus.cuh:189-239
      std::vector<double> hc(t_n * D), hf(t_n), hg(t_n), ha(t_n, 0.0);
      std::vector<int8_t> hs(t_n);
      abi_check(zeus_memcpy_d2h(hc.data(), t_coords.data(), 8 * hc.size()), "trajectory copy");
      abi_check(zeus_memcpy_d2h(hf.data(), t_fval.data(), 8 * hf.size()), "trajectory copy");
      abi_check(zeus_memcpy_d2h(hg.data(), t_grad.data(), 8 * hg.size()), "trajectory copy");
      abi_check(zeus_memcpy_d2h(hs.data(), t_status.data(), hs.size()), "trajectory copy");
      if (tr->alpha)
        abi_check(zeus_memcpy_d2h(ha.data(), t_alpha.data(), 8 * ha.size()), "trajectory copy");
      const int rc = zeus_write_trajectory(tr->file.c_str(), hc.data(), hf.data(), hg.data(),
                                           hs.data(), ha.data(), int32_t(t_starts),
                                           int32_t(t_steps), D);
      if (rc != ZEUS_OK) // reported, the result still returned (zeus.cuh:229-238)
        std::fprintf(stderr, "zeus: trajectory not written to %s: %s\n", tr->file.c_str(),
                     zeus_status_string(rc));
    }
  } catch (const HipError& e) {
    res = Result<D>{};
    res.status = (e.code() == hipErrorOutOfMemory || e.code() == hipErrorMemoryAllocation)
                   ? kStatusMalloc
                   : kStatusKernel;
  }
  return res;
}

// A built-in objective (builtins.hpp) through zeus_run_batch: the library's
// kernels for it, autotuned.  Returns false when the library does not
// instantiate the objective at this dimension (the caller then runs the
// functor path).
template <class F, int D>
bool
run_builtin(const F& f, double lower, double upper, std::size_t N, int MAX_ITER, int PSO_ITER,
            int required, std::string_view fun_name, double tol, int seed, int run_id, int prng,
            Gradient gradient, Result<D>& res)
{
  const std::string name(fun_name);
  zeus_problem p{};
  p.objective = F::zeus_objective;
  p.dim = D;
  p.lower = lower;
  p.upper = upper;
  p.n_starts = int64_t(N);
  p.max_iter = MAX_ITER;
  p.pso_iter = PSO_ITER;
  p.required = required;
  p.seed = uint32_t(seed);
  p.stream = 0;
  p.prng = prng;
  p.tol = tol;
  p.params = f.zeus_params();
  p.n_params = int64_t(f.zeus_n_params());
  p.name = name.c_str();
  zeus_options o{};
  o.gradient = gradient == Gradient::Dual ? ZEUS_GRAD_DUAL : ZEUS_GRAD_ADJOINT;
  o.form = ZEUS_FORM_GROUP;
  zeus_result r{};
  double bx[D];
  const int rc = zeus_run_batch(&p, 1, &o, &r, bx, nullptr, nullptr);
  if (rc == ZEUS_E_UNSUPPORTED)
    return false;
  res = Result<D>{};
  if (rc == ZEUS_E_ALLOC || rc == ZEUS_E_LAUNCH || rc == ZEUS_E_NODEVICE) {
    res.status = rc == ZEUS_E_ALLOC ? kStatusMalloc : kStatusKernel; // zeus.cuh:108-116
    return true;
  }
  if (rc != ZEUS_OK)
    throw std::invalid_argument(std::string("zeus::Zeus: ") + zeus_status_string(rc));
  res.idx = int(r.idx);
  res.status = r.status;
  res.iter = r.iter;
  res.fval = r.fval;
  res.gradientNorm = r.gradient_norm;
  for (int d = 0; d < D; ++d)
    res.coordinates[d] = bx[d];
  res.c.claimed = r.claimed;
  res.c.actual = r.actual;
  res.c.stopped = r.stopped;
  res.c.surrendered = r.surrendered;
  res.ms_opt = r.ms_opt;
  if (results_table_enabled()) { // zeus.cuh:172-183
    const double err = util::calculateEuclideanError(name, res.coordinates, D);
    util::appendResultsToTsv<std::size_t(D)>(D, int(N), name, 0.f, float(r.ms_pso),
                                             float(r.ms_opt), 0.f, MAX_ITER, PSO_ITER, err,
                                             run_id, res);
  }
  return true;
}

} // namespace detail

template <typename Function, std::size_t ZEUS_DIM = FnTraits<Function>::arity>
  requires ZeusObjective<Function, ZEUS_DIM>
auto
Zeus(Function const& f, double lower, double upper, std::size_t N, int MAX_ITER, int PSO_ITER,
     int requiredConverged, std::string_view fun_name, double tolerance, int seed, int run,
     bool parallel = true, PRNGType prng = PRNGType::XORWOW,
     std::string_view trajectory_file = {}, int nzerosteps = 0,
     Gradient gradient = Gradient::Adjoint, int trajectory_steps = -1,
     int trajectory_starts = -1) -> Result<ZEUS_DIM>
{
  const int kind = prng == PRNGType::PHILOX ? rng::kPhilox
                   : prng == PRNGType::SOBOL ? rng::kSobol
                                             : rng::kXorwow;
  (void)nzerosteps; // inert: the line search never returns a zero step
  constexpr int D = int(ZEUS_DIM);
  // trajectory_file: every start's iterations, written as .tsv (bounded by
  // trajectory_steps / trajectory_starts; -1 = MAX_ITER / N)
  detail::TrajSpec spec;
  const detail::TrajSpec* tr = nullptr;
  if (!trajectory_file.empty() && N > 0 && MAX_ITER > 0) {
    spec.file = std::string(trajectory_file);
    spec.steps = trajectory_steps < 0 ? MAX_ITER : std::min(trajectory_steps, MAX_ITER);
    spec.starts = trajectory_starts < 0 ? int(N) : int(std::min<std::size_t>(trajectory_starts, N));
    spec.alpha = parallel;
    tr = &spec;
  }
  // built-in objectives run on the library's kernels (parallel form, no
  // trajectory), unless the library lacks this dimension
  if constexpr (IsBuiltin<Function>)
    if (parallel && tr == nullptr && N > 0 && MAX_ITER >= 0) {
      Result<ZEUS_DIM> res;
      if (detail::run_builtin<Function, D>(f, lower, upper, N, MAX_ITER, PSO_ITER,
                                           requiredConverged, fun_name, tolerance, seed, run,
                                           kind, gradient, res))
        return res;
    }
  if constexpr (HasAdjoint<Function>)
    if (gradient == Gradient::Adjoint)
      return detail::run<Function, D, Gradient::Adjoint>(f, lower, upper, N, MAX_ITER, PSO_ITER,
                                                         requiredConverged, fun_name, tolerance,
                                                         seed, parallel, tr, run, kind);
  return detail::run<Function, D, Gradient::Dual>(f, lower, upper, N, MAX_ITER, PSO_ITER,
                                                  requiredConverged, fun_name, tolerance, seed,
                                                  parallel, tr, run, kind);
}

} // namespace zeus
