// bfgs_thread.hpp -- per-thread BFGS: one lane owns one start.
//
// Restates reference include/bfgs_sequential.cuh:16-234 (step order) with
// the termination rules of include/bfgs_common.cuh:108-275.  Differences
// by design: H (symmetric by construction) is stored as a packed upper
// triangle -- in registers for small D, otherwise in an interleaved global
// slab -- and updated with the O(D^2) rank-2 form of the reference's
// O(D^4) product; no device malloc; the accepted line-search value is
// reused as f(x_new).
//
// Arithmetic contract: every operation and its order match the CPU oracle
// (serial left-to-right sums with fma, the same rank-2 formula), so with
// objectives made of + - * / the results are bit-identical to the oracle.
#pragma once

#include "batch.hpp"
#include "lane_ops.hpp"
#include "philox.hpp"
#include "prng.hpp"

namespace zeus::dev {

// Threshold for keeping the packed H of the per-thread form in registers;
// loops are unrolled only below it (above it, state lives in memory).
constexpr int kThreadRegMaxD = 8;

template <int D>
ZEUS_HD constexpr int
packed_index(int i, int j)
{
  return i <= j ? i * D - i * (i - 1) / 2 + (j - i) : j * D - j * (j - 1) / 2 + (i - j);
}

// Packed symmetric H in registers (compile-time indices after unrolling).
template <int D>
struct RegPackedH {
  static constexpr int kSize = D * (D + 1) / 2;
  double h[kSize];
  ZEUS_D ZEUS_FI double& at(int i, int j) { return h[packed_index<D>(i, j)]; }
};

// Packed symmetric H in global memory, element k of a start at
// base[k * stride] (starts interleaved so that lanes coalesce).
template <int D>
struct GlobalPackedH {
  static constexpr int kSize = D * (D + 1) / 2;
  double* base;
  int64_t stride;
  ZEUS_D ZEUS_FI double& at(int i, int j)
  {
    return base[int64_t(packed_index<D>(i, j)) * stride];
  }
};

template <class F, int D, Gradient G, class HS>
ZEUS_D void
thread_bfgs(const F& f, const ProblemDesc& P, Control* ctl, int64_t s,
            const double* starts, const StartOut& out, HS& H)
{
  constexpr int kU = D <= kThreadRegMaxD ? D : 1; // unroll factor
  const int64_t sl = s - P.start0; // start index within its problem
  // timing (bfgs_sequential.cuh:50,86-97,149-173,222-232), when requested
  const bool timed = out.ad_ticks != nullptr;
  const uint64_t k_begin = timed ? uint64_t(wall_clock64()) : 0;
  uint64_t ad_t = 0, bfgs_t = 0;
  int ad_n = 0, bfgs_n = 0;
  double x[D], g[D], p[D], xt[D], gn[D], dx[D], dg[D];

#pragma unroll(kU)
  for (int i = 0; i < D; ++i)
#pragma unroll(kU)
    for (int j = i; j < D; ++j)
      H.at(i, j) = (i == j) ? 1.0 : 0.0;

#pragma unroll(kU)
  for (int d = 0; d < D; ++d) {
    if (starts)
      x[d] = starts[s * D + d];
    else if (P.x0)
      x[d] = P.x0[sl * D + d];
    else
      x[d] = start_uniform_in(P, uint64_t(sl), d);
    g[d] = 0.0;
  }

  const double f0 = lane_value<F, D>(f, x);
  double bestVal = f0;
  {
    const uint64_t t0 = timed ? uint64_t(wall_clock64()) : 0;
    lane_gradient<F, D, G>(f, x, g);
    if (timed) {
      ad_t += uint64_t(wall_clock64()) - t0;
      ++ad_n;
    }
  }
  int grads = 1, ls_evals = 0;

  int status = kStatusUnset;
  double fval = 333777.0, gnorm = 69.0;
  int iter = 0;
  bool done = false;
  for (; iter < P.max_iter; ++iter) {
    // trajectory step: the point, its value and gradient norm, status -1
    // (bfgs_sequential.cuh:100-109)
    const int64_t to = traj_slot(P.traj, sl, iter);
    if (to >= 0) {
      double gg = 0.0;
#pragma unroll(kU)
      for (int i = 0; i < D; ++i)
        gg = ::fma(g[i], g[i], gg);
#pragma unroll(kU)
      for (int d = 0; d < D; ++d)
        P.traj.coords[(int64_t(iter) * D + d) * P.traj.starts + sl] = x[d];
      P.traj.fval[to] = bestVal;
      P.traj.grad[to] = ::sqrt(gg);
      P.traj.status[to] = int8_t(kStatusUnset);
    }
    // stop flag raised by another start (bfgs_common.cuh:115-141)
    if (load_flag(&ctl->stop_flag) != 0) {
      status = kStatusStopped;
      fval = bestVal; // == f(x): x is where bestVal was evaluated
      double gg = 0.0;
#pragma unroll(kU)
      for (int i = 0; i < D; ++i)
        gg = ::fma(g[i], g[i], gg);
      gnorm = ::sqrt(gg);
      done = true;
      break;
    }
    // p = -H g (utils.cuh:148-161)
#pragma unroll(kU)
    for (int i = 0; i < D; ++i) {
      double acc = 0.0;
#pragma unroll(kU)
      for (int j = 0; j < D; ++j)
        acc = ::fma(H.at(i, j), g[j], acc);
      p[i] = -acc;
    }
    // backtracking Armijo search (utils.cuh:376-426)
    double ddir = 0.0;
#pragma unroll(kU)
    for (int i = 0; i < D; ++i)
      ddir = ::fma(g[i], p[i], ddir);
    double alpha = 1.0, f1 = 0.0;
    bool accepted = false;
    for (int t = 0; t < kLineSearchTrials; ++t) {
#pragma unroll(kU)
      for (int j = 0; j < D; ++j)
        xt[j] = ::fma(alpha, p[j], x[j]);
      f1 = lane_value<F, D>(f, xt);
      ++ls_evals;
      if (f1 <= ::fma(kArmijoC1 * alpha, ddir, bestVal) && f1 < bestVal) {
        accepted = true;
        break;
      }
      alpha *= 0.5;
    }
    if (to >= 0 && P.traj.alpha)
      P.traj.alpha[to] = alpha;
#pragma unroll(kU)
    for (int i = 0; i < D; ++i) {
      xt[i] = ::fma(alpha, p[i], x[i]);
      dx[i] = xt[i] - x[i];
    }
    double fnew = f1;
    if (!accepted) {
      fnew = lane_value<F, D>(f, xt);
      ++ls_evals;
    }
    {
      const uint64_t t0 = timed ? uint64_t(wall_clock64()) : 0;
      lane_gradient<F, D, G>(f, xt, gn);
      if (timed) {
        ad_t += uint64_t(wall_clock64()) - t0;
        ++ad_n;
      }
    }
    ++grads;
    const uint64_t b0 = timed ? uint64_t(wall_clock64()) : 0;
    double sTy = 0.0;
#pragma unroll(kU)
    for (int i = 0; i < D; ++i) {
      dg[i] = gn[i] - g[i];
      sTy = ::fma(dx[i], dg[i], sTy);
    }
    // rank-2 form of utils.cuh:245-287 (skipped when |s.y| < 1e-14)
    if (!(::fabs(sTy) < kCurvatureSkip)) {
      const double rho = 1.0 / sTy;
      double u[D];
#pragma unroll(kU)
      for (int i = 0; i < D; ++i) {
        double acc = 0.0;
#pragma unroll(kU)
        for (int j = 0; j < D; ++j)
          acc = ::fma(H.at(i, j), dg[j], acc);
        u[i] = acc;
      }
      double yu = 0.0;
#pragma unroll(kU)
      for (int i = 0; i < D; ++i)
        yu = ::fma(dg[i], u[i], yu);
      const double beta = ::fma(rho * rho, yu, rho);
      const double hb = 0.5 * beta;
#pragma unroll(kU)
      for (int i = 0; i < D; ++i)
        u[i] = ::fma(hb, dx[i], -(rho * u[i])); // u now holds a
#pragma unroll(kU)
      for (int i = 0; i < D; ++i)
#pragma unroll(kU)
        for (int j = i; j < D; ++j)
          H.at(i, j) = H.at(i, j) + (dx[i] * u[j] + u[i] * dx[j]);
    }
    if (timed) {
      bfgs_t += uint64_t(wall_clock64()) - b0;
      ++bfgs_n;
    }
    if (fnew < bestVal) {
      bestVal = fnew;
#pragma unroll(kU)
      for (int i = 0; i < D; ++i) {
        x[i] = xt[i];
        g[i] = gn[i];
      }
    }
    double gg = 0.0;
#pragma unroll(kU)
    for (int i = 0; i < D; ++i)
      gg = ::fma(g[i], g[i], gg);
    const double nrm = ::sqrt(gg);
    if (!::isfinite(nrm) || !::isfinite(fnew)) {
      status = kStatusNonFinite;
      fval = fnew;
      gnorm = nrm;
      done = true;
      break;
    }
    if (nrm < P.tol) {
      count_converged(ctl, P.required);
      status = kStatusConverged;
      fval = bestVal;
      gnorm = nrm;
      done = true;
      break;
    }
  }
  if (!done) { // surrendered at MAX_ITER (bfgs_common.cuh:213-238)
    status = kStatusSurrendered;
    fval = bestVal;
    double gg = 0.0;
#pragma unroll(kU)
    for (int i = 0; i < D; ++i)
      gg = ::fma(g[i], g[i], gg);
    gnorm = ::sqrt(gg);
  }
  traj_final_status(P.traj, sl, status, iter, P.max_iter);
  out.status[s] = status;
  out.iter[s] = iter;
  out.fval[s] = fval;
  out.gnorm[s] = gnorm;
#pragma unroll(kU)
  for (int d = 0; d < D; ++d)
    out.x[s * D + d] = x[d];
  out.ls_evals[s] = ls_evals;
  out.grads[s] = grads;
  int8_t near = 0;
  if (status == kStatusConverged) {
    double e2 = 0.0;
    for (int d = 0; d < D; ++d) {
      const double df = x[d] - known_min(P.min_kind, d);
      e2 += df * df;
    }
    near = ::sqrt(e2) < 0.5;
  }
  out.near[s] = near;
  if (timed) {
    out.ad_ticks[s] = ad_t;
    out.ad_calls[s] = ad_n;
    out.bfgs_ticks[s] = bfgs_t;
    out.bfgs_calls[s] = bfgs_n;
    out.total_ticks[s] = uint64_t(wall_clock64()) - k_begin;
  }
}

template <class F, int D, Gradient G, class Prov>
__global__ __launch_bounds__(256) void
bfgs_thread_kernel(ClassArgs a, Prov prov)
{
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= a.total_starts)
    return;
  const int pi = find_problem(a.probs, a.n_probs, s);
  const ProblemDesc P = a.probs[pi];
  decltype(auto) f = prov.template get<F>(P);
  if constexpr (D <= kThreadRegMaxD) {
    RegPackedH<D> H;
    thread_bfgs<F, D, G>(f, P, a.ctl + pi, s, a.starts, a.out, H);
  } else {
    GlobalPackedH<D> H{a.h_slab + s, a.total_starts};
    thread_bfgs<F, D, G>(f, P, a.ctl + pi, s, a.starts, a.out, H);
  }
}

} // namespace zeus::dev
