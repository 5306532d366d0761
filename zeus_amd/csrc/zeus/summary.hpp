// summary.hpp -- control-block setup, objective preparation and the
// per-problem result reduction.
#pragma once

#include "batch.hpp"
#include "pso.hpp"

#include <limits>

namespace zeus::dev {

// Zero every control block (a fill kernel rather than a memset node, so the
// sequence can be captured into a graph and replayed).
static __global__ void
init_controls_kernel(Control* c, int n, double* gbx, int64_t n_gbx)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    reset_control(c + i);
  for (int64_t k = i; k < n_gbx; k += int64_t(gridDim.x) * blockDim.x)
    gbx[k] = 0.0;
}

// Zero n 32-bit words (a fill kernel: capturable, no memset node).
static __global__ void
zero_words_kernel(uint32_t* p, int n)
{
  const int i = int(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = 0u;
}

// Trajectory buffers before a run (reference zeus.cuh:77-103): NaN
// coordinates, values and gradient norms, status -1, alpha 0.
static __global__ void
traj_fill_kernel(TrajOut t, int D)
{
  const int64_t n = int64_t(t.steps) * t.starts, nc = n * D;
  const double nan = __builtin_nan("");
  const int64_t step = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nc; i += step) {
    t.coords[i] = nan;
    if (i < n) {
      t.fval[i] = nan;
      t.grad[i] = nan;
      t.status[i] = int8_t(kStatusUnset);
      if (t.alpha)
        t.alpha[i] = 0.0;
    }
  }
}

// Cs = (C + C^T) / 2 for a Gaussian problem (same quadratic form), row
// major, and the same matrix in MFMA A-operand order: fragment (mb, kb) is
// 64 consecutive doubles, lane l holding Cs[16 mb + l % 16][4 kb + l / 16]
// (zero outside D x D), so a fragment load is one coalesced 512-byte read.
// C is read through the problem descriptor in device memory (uploaded by
// every call), never taken as a kernel argument, so that a replayed graph
// prepares the matrix of the call that replays it.
ZEUS_D ZEUS_FI void
symmetrize_element(const ProblemDesc* desc, int k, double* Cs, double* Cf, int D, int e)
{
  const double* C = desc[k].params;
  if (e < D * D) {
    const int i = e / D, j = e - i * D;
    Cs[e] = 0.5 * (C[int64_t(i) * D + j] + C[int64_t(j) * D + i]);
  }
  const int KB = (D + 3) / 4, MB = (D + 15) / 16;
  if (e < MB * KB * 64) {
    const int lane = e & 63;
    const int f = e >> 6;
    const int mb = f / KB, kb = f - mb * KB;
    const int r = 16 * mb + (lane & 15), c = 4 * kb + (lane >> 4);
    Cf[e] = (r < D && c < D) ? 0.5 * (C[int64_t(r) * D + c] + C[int64_t(c) * D + r]) : 0.0;
  }
  // upper fragments (cs_product.hpp, CsUpper): row block mb keeps k steps
  // 4 mb .. KB - 1, the diagonal block (k steps 4 mb .. 4 mb + 3) times 0.5
  int nu = 0;
  for (int mb = 0; mb < MB; ++mb)
    nu += KB - 4 * mb;
  if (e < nu * 64) {
    const int lane = e & 63;
    int f = e >> 6;
    int mb = 0;
    while (f >= KB - 4 * mb) {
      f -= KB - 4 * mb;
      ++mb;
    }
    const int kb = 4 * mb + f;
    const int r = 16 * mb + (lane & 15), c = 4 * kb + (lane >> 4);
    const double v =
      (r < D && c < D) ? 0.5 * (C[int64_t(r) * D + c] + C[int64_t(c) * D + r]) : 0.0;
    Cf[int64_t(MB) * KB * 64 + e] = kb < 4 * mb + 4 ? 0.5 * v : v;
  }
}

// element indices stay below D * D + (fragments) * 64 < 2^31: 32-bit
// division (64-bit division is a long software sequence)
static __global__ void
symmetrize_kernel(const ProblemDesc* desc, int k, double* Cs, double* Cf, int D)
{
  symmetrize_element(desc, k, Cs, Cf, D, int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x));
}

// init_controls_kernel and symmetrize_kernel in one launch, for a call whose
// first class is one Gaussian problem (the grid covers both element ranges).
static __global__ void
init_symmetrize_kernel(Control* c, int n, double* gbx, int64_t n_gbx, const ProblemDesc* desc,
                       int k, double* Cs, double* Cf, int D)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n)
    reset_control(c + i);
  for (int64_t kk = i; kk < n_gbx; kk += int64_t(gridDim.x) * blockDim.x)
    gbx[kk] = 0.0;
  symmetrize_element(desc, k, Cs, Cf, D, int(i));
}

// Doubles of the Gaussian preparation per problem: Cs, its fragments, then
// the upper fragments.
ZEUS_HD constexpr int64_t
gaussian_aux_doubles(int D)
{
  const int KB = (D + 3) / 4, MB = (D + 15) / 16;
  int64_t nu = 0;
  for (int mb = 0; mb < MB; ++mb)
    nu += KB - 4 * mb;
  return int64_t(D) * D + (int64_t(MB) * KB + nu) * 64;
}

// One workgroup per problem: deterministic argmin of the start values (the
// reference's launchReduction, bfgs_common.cuh:280-319, with NaN skipped
// and ties to the lowest index) plus the status counts of
// util.hpp:151-200 and the work counters used for the roofline.  The
// kernel is bound by instruction issue on its one CU, not by its loads:
// 512 threads measured 9.3 us per call at 4096 starts, 1024 threads 11.1 us
// and 256 threads 10.7 us.
constexpr int kSummaryThreads = 512;

static __global__ __launch_bounds__(kSummaryThreads) void
summary_kernel(const ProblemDesc* probs, Control* ctl, StartOut out, double* best_x,
               const int64_t* best_x_off)
{
  constexpr int NW = kSummaryThreads / kWave;
  __shared__ double sv[NW];
  __shared__ int64_t si[NW];
  __shared__ int64_t sc[NW][8];
  __shared__ int64_t sbest;
  const int pi = blockIdx.x;
  const ProblemDesc P = probs[pi];
  double v = __builtin_nan("");
  int64_t idx = int64_t(1) << 62;
  int my_status = 0, my_iter = 0; // fields of this thread's best start
  double my_gnorm = 0.0;
  int64_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // Four strides at a time: every load of a group is issued (at a clamped,
  // always valid start) before any is used, then the group is reduced with
  // selects; each thread still visits its starts in increasing order.
  constexpr int kU = 4;
  for (int64_t k0 = threadIdx.x; k0 < P.n_starts; k0 += kU * kSummaryThreads) {
    double fv[kU], gv[kU];
    int stv[kU], iv[kU], lv[kU], rv[kU], nv[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t k = k0 + int64_t(u) * kSummaryThreads;
      const int64_t s = P.start0 + (k < P.n_starts ? k : P.n_starts - 1);
      fv[u] = out.fval[s];
      stv[u] = out.status[s];
      iv[u] = out.iter[s];
      gv[u] = out.gnorm[s];
      lv[u] = out.ls_evals[s];
      rv[u] = out.grads[s];
      nv[u] = out.near[s];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t k = k0 + int64_t(u) * kSummaryThreads;
      const bool in = k < P.n_starts;
      const int st = stv[u];
      const bool take = in && better(fv[u], k, v, idx);
      v = take ? fv[u] : v;
      idx = take ? k : idx;
      my_status = take ? st : my_status;
      my_iter = take ? iv[u] : my_iter;
      my_gnorm = take ? gv[u] : my_gnorm;
      cnt[0] += in && st == kStatusConverged;
      cnt[1] += in && st == kStatusStopped;
      cnt[2] += in && st != kStatusConverged && st != kStatusStopped;
      cnt[3] += in && st == kStatusNonFinite;
      cnt[4] += in ? iv[u] : 0;
      cnt[5] += in ? lv[u] : 0;
      cnt[6] += in ? rv[u] : 0;
      cnt[7] += in ? nv[u] : 0;
    }
  }
  const int64_t mine = idx; // this thread's own best before the reduction
  const double mine_v = v;
  block_argmin<kSummaryThreads>(v, idx, sv, si); // winner in thread 0
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int64_t t = wave_sum_i64(cnt[c]);
    if ((threadIdx.x & 63) == 0)
      sc[threadIdx.x >> 6][c] = t;
  }
  if (threadIdx.x == 0) {
    // every thread's first candidate is accepted, so with any start at all
    // the winner is a real start (start 0 when every value is NaN)
    const int64_t b = P.n_starts > 0 ? idx : -1;
    ctl[pi].best_idx = int32_t(b);
    sbest = b;
  }
  __syncthreads();
  const int64_t b = sbest;
  // the thread whose own best is the winner holds its fields: written from
  // registers, with no dependent reload of the winner's outputs
  if (b >= 0 && mine == b) {
    ctl[pi].best_status = my_status;
    ctl[pi].best_iter = my_iter;
    ctl[pi].best_fval = mine_v;
    ctl[pi].best_gnorm = my_gnorm;
  }
  if (threadIdx.x < 8) { // one counter per thread, wavefronts in order
    const int c = threadIdx.x;
    int64_t tot = 0;
    for (int w = 0; w < NW; ++w)
      tot += sc[w][c];
    if (c == 0)
      ctl[pi].claimed = int32_t(tot);
    else if (c == 1)
      ctl[pi].stopped = int32_t(tot);
    else if (c == 2)
      ctl[pi].surrendered = int32_t(tot);
    else if (c == 3)
      ctl[pi].nonfinite = int32_t(tot);
    else if (c == 4)
      ctl[pi].sum_iter = tot;
    else if (c == 5)
      ctl[pi].sum_ls_evals = tot;
    else if (c == 6)
      ctl[pi].sum_grads = tot;
    else
      ctl[pi].actual = int32_t(tot);
  }
  if (b >= 0)
    for (int d = threadIdx.x; d < P.dim; d += kSummaryThreads)
      best_x[best_x_off[pi] + d] = out.x[(P.start0 + b) * P.dim + d];
}

// Objective construction from a problem descriptor (built-in objectives).
struct BuiltinProvider {
  template <class F>
  ZEUS_D ZEUS_FI F get(const ProblemDesc& P) const
  {
    F f{};
    if constexpr (F::id == obj::kGaussian) {
      f.C = P.params;
    } else if constexpr (F::id == obj::kLogLikelihood) {
      f.P = P.params;
      f.n = int(P.n_params / 3);
    } else if constexpr (F::id == obj::kMlp) {
      f.P = P.params;
      f.m = int(P.n_params / 2);
    }
    return f;
  }
};

// A user functor uploaded once per call; kernels read it through a pointer
// (no per-launch copy of the functor's members).
template <class UF>
struct UserProvider {
  const UF* f;
  template <class F>
  ZEUS_D ZEUS_FI const UF& get(const ProblemDesc&) const
  {
    return *f;
  }
};

} // namespace zeus::dev
