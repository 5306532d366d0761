// pso.hpp -- particle-swarm seeding, synchronous form.
//
// Restates reference include/pso.cuh:17-155 (init / iteration kernels) and
// 194-345 (launch): positions U(lo, hi], velocities U(-0.1R, 0.1R], then
// PSO_ITER updates v' = w v + c1 r1 (pbest - x) + c2 r2 (gbest - x),
// x' = x + v' (no clamping), pBest on strict improvement.  The result is
// pBestX.  Designed-out hazards of the reference: gBest is frozen during an
// iteration and replaced after it -- only on strict improvement, ties to the
// lowest particle index -- by the last workgroup of the problem to finish
// (wavefront-shuffle argmin, then LDS across wavefronts, then the
// per-workgroup partials), so runs are reproducible and need no host sync.
//
// Buffers are start-major: X[s * D + d] for global start s of the class.
#pragma once

#include "batch.hpp"
#include "lane_ops.hpp"
#include "cs_product.hpp"
#include "philox.hpp"

namespace zeus::dev {

struct PsoArgs {
  const ProblemDesc* probs;
  int32_t n_probs;
  int32_t iter;        // 0 = init, t >= 1 = iteration t
  Control* ctl;
  double* X;
  double* V;
  double* pbx;
  double* pbv;
  double* gbx;         // n_probs x D
  double* part_val;    // one per workgroup
  int64_t* part_idx;
  const int32_t* block_problem; // workgroup -> problem
};

// (value, index) order for argmin: NaN loses to any number, ties and the
// all-NaN case go to the lowest index.
ZEUS_HD ZEUS_FI bool
better(double a, int64_t ia, double b, int64_t ib)
{
  const bool na = a != a, nb = b != b;
  if (na != nb)
    return nb;
  if (!na && a != b)
    return a < b;
  return ia < ib;
}

ZEUS_D ZEUS_FI void
wave_argmin(double& v, int64_t& i)
{
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, kWave);
    const int64_t oi = __shfl_xor(i, off, kWave);
    if (better(ov, oi, v, i)) {
      v = ov;
      i = oi;
    }
  }
}

// Workgroup argmin; result valid in thread 0.
template <int NT>
ZEUS_D ZEUS_FI void
block_argmin(double& v, int64_t& i, double* sv, int64_t* si)
{
  wave_argmin(v, i);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    sv[w] = v;
    si[w] = i;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < NT / kWave; ++k)
      if (better(sv[k], si[k], v, i)) {
        v = sv[k];
        i = si[k];
      }
  }
}

// After the workgroup's particles are final: publish the workgroup's best,
// and let the last workgroup of the problem fold the partials into gBest.
template <int NT, int PPB, int D>
ZEUS_D void
publish_best(const PsoArgs& a, int pi, const ProblemDesc& P, double v, int64_t idx)
{
  __shared__ double sv[NT / kWave];
  __shared__ int64_t si[NT / kWave];
  __shared__ int last;
  block_argmin<NT>(v, idx, sv, si);
  const int nblk = int((P.n_starts + PPB - 1) / PPB); // PPB particles per workgroup
  if (threadIdx.x == 0) {
    a.part_val[blockIdx.x] = v;
    a.part_idx[blockIdx.x] = idx;
    __threadfence();
    const uint32_t t = __hip_atomic_fetch_add(&a.ctl[pi].pso_arrive, 1u, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_AGENT);
    last = (t == uint32_t(nblk - 1));
  }
  __syncthreads();
  if (!last)
    return;
  __threadfence();
  // fold the partials with the whole workgroup; `better` is a total order,
  // so the result does not depend on the folding order
  double bv = __builtin_nan("");
  int64_t bi = int64_t(1) << 62;
  for (int b = threadIdx.x; b < nblk; b += NT) {
    const double ov = a.part_val[P.block0 + b];
    const int64_t oi = a.part_idx[P.block0 + b];
    if (better(ov, oi, bv, bi)) {
      bv = ov;
      bi = oi;
    }
  }
  __syncthreads(); // sv/si are reused by the second argmin
  block_argmin<NT>(bv, bi, sv, si);
  if (threadIdx.x == 0) {
    a.ctl[pi].pso_arrive = 0u;
    if (bv < a.ctl[pi].gbest_val) {
      a.ctl[pi].gbest_val = bv;
      sv[0] = 1.0;
      si[0] = bi;
    } else {
      sv[0] = 0.0;
    }
  }
  __syncthreads();
  if (sv[0] != 0.0) {
    const int64_t src = P.start0 + si[0]; // winner's local index -> global row
    for (int d = threadIdx.x; d < D; d += NT)
      a.gbx[int64_t(pi) * D + d] = a.pbx[src * D + d];
  }
}

// Position/velocity of one coordinate: init (t = 0) or update (t >= 1).
ZEUS_D ZEUS_FI void
pso_coord(const PsoArgs& a, const ProblemDesc& P, int pi, int D, int64_t s, int d)
{
  const int64_t sl = s - P.start0;
  const int64_t o = s * D + d;
  if (a.iter == 0) {
    const double vr = (P.upper - P.lower) * 0.1;
    const double rx = rng::uniform_in(P.seed, P.stream, uint64_t(sl),
                                      rng::pso_draw(D, 0, d, 0), P.lower, P.upper);
    const double rv =
      rng::uniform_in(P.seed, P.stream, uint64_t(sl), rng::pso_draw(D, 0, d, 1), -vr, vr);
    a.X[o] = rx;
    a.V[o] = rv;
    a.pbx[o] = rx;
  } else {
    const double r1 = rng::uniform_in(P.seed, P.stream, uint64_t(sl),
                                      rng::pso_draw(D, a.iter, d, 0), 0.0, 1.0);
    const double r2 = rng::uniform_in(P.seed, P.stream, uint64_t(sl),
                                      rng::pso_draw(D, a.iter, d, 1), 0.0, 1.0);
    const double xi = a.X[o], vi = a.V[o], pb = a.pbx[o];
    const double gb = a.gbx[int64_t(pi) * D + d];
    const double nv = kPsoW * vi + kPsoC1 * r1 * (pb - xi) + kPsoC2 * r2 * (gb - xi);
    a.V[o] = nv;
    a.X[o] = xi + nv;
  }
}

// Generic form: one lane per particle, objective evaluated by that lane.
template <class F, int D, class Prov>
__global__ __launch_bounds__(256) void
pso_lane_kernel(PsoArgs a, Prov prov)
{
  const int pi = a.block_problem[blockIdx.x];
  const ProblemDesc P = a.probs[pi];
  if (a.iter > P.pso_iter)
    return; // this problem runs fewer iterations (whole problem returns)
  const int64_t sl = int64_t(blockIdx.x - P.block0) * 256 + threadIdx.x;
  const bool live = sl < P.n_starts;
  const int64_t s = P.start0 + sl;
  double v = __builtin_nan("");
  if (live) {
    decltype(auto) f = prov.template get<F>(P);
    for (int d = 0; d < D; ++d)
      pso_coord(a, P, pi, D, s, d);
    const double fv = lane_value<F, D>(f, a.X + s * D);
    if (a.iter == 0) {
      a.pbv[s] = fv;
    } else if (fv < a.pbv[s]) {
      a.pbv[s] = fv;
      for (int d = 0; d < D; ++d)
        a.pbx[s * D + d] = a.X[s * D + d];
    }
    v = a.pbv[s];
  }
  publish_best<256, 256, D>(a, pi, P, v, live ? sl : int64_t(1) << 62);
}

// Gaussian form: a 1024-thread workgroup owns 16 particles.  Their objective
// values come from G = Cs Xt on FP64 MFMA (cs_product.hpp), Xt the 16
// positions in LDS; f_i = 0.5 x_i . (Cs x_i), the per-unit partial dot
// products added in a fixed order.
constexpr int kPsoTile = kCsTile; // particles per workgroup (MFMA N)
constexpr int kPsoTileThreads = 1024;

template <int D>
__global__ __launch_bounds__(kPsoTileThreads) void
pso_gaussian_mfma_kernel(PsoArgs a)
{
  using S = CsShape<D>;
  constexpr int NW = kPsoTileThreads / kWave;
  __shared__ double xs[S::KP * S::XS]; // positions, [dim][particle]
  __shared__ double fpart[S::U * kPsoTile];
  __shared__ int improved[kPsoTile];

  const int pi = a.block_problem[blockIdx.x];
  const ProblemDesc P = a.probs[pi];
  if (a.iter > P.pso_iter)
    return;
  const int64_t base = int64_t(blockIdx.x - P.block0) * kPsoTile; // local index
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;

  // update / init coordinates (coalesced over d) and stage the tile
  for (int e = tid; e < S::KP * S::XS; e += kPsoTileThreads)
    xs[e] = 0.0;
  __syncthreads();
#pragma unroll
  for (int e = tid; e < kPsoTile * D; e += kPsoTileThreads) {
    const int j = e / D, d = e % D;
    const int64_t sl = base + j;
    if (sl < P.n_starts) {
      const int64_t s = P.start0 + sl;
      pso_coord(a, P, pi, D, s, d);
      xs[d * S::XS + j] = a.X[s * D + d];
    }
  }
  __syncthreads();

  const double* Cf = P.aux_frag;
  for (int u = w; u < S::U; u += NW) {
    const mfma_acc_t acc = cs_unit<D>(Cf, xs, u, li, lk);
    double fp = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = 16 * (u >> 1) + 4 * q + lk;
      if (rr < D)
        fp = ::fma(xs[rr * S::XS + li], acc[q], fp);
    }
    fp += __shfl_xor(fp, 16, kWave);
    fp += __shfl_xor(fp, 32, kWave);
    if (lk == 0)
      fpart[u * kPsoTile + li] = fp;
  }
  __syncthreads();

  // f_i, pBest on strict improvement
  double v = __builtin_nan("");
  int64_t vi = int64_t(1) << 62;
  if (tid < kPsoTile) {
    const int64_t sl = base + tid;
    int imp = 0;
    if (sl < P.n_starts) {
      double t = 0.0;
      for (int u = 0; u < S::U; ++u)
        t += fpart[u * kPsoTile + tid];
      const double fv = 0.5 * t;
      const int64_t s = P.start0 + sl;
      if (a.iter == 0) {
        a.pbv[s] = fv;
      } else if (fv < a.pbv[s]) {
        a.pbv[s] = fv;
        imp = 1;
      }
      v = a.pbv[s];
      vi = sl;
    }
    improved[tid] = imp;
  }
  __syncthreads();
  for (int e = tid; e < kPsoTile * D; e += kPsoTileThreads) {
    const int j = e / D, d = e % D;
    if (improved[j])
      a.pbx[(P.start0 + base + j) * D + d] = xs[d * S::XS + j];
  }
  publish_best<kPsoTileThreads, kPsoTile, D>(a, pi, P, v, vi);
}

} // namespace zeus::dev
