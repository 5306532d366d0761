// pso.hpp -- particle-swarm seeding, synchronous form.
//
// Restates reference include/pso.cuh:17-155 (init / iteration kernels) and
// 194-345 (launch): positions U(lo, hi], velocities U(-0.1R, 0.1R], then
// PSO_ITER updates v' = w v + c1 r1 (pbest - x) + c2 r2 (gbest - x),
// x' = x + v' (no clamping), pBest on strict improvement.  The result is
// pBestX.  Designed-out hazards of the reference: gBest is frozen during an
// iteration and replaced after it -- only on strict improvement, ties to the
// lowest particle index -- so runs are reproducible and need no host sync.
//
// One launch per iteration.  A launch ends by publishing, per workgroup, its
// best particle (value, index, coordinates) in partial slots kept by
// iteration parity.  The next launch starts with every workgroup of a
// problem folding that problem's partials itself, in one total order, into
// the gBest the iteration uses; the problem's first workgroup also stores
// it, for the fold after.  Visibility between launches is the kernel
// boundary: no device-scope atomics or fences inside the kernels (on a
// multi-die GPU those write back and invalidate the L2 of every die, at
// tens of microseconds each).
//
// Buffers are start-major: X[s * D + d] for global start s of the class.
#pragma once

#include "batch.hpp"
#include "bfgs_group.hpp"
#include "cs_product.hpp"
#include "lane_ops.hpp"
#include "philox.hpp"
#include "prng.hpp"

namespace zeus::dev {

struct PsoArgs {
  const ProblemDesc* probs;
  int32_t n_probs;
  int32_t iter;        // 0 = init, t >= 1 = iteration t
  int32_t n_blocks;    // PSO workgroups of the class (partial slots per parity)
  Control* ctl;
  double* X;
  double* V;
  double* pbx;
  double* pbv;
  double* gbx;         // n_probs x D: gBest after the last iteration (final fold)
  double* part_val;    // [2][n_blocks]: best pBest value of each workgroup
  int64_t* part_idx;   // [2][n_blocks]: its problem-local particle index
  double* part_x;      // [2][n_blocks][D]: its coordinates
  double* gstate;      // n_probs x [2][1 + D]: gBest (value, x) by iteration parity
  const int32_t* block_problem; // workgroup -> problem
  // XORWOW states of the class's (start, coordinate) pairs (prng.hpp), word
  // k at xw[k * xw_stride + s * D + d]; nullptr when no problem uses XORWOW
  uint32_t* xw;
  int64_t xw_stride;
  int32_t keep_rng; // store the states after the last iteration too (PSO phases)
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

// Argmin over the 64 lanes of a wavefront.  `better` is a strict total
// order on (value, index) pairs of distinct indices, so the winner does not
// depend on how lanes are paired: DPP exchanges inside each 16-lane row
// (lane xor 1, xor 2, half-row mirror, row mirror; each step pairs lanes
// that hold the winners of disjoint lane sets), then the gfx950 row swaps
// across rows (register moves, no LDS).  Every lane ends with the winner.
// All lanes must be active.
ZEUS_D ZEUS_FI void
wave_argmin(double& v, int64_t& i)
{
  auto take = [&](double ov, int64_t oi) {
    if (better(ov, oi, v, i)) {
      v = ov;
      i = oi;
    }
  };
  take(dpp_b64<0xB1>(v), dpp_b64<0xB1>(i));   // quad_perm [1,0,3,2]
  take(dpp_b64<0x4E>(v), dpp_b64<0x4E>(i));   // quad_perm [2,3,0,1]
  take(dpp_b64<0x141>(v), dpp_b64<0x141>(i)); // row_half_mirror
  take(dpp_b64<0x140>(v), dpp_b64<0x140>(i)); // row_mirror
  double v0, v1;
  int64_t i0, i1;
  row_swap_b64<16>(v, v0, v1);
  row_swap_b64<16>(i, i0, i1);
  v = v0;
  i = i0;
  take(v1, i1);
  row_swap_b64<32>(v, v0, v1);
  row_swap_b64<32>(i, i0, i1);
  v = v0;
  i = i0;
  take(v1, i1);
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

// gBest after iteration t_prev, computed by every workgroup of the problem
// from the same inputs: the gBest before it (stored during iteration t_prev;
// +inf before iteration 0) and the partials iteration t_prev published.
// Replaced only on strict improvement.  Writes gb[0..D) (LDS) and returns
// the value; with `store` (the problem's first PSO workgroup) the result is
// kept for the next fold.  All threads of the workgroup call.
//
// Two steps: pso_fold_read issues every load that does not depend on the
// winner (this thread's share of the partials, the previous gBest value and,
// when a thread per coordinate is available, its coordinates), so that a
// caller can issue further independent loads before pso_fold_finish; only
// the winner's coordinates are read after the reduction.
struct FoldIn {
  double bv;
  int64_t bi;
  double gpv; // previous gBest value (thread 0)
  double gpx; // previous gBest coordinate threadIdx.x (NT >= D)
};

template <int NT, int PPB, int D>
ZEUS_D ZEUS_FI FoldIn
pso_fold_read(const PsoArgs& a, int pi, const ProblemDesc& P, int t_prev)
{
  const int q = t_prev & 1;
  const int nblk = int((P.n_starts + PPB - 1) / PPB);
  FoldIn f{__builtin_nan(""), int64_t(1) << 62, __builtin_huge_val(), 0.0};
  // this thread's first partial replaces the initial (NaN, 2^62) in any case
  // (NaN loses to a number; two NaNs go to the lower index), so it is only
  // loaded here and compared in pso_fold_finish: nothing waits for these
  // loads before the caller issues its other loads
  int b = threadIdx.x;
  if (b < nblk) {
    const int64_t o = int64_t(q) * a.n_blocks + P.block0 + b;
    f.bv = a.part_val[o];
    f.bi = a.part_idx[o];
  }
  for (b += NT; b < nblk; b += NT) {
    const int64_t o = int64_t(q) * a.n_blocks + P.block0 + b;
    const double v = a.part_val[o];
    const int64_t i = a.part_idx[o];
    if (better(v, i, f.bv, f.bi)) {
      f.bv = v;
      f.bi = i;
    }
  }
  const double* gprev = a.gstate + int64_t(pi) * 2 * (D + 1) + (q ^ 1) * (D + 1);
  if (threadIdx.x == 0 && t_prev > 0)
    f.gpv = gprev[0];
  if constexpr (NT >= D)
    if (int(threadIdx.x) < D && t_prev > 0)
      f.gpx = gprev[1 + threadIdx.x];
  return f;
}

template <int NT, int PPB, int D>
ZEUS_D double
pso_fold_finish(const PsoArgs& a, int pi, const ProblemDesc& P, int t_prev, double* gb,
                bool store, FoldIn f)
{
  double bv = f.bv;
  int64_t bi = f.bi;
  __shared__ double sv[NT / kWave];
  __shared__ int64_t si[NT / kWave];
  __shared__ double s_val;
  __shared__ int s_src; // -1: keep the previous gBest, else winning workgroup
  const int q = t_prev & 1;
  block_argmin<NT>(bv, bi, sv, si);
  double* gs = a.gstate + int64_t(pi) * 2 * (D + 1);
  const double* gprev = gs + (q ^ 1) * (D + 1); // stored during iteration t_prev
  if (threadIdx.x == 0) {
    const double pv = f.gpv; // +inf before iteration 0
    s_src = -1;
    s_val = pv;
    if (bv < pv) {
      s_val = bv;
      s_src = int(bi / PPB);
    }
  }
  __syncthreads();
  if (s_src >= 0) {
    const double* src = a.part_x + (int64_t(q) * a.n_blocks + P.block0 + s_src) * D;
    for (int d = threadIdx.x; d < D; d += NT)
      gb[d] = src[d];
  } else if constexpr (NT >= D) {
    if (int(threadIdx.x) < D)
      gb[threadIdx.x] = f.gpx;
  } else {
    for (int d = threadIdx.x; d < D; d += NT)
      gb[d] = gprev[1 + d];
  }
  const double val = s_val;
  if (store) {
    double* gnew = gs + q * (D + 1);
    if (threadIdx.x == 0)
      gnew[0] = val;
    for (int d = threadIdx.x; d < D; d += NT)
      gnew[1 + d] = gb[d];
  }
  __syncthreads();
  return val;
}

template <int NT, int PPB, int D>
ZEUS_D double
pso_fold(const PsoArgs& a, int pi, const ProblemDesc& P, int t_prev, double* gb, bool store)
{
  return pso_fold_finish<NT, PPB, D>(a, pi, P, t_prev, gb, store,
                                     pso_fold_read<NT, PPB, D>(a, pi, P, t_prev));
}

// Position/velocity of one coordinate: init (t = 0) or update (t >= 1).
// Returns the new position.
// Both uniforms (j = 0, 1) of coordinate d of start s in this iteration
// (draws 2Dt + 2d + j): one Philox block, two Sobol points, or the next two
// draws of the (start, coordinate) XORWOW state, which is advanced.
ZEUS_D ZEUS_FI void
pso_uniforms(const PsoArgs& a, const ProblemDesc& P, int D, int64_t s, int d, double& u0,
             double& u1)
{
  const uint64_t sl = uint64_t(s - P.start0);
  const uint64_t k = rng::pso_draw(D, a.iter, d, 0);
  if (P.prng == rng::kXorwow) {
    const int64_t o = s * D + d;
    rng::Xorwow st = a.iter == 0 ? rng::xorwow_seed(P.seed, P.stream, sl, uint32_t(d))
                                 : xw_load(a.xw, a.xw_stride, o);
    u0 = rng::xorwow_uniform(st);
    u1 = rng::xorwow_uniform(st);
    if (a.iter < P.pso_iter || a.keep_rng)
      xw_store(a.xw, a.xw_stride, o, st);
  } else if (P.prng == rng::kSobol) {
    u0 = rng::sobol_uniform(sl, k);
    u1 = rng::sobol_uniform(sl, k + 1);
  } else {
    rng::uniform_pair(P.seed, P.stream, sl, k >> 1, u0, u1);
  }
}

// The operands of one coordinate's update that do not depend on gBest
// (position, velocity, pBest and both uniforms), loaded ahead of the fold.
struct PsoCoordIn {
  double x, v, pb, u0, u1;
};

ZEUS_D ZEUS_FI PsoCoordIn
pso_coord_load(const PsoArgs& a, const ProblemDesc& P, int D, int64_t s, int d)
{
  PsoCoordIn c{0.0, 0.0, 0.0, 0.0, 0.0};
  pso_uniforms(a, P, D, s, d, c.u0, c.u1);
  if (a.iter > 0) {
    const int64_t o = s * D + d;
    c.x = a.X[o];
    c.v = a.V[o];
    c.pb = a.pbx[o];
  }
  return c;
}

// Initial position and velocity of one coordinate from its two uniforms
// (reference pso.cuh initKernel).
ZEUS_D ZEUS_FI void
pso_init_coord(const ProblemDesc& P, double u0, double u1, double& x, double& v)
{
  const double vr = (P.upper - P.lower) * 0.1;
  x = P.lower + (P.upper + (-P.lower)) * u0;
  v = -vr + (vr + (-(-vr))) * u1;
}

// New velocity of one coordinate (reference pso.cuh iterKernel); the new
// position is xi + the result.
ZEUS_D ZEUS_FI double
pso_velocity(double xi, double vi, double pb, double g, double u0, double u1)
{
  const double r1 = 0.0 + (1.0 + (-0.0)) * u0;
  const double r2 = 0.0 + (1.0 + (-0.0)) * u1;
  return kPsoW * vi + kPsoC1 * r1 * (pb - xi) + kPsoC2 * r2 * (g - xi);
}

// pso_coord from preloaded operands: the same arithmetic, same order.
ZEUS_D ZEUS_FI double
pso_coord_apply(const PsoArgs& a, const ProblemDesc& P, int D, int64_t s, int d,
                const PsoCoordIn& c, const double* gb)
{
  const int64_t o = s * D + d;
  if (a.iter == 0) {
    double rx, rv;
    pso_init_coord(P, c.u0, c.u1, rx, rv);
    a.X[o] = rx;
    a.V[o] = rv;
    a.pbx[o] = rx;
    return rx;
  }
  const double xi = c.x;
  const double nv = pso_velocity(xi, c.v, c.pb, gb[d], c.u0, c.u1);
  a.V[o] = nv;
  a.X[o] = xi + nv;
  return xi + nv;
}

ZEUS_D ZEUS_FI double
pso_coord(const PsoArgs& a, const ProblemDesc& P, int D, int64_t s, int d, const double* gb)
{
  return pso_coord_apply(a, P, D, s, d, pso_coord_load(a, P, D, s, d), gb);
}

// Problem of this workgroup.  A class of one problem (the common case)
// needs no table lookup, so the problem's descriptor is the first load.
ZEUS_D ZEUS_FI int
pso_block_problem(const PsoArgs& a)
{
  return a.n_probs == 1 ? 0 : a.block_problem[blockIdx.x];
}

// Publish the workgroup's best particle into this iteration's partial slot
// (value and index; the caller copies the coordinates).
ZEUS_D ZEUS_FI void
pso_publish_best(const PsoArgs& a, double v, int64_t idx)
{
  const int64_t o = int64_t(a.iter & 1) * a.n_blocks + blockIdx.x;
  a.part_val[o] = v;
  a.part_idx[o] = idx;
}

// Generic form: one lane per particle, objective evaluated by that lane.
template <class F, int D, class Prov>
__global__ __launch_bounds__(256) void
pso_lane_kernel(PsoArgs a, Prov prov)
{
  __shared__ double gb[D];
  __shared__ double sv[4];
  __shared__ int64_t si[4];
  __shared__ int64_t s_best;
  const int pi = pso_block_problem(a);
  const ProblemDesc P = a.probs[pi];
  if (a.iter > P.pso_iter)
    return; // this problem runs fewer iterations (whole workgroup returns)
  if (a.iter > 0)
    (void)pso_fold<256, 256, D>(a, pi, P, a.iter - 1, gb, int(blockIdx.x) == P.block0);
  const int64_t sl = int64_t(blockIdx.x - P.block0) * 256 + threadIdx.x;
  const bool live = sl < P.n_starts;
  const int64_t s = P.start0 + sl;
  double v = __builtin_nan("");
  if (live) {
    decltype(auto) f = prov.template get<F>(P);
    for (int d = 0; d < D; ++d)
      (void)pso_coord(a, P, D, s, d, gb);
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
  int64_t idx = live ? sl : int64_t(1) << 62;
  block_argmin<256>(v, idx, sv, si);
  if (threadIdx.x == 0) {
    pso_publish_best(a, v, idx);
    s_best = idx;
  }
  __syncthreads();
  // the best particle's coordinates: its pBest row, final for this launch
  const int64_t bs = P.start0 + s_best;
  double* px = a.part_x + (int64_t(a.iter & 1) * a.n_blocks + blockIdx.x) * D;
  for (int d = threadIdx.x; d < D; d += 256)
    px[d] = a.pbx[bs * D + d];
}

// Cooperative form: a workgroup owns PPB particles and evaluates each with
// all of its lanes (Eval is a workgroup evaluator whose value() needs only
// the Block reductions, e.g. the MLP loss with one data point per lane).
// Per particle: new coordinates into LDS, value, pBest on strict
// improvement; the workgroup's best pBest is then published like the lane
// form's.
constexpr int kPsoGroupPpb = 2;

template <class F, int D, int NT, int PPB, class Eval, class Prov>
__global__ __launch_bounds__(NT) void
pso_group_kernel(PsoArgs a, Prov prov)
{
  __shared__ double gb[D];
  __shared__ double xs[D];
  __shared__ double red[8 * (NT / kWave)];
  __shared__ double stash[D];
  __shared__ double scalar[4];
  __shared__ int64_t s_best;
  const int pi = pso_block_problem(a);
  const ProblemDesc P = a.probs[pi];
  if (a.iter > P.pso_iter)
    return;
  if (a.iter > 0)
    (void)pso_fold<NT, PPB, D>(a, pi, P, a.iter - 1, gb, int(blockIdx.x) == P.block0);
  Block<NT> blk{red};
  Eval ev;
  ev.sc = GroupScratch{nullptr, stash, scalar};
  ev.bind(P);
  decltype(auto) f = prov.template get<F>(P);
  const int tid = threadIdx.x;
  double bv = __builtin_nan("");
  int64_t bi = int64_t(1) << 62;
  for (int j = 0; j < PPB; ++j) {
    const int64_t sl = int64_t(blockIdx.x - P.block0) * PPB + j;
    if (sl >= P.n_starts)
      break; // uniform across the workgroup
    const int64_t s = P.start0 + sl;
    for (int d = tid; d < D; d += NT)
      xs[d] = pso_coord(a, P, D, s, d, gb);
    const double old = a.iter == 0 ? 0.0 : a.pbv[s];
    __syncthreads();
    const double fv = ev.value(f, xs, blk);
    const bool imp = a.iter == 0 || fv < old;
    if (imp) {
      if (tid == 0)
        a.pbv[s] = fv;
      if (a.iter > 0)
        for (int d = tid; d < D; d += NT)
          a.pbx[s * D + d] = xs[d];
    }
    const double pv = imp ? fv : old;
    if (better(pv, sl, bv, bi)) {
      bv = pv;
      bi = sl;
    }
    __syncthreads(); // xs is rewritten by the next particle
  }
  if (tid == 0) {
    pso_publish_best(a, bv, bi);
    s_best = bi;
  }
  __syncthreads();
  const int64_t bs = P.start0 + s_best;
  double* px = a.part_x + (int64_t(a.iter & 1) * a.n_blocks + blockIdx.x) * D;
  for (int d = tid; d < D; d += NT)
    px[d] = a.pbx[bs * D + d];
}

// Gaussian form: a 1024-thread workgroup owns 16 particles.  Their objective
// values f_i = 0.5 x_i . (Cs x_i) come from the upper block triangle of Cs on
// FP64 MFMA (cs_product.hpp, CsUpperSeg), Xt the 16 positions in LDS; the
// per-segment partial dot products are added in a fixed order.
constexpr int kPsoTile = kCsTile; // particles per workgroup (MFMA N)
constexpr int kPsoTileThreads = 1024;

template <int D>
__global__ __launch_bounds__(kPsoTileThreads) void
pso_gaussian_mfma_kernel(PsoArgs a)
{
  using S = CsShape<D>;
  constexpr int NW = kPsoTileThreads / kWave;
  using Seg = CsUpperSeg<D, CsUpper<D>::template na<NW>()>;
  __shared__ double xs[S::KP * S::XS]; // positions, [dim][particle]
  __shared__ double fpart[Seg::NF * kPsoTile];
  __shared__ double gb[D];
  __shared__ int improved[kPsoTile];
  __shared__ int best_j;

  const int pi = pso_block_problem(a);
  const ProblemDesc P = a.probs[pi];
  if (a.iter > P.pso_iter)
    return;
  const int64_t base = int64_t(blockIdx.x - P.block0) * kPsoTile; // local index
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;

  for (int e = tid; e < S::KP * S::XS; e += kPsoTileThreads)
    xs[e] = 0.0;
  // the fold's partials are read first; then this wavefront's fragments of
  // the product and this thread's coordinate operands and draws (coalesced
  // over d), none of which depends on gBest, so that their latency overlaps
  // the rest of the fold (loads complete in issue order)
  FoldIn fin{};
  if (a.iter > 0)
    fin = pso_fold_read<kPsoTileThreads, kPsoTile, D>(a, pi, P, a.iter - 1);
  Seg seg;
  seg.load(w, P.aux_frag + CsUpper<D>::kFragOffset, li, lk);
  constexpr int kE = (kPsoTile * D + kPsoTileThreads - 1) / kPsoTileThreads;
  PsoCoordIn cin[kE];
#pragma unroll
  for (int k = 0; k < kE; ++k) {
    const int e = tid + k * kPsoTileThreads;
    const int j = e / D, d = e % D;
    if (e < kPsoTile * D && base + j < P.n_starts)
      cin[k] = pso_coord_load(a, P, D, P.start0 + base + j, d);
  }
  double old_pbv = 0.0;
  if (tid < kPsoTile && a.iter > 0 && base + tid < P.n_starts)
    old_pbv = a.pbv[P.start0 + base + tid];
  if (a.iter > 0)
    (void)pso_fold_finish<kPsoTileThreads, kPsoTile, D>(a, pi, P, a.iter - 1, gb,
                                                        int(blockIdx.x) == P.block0, fin);
  __syncthreads();
  // Every load issued so far (fragments, operands) is complete before the
  // coordinate stores are issued: vector memory counts loads and stores
  // together, so a fragment still counted at the first MFMA would make it wait
  // for these stores as well.  The operands are needed here anyway.
  __builtin_amdgcn_s_waitcnt(0x0F70); // vmcnt(0) (expcnt, lgkmcnt untouched)
  // update / init coordinates and stage the tile
#pragma unroll
  for (int k = 0; k < kE; ++k) {
    const int e = tid + k * kPsoTileThreads;
    const int j = e / D, d = e % D;
    if (e < kPsoTile * D && base + j < P.n_starts)
      xs[d * S::XS + j] = pso_coord_apply(a, P, D, P.start0 + base + j, d, cin[k], gb);
  }
  __syncthreads();

  seg.mma(w, xs, fpart, li, lk);
  __syncthreads();

  // f_i, pBest on strict improvement (lanes 0-15 of wavefront 0), then the
  // best particle of the workgroup from the same registers (NaN loses, ties
  // to the lowest index)
  if (w == 0) {
    double v = __builtin_nan("");
    if (tid < kPsoTile) {
      const int64_t sl = base + tid;
      int imp = 0;
      if (sl < P.n_starts) {
        double fv = 0.0; // the upper form carries the factor 0.5
#pragma unroll
        for (int k = 0; k < Seg::NF; ++k)
          fv += fpart[k * kPsoTile + tid];
        const int64_t s = P.start0 + sl;
        if (a.iter == 0) {
          a.pbv[s] = fv;
          v = fv;
        } else {
          v = old_pbv;
          if (fv < v) {
            a.pbv[s] = fv;
            v = fv;
            imp = 1;
          }
        }
      }
      improved[tid] = imp;
    }
    const bool ok = lane < kPsoTile && base + lane < P.n_starts;
    double bv = ok ? v : __builtin_nan("");
    int64_t bi = ok ? base + lane : int64_t(1) << 62;
    wave_argmin(bv, bi);
    if (lane == 0) {
      best_j = int(bi - base);
      pso_publish_best(a, bv, bi);
    }
  }
  __syncthreads();
  // pBest rows of improved particles, and the best particle's coordinates:
  // its pBest is the staged position if it was set now, else its pBest row
  // (not written by this launch)
  double* px = a.part_x + (int64_t(a.iter & 1) * a.n_blocks + blockIdx.x) * D;
  for (int e = tid; e < kPsoTile * D; e += kPsoTileThreads) {
    const int j = e / D, d = e % D;
    const int64_t s = P.start0 + base + j;
    if (base + j < P.n_starts) {
      if (improved[j])
        a.pbx[s * D + d] = xs[d * S::XS + j];
      if (j == best_j)
        px[d] = (improved[j] || a.iter == 0) ? xs[d * S::XS + j] : a.pbx[s * D + d];
    }
  }
}

// Make the fold of iteration t_prev + 1 of problem 0 return the caller's
// gBest (value *gbv, coordinates a.gbx): it is the only partial of
// iteration t_prev (slot 0; the other slots hold NaN, which loses every
// comparison) and also the gBest stored as the one used during iteration
// t_prev, so the fold keeps it whichever of the two it compares (before
// iteration 1 the stored one is not read).  One workgroup.
static __global__ void
pso_seed_kernel(PsoArgs a, int t_prev, const double* gbv, int D)
{
  const int q = t_prev & 1;
  const double v = *gbv;
  for (int b = threadIdx.x; b < a.n_blocks; b += blockDim.x) {
    a.part_val[int64_t(q) * a.n_blocks + b] = b == 0 ? v : __builtin_nan("");
    a.part_idx[int64_t(q) * a.n_blocks + b] = b == 0 ? 0 : int64_t(1) << 62;
  }
  double* g = a.gstate + (q ^ 1) * (D + 1);
  double* px = a.part_x + int64_t(q) * a.n_blocks * D;
  if (threadIdx.x == 0)
    g[0] = v;
  for (int d = threadIdx.x; d < D; d += blockDim.x) {
    g[1 + d] = a.gbx[d];
    px[d] = a.gbx[d];
  }
}

// gBest after a problem's last iteration, for callers that want it: one
// workgroup per problem, writing the gbx row and Control::gbest_val.
template <int NT, int PPB, int D>
__global__ __launch_bounds__(NT) void
pso_final_kernel(PsoArgs a)
{
  __shared__ double gb[D];
  const int pi = blockIdx.x;
  const ProblemDesc P = a.probs[pi];
  if (P.pso_iter < 0 || P.n_starts == 0)
    return;
  const double val = pso_fold<NT, PPB, D>(a, pi, P, P.pso_iter, gb, false);
  for (int d = threadIdx.x; d < D; d += NT)
    a.gbx[int64_t(pi) * D + d] = gb[d];
  if (threadIdx.x == 0)
    a.ctl[pi].gbest_val = val;
}

} // namespace zeus::dev
