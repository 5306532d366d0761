// bfgs_tile.hpp -- tile form of the BFGS for the Gaussian objective: one
// 1024-thread workgroup advances a tile of 16 starts of one problem in
// lock-step rounds.
//
// Round structure.  Every start of the tile submits one point per round (its
// first point, a line-search trial x + alpha p, or the unconditional point
// after 20 rejected trials).  The 16 points Xt (D x 16, LDS) are evaluated
// together: G = Cs Xt on FP64 MFMA (v_mfma_f64_16x16x4f64), Cs = (C + C^T)/2
// as A operands, f_j = 0.5 Xt_j . G_j, and G_j is the adjoint gradient of
// start j at its point.  The product's MB x KB steps are cut into 16
// balanced segments, one per wavefront (cs_product.hpp, CsFullSeg), whose
// Cs fragments stay in registers for the whole tile (Cs does not change);
// each segment leaves partial blocks of G in LDS, and row d of G is the sum
// of the blocks of its row block in segment order.  Wavefront w then runs
// start w's BFGS logic for the round -- line-search decision, update,
// convergence test -- with DPP wavefront reductions only, and submits its
// next point.  Two workgroup barriers per round.
//
// Method.  The reference's per-start BFGS (include/bfgs_sequential.cuh:
// 46-221, utils.cuh:245-287, 376-426) with the same line search, update
// skip rule, statuses and counters as the other forms.  The inverse Hessian
// is kept as H = B + sum_j (s_j a_j^T + a_j s_j^T): the rank-2 pairs of the
// update H' = H + s a^T + a s^T (a = (beta/2) s - rho u, u = H y), which is
// the reference's update expanded; B = I until the pair list is full, then
// the pairs are folded into an explicit B (global memory) and the list
// restarts.  H g needs 2k dot products and 2k axpys instead of a pass over
// D^2 entries.  Sums run in a fixed order (results are deterministic), but
// in another order than the serial oracle: parity is to tolerance.
#pragma once

#include "batch.hpp"
#include "block.hpp"
#include "cs_product.hpp"
#include "philox.hpp"
#include "prng.hpp"

namespace zeus::dev {

constexpr int kTileStarts = kCsTile; // starts per tile = MFMA N dimension
// Workgroup size of the tile form: 16 wavefronts, one per start of the
// tile (a 128-register budget per lane).
constexpr int kTileThreads = 1024;
constexpr int kTileWaves = kTileThreads / kWave;
static_assert(kTileStarts % kTileWaves == 0, "whole slots per wavefront");

template <int D>
struct TileShape : CsShape<D> {
  static constexpr int EL = (D + kWave - 1) / kWave; // rows per lane in start logic
  static constexpr int kHist = 32; // rank-2 pairs kept before folding into B
  static constexpr int64_t kSlotDoubles = int64_t(kHist) * 2 * D + int64_t(D) * D;
};

struct TileArgs {
  const ProblemDesc* probs;
  const int32_t* tile_problem; // tile -> problem (tiles of a problem are consecutive)
  int64_t n_tiles;
  Control* ctl;
  StartOut out;
  double* slot_mem; // per (workgroup, start slot): kHist pairs, then B (D x D)
};

// Per-start state of a slot (LDS; read and written by its wavefront only).
struct TileSlot {
  int64_t s;   // global start index
  int phase;   // 0 idle, 1 first point, 2 trial, 3 unconditional point
  int trial, iter, grads, ls_evals, k;
  int base;    // explicit B in use
  double best, alpha, ddir;
};

// The product's segments: 16, one per wavefront, every fragment of a
// segment held in registers across the rounds of a tile (kTileKeep >= the
// longest segment; 24 fragments = 48 VGPRs per lane at D = 150).
constexpr int kTileKeep = 64;
constexpr int kHgChunk = 2; // rank-2 pairs read together in H g
template <int D>
using TileSeg = CsFullSeg<D, kTileWaves, kTileKeep>;

template <int D>
struct TileLds {
  using S = TileShape<D>;
  using Seg = TileSeg<D>;
  double xt[S::KP * S::XS]; // submitted points, [row][slot]
  double gpart[Seg::NP * Seg::kSlot]; // partial blocks of Cs Xt, [part][row % 16][slot]
  double fpart[Seg::NP * kTileStarts];
  double x[kTileStarts * D], g[kTileStarts * D], p[kTileStarts * D];
  // one byte per slot, so that a wavefront reads all 16 flags in one access
  alignas(16) uint8_t live[kTileStarts];
  alignas(16) uint8_t newconv[kTileStarts]; // slot converged in the last round, not yet counted
  // the problem's converged count as read in a round (the reading
  // wavefronts' counters summed), by round parity: one is summed while the
  // other, last read in the round before, is cleared
  int conv_sum[2];
  TileSlot slot[kTileStarts];
  ProblemDesc prob; // the tile's problem
};

// Row d of G for slot j: the partial blocks of the segments covering d's row
// block (consecutive parts), added in segment order.  Every part is read at
// once (past the block's count the first part again, not added).
template <int D>
ZEUS_D ZEUS_FI double
tile_grad(const TileLds<D>& L, int d, int j)
{
  using Seg = TileSeg<D>;
  using S = TileShape<D>;
  const int mb = (d >> 4) < S::MB - 1 ? (d >> 4) : S::MB - 1;
  const int p0 = int((Seg::kFirstBits >> (6 * mb)) & 63);
  const int n = int((Seg::kCountBits >> (4 * mb)) & 15);
  const double* gp = L.gpart + p0 * Seg::kSlot + (d & 15) * S::XS + j;
  double v[Seg::kCover];
#pragma unroll
  for (int c = 0; c < Seg::kCover; ++c)
    v[c] = gp[(c < n ? c : 0) * Seg::kSlot];
  double g = v[0];
#pragma unroll
  for (int c = 1; c < Seg::kCover; ++c)
    if (c < n)
      g += v[c];
  return g;
}

template <int EL>
ZEUS_D ZEUS_FI double
tile_dot(const double (&a)[EL], const double (&b)[EL])
{
  double t = 0.0;
#pragma unroll
  for (int e = 0; e < EL; ++e)
    t = ::fma(a[e], b[e], t);
  return wave_sum(t);
}

// Submits slot w's next point x + alpha p into the product's operand Xt.
// Every row is read before the first store (the compiler cannot tell the LDS
// rows apart); full rows need no mask (lane < 64).
template <int D>
ZEUS_D ZEUS_FI void
tile_submit(TileLds<D>& L, const int w, int lane, const double alpha)
{
  using S = TileShape<D>;
  constexpr int EL = S::EL;
  // opaque lane: the row addresses are recomputed here, not held (or
  // spilled) across the round loop
  asm volatile("" : "+v"(lane));
  const double* X = L.x + w * D;
  const double* Pd = L.p + w * D;
  double x[EL], p[EL];
#pragma unroll
  for (int e = 0; e < EL; ++e)
    if ((e + 1) * kWave <= D || lane + e * kWave < D) {
      x[e] = X[lane + e * kWave];
      p[e] = Pd[lane + e * kWave];
    }
#pragma unroll
  for (int e = 0; e < EL; ++e)
    if ((e + 1) * kWave <= D || lane + e * kWave < D)
      L.xt[(lane + e * kWave) * S::XS + w] = ::fma(alpha, p[e], x[e]);
}

// Armijo acceptance of a line-search trial with value f at step alpha
// (reference utils.cuh:376-426).
ZEUS_D ZEUS_FI bool
tile_accepts(double f, double alpha, double ddir, double best)
{
  return f <= ::fma(kArmijoC1 * alpha, ddir, best) && f < best;
}

// Logic of start slot w for one round, run by one wavefront; `f` is the
// value at the slot's submitted point, `stop` the problem's early stop as
// seen this round.
template <int D>
ZEUS_D void
tile_slot_step(TileLds<D>& L, TileSlot& st, const int w, const ProblemDesc& P,
               const StartOut& out, double* mem, double f, const bool stop)
{
  using S = TileShape<D>;
  constexpr int EL = S::EL;
  // the lane index is made opaque here, so that the per-lane addresses below
  // are recomputed each round instead of being held (or spilled) across the
  // whole round loop
  int lane = threadIdx.x % kWave;
  asm volatile("" : "+v"(lane));
  double* X = L.x + w * D;
  double* G = L.g + w * D;
  double* Pd = L.p + w * D;
  double* hist = mem;
  double* B = mem + int64_t(S::kHist) * 2 * D;
  int row[EL];
  bool ok[EL];
#pragma unroll
  for (int e = 0; e < EL; ++e) {
    row[e] = lane + e * kWave;
    ok[e] = row[e] < D;
  }
  // rows every lane has (d < 64 (e + 1) <= D) need no mask; the lane index
  // is opaque to the compiler, so it cannot tell by itself
  auto okr = [&](int e) { return (e + 1) * kWave <= D || ok[e]; };
  auto load = [&](const double* v, double (&r)[EL]) {
#pragma unroll
    for (int e = 0; e < EL; ++e)
      r[e] = okr(e) ? v[row[e]] : 0.0;
  };
  auto load_grad = [&](double (&r)[EL]) { // G column from its partial blocks
#pragma unroll
    for (int e = 0; e < EL; ++e)
      r[e] = okr(e) ? tile_grad<D>(L, row[e], w) : 0.0;
  };
  auto submit = [&](int phase) {
    tile_submit<D>(L, w, lane, st.alpha);
    st.phase = phase;
  };
  auto finish = [&](int status, double fval, double gnorm) {
    double e2 = 0.0;
    if (status == kStatusConverged) {
      double x[EL], d[EL];
      load(X, x);
#pragma unroll
      for (int e = 0; e < EL; ++e)
        d[e] = ok[e] ? x[e] - known_min(P.min_kind, row[e]) : 0.0;
      e2 = tile_dot<EL>(d, d);
    }
    if (lane == 0) {
      out.status[st.s] = status;
      out.iter[st.s] = st.iter;
      out.fval[st.s] = fval;
      out.gnorm[st.s] = gnorm;
      out.ls_evals[st.s] = st.ls_evals;
      out.grads[st.s] = st.grads;
      out.near[st.s] = status == kStatusConverged && ::sqrt(e2) < 0.5;
    }
#pragma unroll
    for (int e = 0; e < EL; ++e)
      if (ok[e])
        out.x[st.s * D + row[e]] = X[row[e]];
    st.phase = 0;
  };
  auto gnorm_now = [&]() {
    double g[EL];
    load(G, g);
    return ::sqrt(tile_dot<EL>(g, g));
  };
  // top of a BFGS iteration: MAX_ITER, stop flag, first trial
  auto begin_iteration = [&]() {
    if (st.iter >= P.max_iter) {
      finish(kStatusSurrendered, st.best, gnorm_now());
      return;
    }
    if (stop) {
      finish(kStatusStopped, st.best, gnorm_now());
      return;
    }
    double g[EL], p[EL];
    load(G, g);
    load(Pd, p);
    st.ddir = tile_dot<EL>(g, p);
    st.alpha = 1.0;
    st.trial = 0;
    submit(2);
  };

  if (st.phase == 1) { // first point: value and gradient
    double g[EL];
    load_grad(g);
#pragma unroll
    for (int e = 0; e < EL; ++e)
      if (ok[e]) {
        G[row[e]] = g[e];
        Pd[row[e]] = -g[e];
      }
    st.best = f;
    st.grads = 1;
    begin_iteration();
    return;
  }
  if (st.phase == 2) {
    ++st.ls_evals;
    const bool accept = tile_accepts(f, st.alpha, st.ddir, st.best);
    if (!accept) {
      st.alpha *= 0.5;
      if (++st.trial == kLineSearchTrials)
        submit(3);
      else
        submit(2);
      return;
    }
  } else { // phase 3: point after the last rejected trial
    ++st.ls_evals;
  }

  // ---- end of the iteration: gradient at the new point and the update ----
  // (short live ranges: vectors are re-read from LDS where they are needed)
  const double fnew = f;
  ++st.grads;
  const bool acc = fnew < st.best;
  double gn[EL], v[EL];
  load_grad(gn);
  // v = H gn, H = B + sum_j (s_j a_j^T + a_j s_j^T)
  if (st.base) {
#pragma unroll
    for (int e = 0; e < EL; ++e)
      v[e] = 0.0;
#pragma unroll 1
    for (int c = 0; c < D; ++c) {
      const double gc = tile_grad<D>(L, c, w);
#pragma unroll
      for (int e = 0; e < EL; ++e)
        if (ok[e])
          v[e] = ::fma(B[int64_t(c) * D + row[e]], gc, v[e]);
    }
  } else {
#pragma unroll
    for (int e = 0; e < EL; ++e)
      v[e] = gn[e];
  }
  // pairs in chunks of kHgChunk: every load of a chunk is issued before its
  // dot products (one memory latency per chunk; the pairs of the whole chip
  // are read in the same rounds, mostly beyond the L2), the chunk's 2
  // kHgChunk wavefront sums run side by side, then v accumulates in order j
  constexpr int kC = kHgChunk;
  // no LDS access is moved across the pair loop (the values it would keep
  // live through the loop cost registers the chunk needs)
  asm volatile("" ::: "memory");
#pragma unroll 1
  for (int j0 = 0; j0 < st.k; j0 += kC) {
    double sj[kC][EL], aj[kC][EL];
#pragma unroll
    for (int c = 0; c < kC; ++c) {
      // past the list's end the last pair is loaded again; its terms are not
      // used.  Rows past D read the next vector of the slot (in bounds: the
      // explicit B follows the list) and are zeroed after the load, so the
      // loads are unconditional, addressed by a uniform base and the lane.
      const int j = j0 + c < st.k ? j0 + c : st.k - 1;
      const gdouble_ptr pj = as_global(hist) + int64_t(j) * 2 * D + lane;
#pragma unroll
      for (int e = 0; e < EL; ++e) {
        sj[c][e] = pj[e * kWave];
        aj[c][e] = pj[D + e * kWave];
      }
    }
#pragma unroll
    for (int c = 0; c < kC; ++c)
#pragma unroll
      for (int e = 0; e < EL; ++e) {
        sj[c][e] = okr(e) ? sj[c][e] : 0.0;
        aj[c][e] = okr(e) ? aj[c][e] : 0.0;
      }
    double cj[kC], dj[kC];
#pragma unroll
    for (int c = 0; c < kC; ++c) {
      double t0 = 0.0, t1 = 0.0;
#pragma unroll
      for (int e = 0; e < EL; ++e) {
        t0 = ::fma(aj[c][e], gn[e], t0);
        t1 = ::fma(sj[c][e], gn[e], t1);
      }
      cj[c] = t0;
      dj[c] = t1;
    }
#pragma unroll
    for (int c = 0; c < kC; ++c) {
      cj[c] = wave_sum(cj[c]);
      dj[c] = wave_sum(dj[c]);
    }
#pragma unroll
    for (int c = 0; c < kC; ++c)
      if (j0 + c < st.k)
#pragma unroll
        for (int e = 0; e < EL; ++e)
          v[e] = ::fma(sj[c][e], cj[c], ::fma(aj[c][e], dj[c], v[e]));
  }
  asm volatile("" ::: "memory");
  // s = xt - x, y = gn - g, u = H y = H gn - H g = v + p; and |g'|^2 for
  // the gradient g' after the iteration (gn if accepted, else g), summed as
  // the norm of g' would be once it is stored
  double s[EL];
  double sTy, yu, gg;
  {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
    for (int e = 0; e < EL; ++e) {
      const double xv = okr(e) ? X[row[e]] : 0.0;
      const double xtv = okr(e) ? L.xt[row[e] * S::XS + w] : 0.0;
      const double gv = okr(e) ? G[row[e]] : 0.0;
      const double pv = okr(e) ? Pd[row[e]] : 0.0;
      s[e] = xtv - xv;
      const double y = gn[e] - gv;
      t0 = ::fma(s[e], y, t0);
      t1 = ::fma(y, v[e] + pv, t1);
      const double ga = acc ? gn[e] : gv;
      t2 = ::fma(ga, ga, t2);
    }
    sTy = wave_sum(t0);
    yu = wave_sum(t1);
    gg = wave_sum(t2);
  }
  double dd = 0.0; // g'.p' (the next iteration's directional derivative), in row order
  const bool upd = !(::fabs(sTy) < kCurvatureSkip);
  if (upd) {
    const double rho = 1.0 / sTy;
    const double beta = ::fma(rho * rho, yu, rho);
    const double hb = 0.5 * beta;
    double a[EL];
    double q0, q1;
    {
      double t0 = 0.0, t1 = 0.0;
#pragma unroll
      for (int e = 0; e < EL; ++e) {
        const double pv = okr(e) ? Pd[row[e]] : 0.0;
        a[e] = okr(e) ? ::fma(hb, s[e], -(rho * (v[e] + pv))) : 0.0;
        const double ga = acc ? gn[e] : (okr(e) ? G[row[e]] : 0.0);
        t0 = ::fma(a[e], ga, t0);
        t1 = ::fma(s[e], ga, t1);
      }
      q0 = wave_sum(t0);
      q1 = wave_sum(t1);
    }
    // p' = -H' g = -(H g + s (a.g) + a (s.g)), H g = v (accepted) or -p
#pragma unroll
    for (int e = 0; e < EL; ++e) {
      const double ga = acc ? gn[e] : (okr(e) ? G[row[e]] : 0.0);
      double pn = 0.0;
      if (okr(e)) {
        const double hg = acc ? v[e] : -Pd[row[e]];
        pn = -::fma(a[e], q1, ::fma(s[e], q0, hg));
        Pd[row[e]] = pn;
      }
      dd = ::fma(ga, pn, dd);
    }
    // append the pair; a full list is folded into B first
    if (st.k == S::kHist) {
#pragma unroll 1
      for (int c = 0; c < D; ++c) {
        double bc[EL];
#pragma unroll
        for (int e = 0; e < EL; ++e)
          bc[e] = st.base ? (ok[e] ? B[int64_t(c) * D + row[e]] : 0.0)
                          : (row[e] == c ? 1.0 : 0.0);
#pragma unroll 1
        for (int j = 0; j < S::kHist; ++j) {
          const double* sj = hist + int64_t(j) * 2 * D;
          const double* aj = sj + D;
          const double sc = sj[c], ac = aj[c];
#pragma unroll
          for (int e = 0; e < EL; ++e)
            if (ok[e])
              bc[e] = bc[e] + (sj[row[e]] * ac + aj[row[e]] * sc);
        }
#pragma unroll
        for (int e = 0; e < EL; ++e)
          if (ok[e])
            B[int64_t(c) * D + row[e]] = bc[e];
      }
      st.base = 1;
      st.k = 0;
    }
    double* sk = hist + int64_t(st.k) * 2 * D;
#pragma unroll
    for (int e = 0; e < EL; ++e)
      if (okr(e)) {
        sk[row[e]] = s[e];
        sk[D + row[e]] = a[e];
      }
    ++st.k;
  } else if (acc) {
#pragma unroll
    for (int e = 0; e < EL; ++e) {
      const double pn = okr(e) ? -v[e] : 0.0;
      if (okr(e))
        Pd[row[e]] = pn;
      dd = ::fma(gn[e], pn, dd);
    }
  } else {
#pragma unroll
    for (int e = 0; e < EL; ++e)
      dd = ::fma(okr(e) ? G[row[e]] : 0.0, okr(e) ? Pd[row[e]] : 0.0, dd);
  }
  if (acc) {
    st.best = fnew;
#pragma unroll
    for (int e = 0; e < EL; ++e)
      if (okr(e)) {
        X[row[e]] = L.xt[row[e] * S::XS + w];
        G[row[e]] = gn[e];
      }
  }
  const double nrm = ::sqrt(gg);
  const double ddir = wave_sum(dd);
  if (!::isfinite(nrm) || !::isfinite(fnew)) {
    finish(kStatusNonFinite, fnew, nrm);
    return;
  }
  if (nrm < P.tol) {
    if (lane == 0)
      L.newconv[w] = 1; // counted for the tile at the top of the next round
    finish(kStatusConverged, st.best, nrm);
    return;
  }
  // top of the next iteration (begin_iteration with the norm and the
  // directional derivative already at hand)
  ++st.iter;
  if (st.iter >= P.max_iter) {
    finish(kStatusSurrendered, st.best, nrm);
    return;
  }
  if (stop) {
    finish(kStatusStopped, st.best, nrm);
    return;
  }
  st.ddir = ddir;
  st.alpha = 1.0;
  st.trial = 0;
  submit(2);
}

constexpr int kStopReader = kTileThreads - kWave; // thread that adds the tile's count

// Adds the slots flagged in newconv to the workgroup's counter of the
// problem's converged count (one relaxed atomic, none when no slot
// converged) and clears the flags.
template <int D>
ZEUS_D ZEUS_FI void
tile_count_converged(TileLds<D>& L, Control* ctl)
{
  static_assert(kTileStarts == 16, "the flags are one 16-byte word");
  uint4* fl = reinterpret_cast<uint4*>(L.newconv);
  const uint4 v = *fl;
  *fl = uint4{0u, 0u, 0u, 0u};
  // bytes are 0 or 1: the byte sum of the four words' sum is the count
  const unsigned b = v.x + v.y + v.z + v.w;
  const int n = int((b * 0x01010101u) >> 24);
  if (n > 0)
    (void)__hip_atomic_fetch_add(&ctl->conv_lines[(blockIdx.x % kConvLines) * kConvStride], n,
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int D>
__global__ __launch_bounds__(kTileThreads) void
bfgs_tile_kernel(TileArgs a)
{
  using S = TileShape<D>;
  __shared__ TileLds<D> L;
  const int tid = threadIdx.x, lane = tid % kWave;
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int li0 = lane & 15, lk0 = lane >> 4;
  auto slot_mem = [&](int j) {
    return a.slot_mem + (int64_t(blockIdx.x) * kTileStarts + j) * S::kSlotDoubles;
  };

  using Seg = TileSeg<D>;
  for (int e = tid; e < S::KP * S::XS; e += kTileThreads)
    L.xt[e] = 0.0; // padding rows stay zero
  Seg seg;

  for (int64_t tile = blockIdx.x; tile < a.n_tiles; tile += gridDim.x) {
    const int pi = wave_uniform(a.tile_problem[tile]);
    if (tid == 0)
      L.prob = a.probs[pi];
    __syncthreads();
    const ProblemDesc& P = L.prob;
    const double* Cf = P.aux_frag;
    // slots w, w + kTileWaves, ...: every lane of the wavefront writes the
    // same values.  A problem's tiles are consecutive from
    // ProblemDesc::block0 (its PSO workgroups, which hold the same starts).
    for (int j = w; j < kTileStarts; j += kTileWaves) {
      TileSlot& st = L.slot[j];
      st = TileSlot{-1, 0, 0, 0, 0, 0, 0, 0, 0.0, 1.0, 0.0};
      L.newconv[j] = 0;
      const int64_t sl = (tile - P.block0) * kTileStarts + j; // local start
      if (sl < P.n_starts) {
        st.s = P.start0 + sl;
        double* X = L.x + j * D;
        for (int d = lane; d < D; d += kWave) {
          const double v =
            P.x0 ? P.x0[sl * D + d]
                 : start_uniform_in(P, uint64_t(sl), d);
          X[d] = v;
          L.xt[d * S::XS + j] = v;
        }
        st.phase = 1;
      }
    }
    Control* ctl = a.ctl + pi;
    if constexpr (Seg::kKeep > 0) // this problem's fragments kept across rounds
      seg.load(w, Cf, li0, lk0, 0, Seg::kKeep);
    if (tid < 2)
      L.conv_sum[tid] = 0;
    __syncthreads();

    int par = 0; // this round's conv_sum (the same in every wavefront)
    for (;; par ^= 1) {
      // Early stop (reference bfgs_common.cuh:170-208): the problem stops
      // once `required` starts have converged; a start sees it at the top of
      // its next iteration, as in the other forms.
      // per-lane indices made opaque each round: addresses derived from them
      // are recomputed here rather than held across the round loop
      int li = li0, lk = lk0;
      asm volatile("" : "+v"(li), "+v"(lk));
      // The starts that converged in the last round are added to one of the
      // problem's eight counters here, one atomic per tile and round (one
      // per start, all on one address, had queued up behind each other in
      // the rounds where most starts of the chip converge together).
      // Wavefronts 8-15 each read one counter a third of the way through
      // their product, when the other tiles' sums of the last round have
      // been added (the load's latency hides behind the rest of the
      // product), and add it to this round's LDS sum.
      if (tid == kStopReader) {
        tile_count_converged(L, ctl);
        L.conv_sum[par ^ 1] = 0; // last read in the round before
      }
      constexpr int kReader0 = kTileWaves - kConvLines;
      int conv = 0;
      auto read_count = [&]() {
        if (w >= kReader0)
          conv = __hip_atomic_load(&ctl->conv_lines[(w - kReader0) * kConvStride],
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      };
      // ---- G = Cs Xt in balanced segments, partial blocks and values ----
      if constexpr (Seg::kKeep < Seg::LMAX)
        seg.load(w, Cf, li, lk, Seg::kKeep, Seg::LMAX);
      __builtin_amdgcn_sched_barrier(0);
      seg.mma(w, L.xt, L.gpart, L.fpart, li, lk, read_count);
      if (w >= kReader0 && lane == 0)
        (void)__hip_atomic_fetch_add(&L.conv_sum[par], conv, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
      __syncthreads();
      // ---- start logic: wavefront w runs its slots one after the other ----
      for (int j = w; j < kTileStarts; j += kTileWaves) {
        TileSlot& sl = L.slot[j];
        // what a rejected line-search trial needs, read together with the
        // value's partials: most rounds end that way, and then the slot
        // takes the backtracking step here (tile_slot_step's, the same
        // arithmetic) instead of re-reading its state step by step
        const int ph = sl.phase, tr = sl.trial, le = sl.ls_evals;
        const double al = sl.alpha, dd = sl.ddir, bs = sl.best;
        // f = 0.5 sum_p fpart[p]: lane p holds part p's partial, one
        // fixed-order wavefront sum
        static_assert(Seg::NP <= kWave, "one part per lane");
        const double fu = lane < Seg::NP ? L.fpart[lane * kTileStarts + j] : 0.0;
        if (ph != 0) {
          const double f = 0.5 * wave_sum(fu);
          if (ph == 2 && !tile_accepts(f, al, dd, bs)) {
            sl.ls_evals = le + 1;
            sl.alpha = al * 0.5;
            sl.trial = tr + 1;
            sl.phase = tr + 1 == kLineSearchTrials ? 3 : 2;
            tile_submit<D>(L, j, lane, al * 0.5);
          } else {
            const bool stop = P.required > 0 && L.conv_sum[par] >= P.required;
            tile_slot_step<D>(L, sl, j, P, a.out, slot_mem(j), f, stop);
          }
        }
        if (lane == 0)
          L.live[j] = sl.phase != 0;
      }
      __syncthreads();
      const uint4 lv = *reinterpret_cast<const uint4*>(L.live);
      if ((lv.x | lv.y | lv.z | lv.w) == 0u)
        break;
    }
    if (tid == kStopReader)
      tile_count_converged(L, ctl);
    __syncthreads(); // LDS reuse by the next tile
  }
}

} // namespace zeus::dev
