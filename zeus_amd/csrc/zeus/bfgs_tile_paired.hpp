// bfgs_tile_paired.hpp -- paired tile form of the BFGS for the Gaussian
// objective: one 1024-thread workgroup advances 16 starts as two half-tiles
// of 8 starts, each start evaluating two line-search trials per round.
//
// Why.  In the single tile form (bfgs_tile.hpp) a round is the MFMA product
// followed by the starts' logic, separated by barriers, so the matrix cores
// idle while the logic runs and the logic waits for the product.  Here the
// two half-tiles run out of phase: in half-step h, the 8 wavefronts of
// half-tile h % 2 run the product of their starts' points while the other 8
// wavefronts run their starts' logic on the product of the previous
// half-step.  Each half-tile still fills the MFMA's 16 columns: every start
// submits two points per round, the next two trials of its line search
// (alpha_t and alpha_t / 2; the line search takes them in order, so a
// second trial is used only when the first is rejected, and otherwise its
// evaluation is discarded and not counted).  A start's line search over E
// trials takes ceil(E / 2) rounds.
//
// Same method and arithmetic as the single tile form: the same product
// units (16-row block of Cs, k half) with the same MFMA accumulation order,
// the same partial-value sums, the same per-start logic.  A start's points,
// values and gradients are therefore bit for bit those of the single tile
// form, and without an early stop so is every per-start result; with an
// early stop only the moment at which starts see the stop flag moves.
//
// Layout (LDS).  x, p, g of the 16 starts (rows D .. KP-1 of x and p zero),
// the two trial steps of each start, and per half-tile the two k-half
// partial products and the units' partial values.  The product's B operand
// (point j of the half-tile at row k) is formed on the fly as
// fma(alpha_j, p[k], x[k]), exactly the point a start submits.
#pragma once

#include "bfgs_tile.hpp"

namespace zeus::dev {

constexpr int kPairHalves = 2;                          // half-tiles
constexpr int kPairStarts = kTileStarts / kPairHalves;  // starts per half-tile
constexpr int kPairWaves = kTileWaves / kPairHalves;    // wavefronts per half-tile
static_assert(kPairStarts * 2 == kCsTile, "two points per start fill the MFMA columns");
static_assert(kPairWaves == kPairStarts, "one wavefront per start of a half-tile");

template <int D>
struct PairLds {
  using S = TileShape<D>;
  static constexpr int VS = S::KP; // per-start stride of x, p, g
  double x[kTileStarts * VS];
  double p[kTileStarts * VS];
  double g[kTileStarts * VS];
  double gt[kPairHalves][2 * D * S::XS]; // per half-tile: [k half][row][column]
  double fpart[kPairHalves][S::U * kCsTile];
  double alpha[kTileStarts][2]; // the two trial steps a start submitted
  // live[h % 2][j]: slot j still runs after half-step h, written by slot j's
  // wavefront in half-step h and read by every wavefront after its barrier;
  // the next write to the same row comes two barriers later
  int live[2][kTileStarts];
  // early stop: written by one thread at the end of half-step h into
  // stop[h % 2] and read by the logic of half-step h + 1, so a write never
  // meets a read of the same word within a half-step
  int stop[2];
  // next product unit to take in half-step h (index h % 2); reset in
  // half-step h - 1, after its last use in h - 2
  int next_unit[2];
  TileSlot slot[kTileStarts];
  ProblemDesc prob;
};

// Product unit u (rows 16 (u / 2) .. + 15, k half u % 2) of half-tile `half`:
// column c is point c % 2 of the half-tile's start c / 2.
template <int D, int KB0, int KB1>
ZEUS_D ZEUS_FI mfma_acc_t
pair_unit_range(const double* Cf, const double* xs, const double* ps, double a, int mb, int li,
                int lk)
{
  using S = CsShape<D>;
  const gdouble_ptr frag = as_global(Cf) + (int64_t(mb) * S::KB << 6) + (li + 16 * lk);
  double av[KB1 - KB0];
#pragma unroll
  for (int kb = KB0; kb < KB1; ++kb)
    av[kb - KB0] = frag[kb << 6];
  __builtin_amdgcn_sched_barrier(0); // every fragment load issued before the chain
  mfma_acc_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kb = KB0; kb < KB1; ++kb) {
    const int r = 4 * kb + lk;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kb - KB0], ::fma(a, ps[r], xs[r]), acc, 0, 0,
                                               0);
  }
  return acc;
}

// Logic of start slot w for one round, run by its wavefront.  f0 and f1 are
// the values at the two submitted points (columns 2 (w % 8) and + 1 of the
// half-tile's product).
template <int D>
ZEUS_D void
pair_slot_step(PairLds<D>& L, const int w, const ProblemDesc& P, Control* ctl,
               const StartOut& out, double* mem, double f0, double f1, int stop)
{
  using S = TileShape<D>;
  constexpr int EL = S::EL;
  constexpr int VS = PairLds<D>::VS;
  TileSlot& st = L.slot[w];
  const int lane = threadIdx.x % kWave;
  const int half = w / kPairStarts;
  const int c0 = 2 * (w % kPairStarts); // first column of this start
  const double* gt = L.gt[half];
  double* X = L.x + w * VS;
  double* G = L.g + w * VS;
  double* Pd = L.p + w * VS;
  double* hist = mem;
  double* B = mem + int64_t(S::kHist) * 2 * D;
  int row[EL];
  bool ok[EL];
#pragma unroll
  for (int e = 0; e < EL; ++e) {
    row[e] = lane + e * kWave;
    ok[e] = row[e] < D;
  }
  auto load = [&](const double* v, double (&r)[EL]) {
#pragma unroll
    for (int e = 0; e < EL; ++e)
      r[e] = ok[e] ? v[row[e]] : 0.0;
  };
  auto load_grad = [&](int col, double (&r)[EL]) { // k halves in fixed order
#pragma unroll
    for (int e = 0; e < EL; ++e)
      r[e] = ok[e] ? gt[row[e] * S::XS + col] + gt[(D + row[e]) * S::XS + col] : 0.0;
  };
  // the next two trials (alpha, alpha / 2), or one point in both columns
  auto submit = [&](int phase, bool two) {
    if (lane == 0) {
      L.alpha[w][0] = st.alpha;
      L.alpha[w][1] = two ? st.alpha * 0.5 : st.alpha;
    }
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
  // top of a BFGS iteration: MAX_ITER, stop flag, first two trials
  auto begin_iteration = [&]() {
    if (st.iter >= P.max_iter) {
      finish(kStatusSurrendered, st.best, gnorm_now());
      return;
    }
    if (stop != 0) {
      finish(kStatusStopped, st.best, gnorm_now());
      return;
    }
    double g[EL], p[EL];
    load(G, g);
    load(Pd, p);
    st.ddir = tile_dot<EL>(g, p);
    st.alpha = 1.0;
    st.trial = 0;
    submit(2, true);
  };

  if (st.phase == 1) { // first point (column c0): value and gradient
    double g[EL];
    load_grad(c0, g);
#pragma unroll
    for (int e = 0; e < EL; ++e)
      if (ok[e]) {
        G[row[e]] = g[e];
        Pd[row[e]] = -g[e];
      }
    st.best = f0;
    st.grads = 1;
    begin_iteration();
    return;
  }
  int col = 0; // column of the point that ends the iteration
  double f = f0;
  if (st.phase == 2) { // trials t (column 0) and t + 1 (column 1), t even
    ++st.ls_evals;
    bool accept = f0 <= ::fma(kArmijoC1 * st.alpha, st.ddir, st.best) && f0 < st.best;
    if (!accept) {
      st.alpha *= 0.5;
      ++st.trial; // t + 1 <= kLineSearchTrials - 1: a tested trial
      ++st.ls_evals;
      col = 1;
      f = f1;
      accept = f1 <= ::fma(kArmijoC1 * st.alpha, st.ddir, st.best) && f1 < st.best;
      if (!accept) {
        st.alpha *= 0.5;
        if (++st.trial == kLineSearchTrials)
          submit(3, false); // the point after the last rejected trial
        else
          submit(2, true);
        return;
      }
    }
  } else { // phase 3: point after the last rejected trial (column 0)
    ++st.ls_evals;
  }

  // ---- end of the iteration: gradient at the new point and the update ----
  const double fnew = f;
  const int cn = c0 + col;
  const double an = L.alpha[w][col];
  ++st.grads;
  const bool acc = fnew < st.best;
  double gn[EL], v[EL];
  load_grad(cn, gn);
  // v = H gn, H = B + sum_j (s_j a_j^T + a_j s_j^T)
  if (st.base) {
#pragma unroll
    for (int e = 0; e < EL; ++e)
      v[e] = 0.0;
#pragma unroll 1
    for (int c = 0; c < D; ++c) {
      const double gc = gt[c * S::XS + cn] + gt[(D + c) * S::XS + cn];
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
#pragma unroll 2
  for (int j = 0; j < st.k; ++j) { // v accumulates in order j
    double sj[EL], aj[EL];
    load(hist + int64_t(j) * 2 * D, sj);
    load(hist + int64_t(j) * 2 * D + D, aj);
    const double cj = tile_dot<EL>(aj, gn);
    const double dj = tile_dot<EL>(sj, gn);
#pragma unroll
    for (int e = 0; e < EL; ++e)
      v[e] = ::fma(sj[e], cj, ::fma(aj[e], dj, v[e]));
  }
  // s = xt - x, y = gn - g, u = H y = H gn - H g = v + p
  double s[EL], xt[EL];
  double sTy, yu;
  {
    double t0 = 0.0, t1 = 0.0;
#pragma unroll
    for (int e = 0; e < EL; ++e) {
      const double xv = ok[e] ? X[row[e]] : 0.0;
      const double pv = ok[e] ? Pd[row[e]] : 0.0;
      const double gv = ok[e] ? G[row[e]] : 0.0;
      xt[e] = ok[e] ? ::fma(an, pv, xv) : 0.0;
      s[e] = xt[e] - xv;
      const double y = gn[e] - gv;
      t0 = ::fma(s[e], y, t0);
      t1 = ::fma(y, v[e] + pv, t1);
    }
    sTy = wave_sum(t0);
    yu = wave_sum(t1);
  }
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
        const double pv = ok[e] ? Pd[row[e]] : 0.0;
        a[e] = ok[e] ? ::fma(hb, s[e], -(rho * (v[e] + pv))) : 0.0;
        const double ga = acc ? gn[e] : (ok[e] ? G[row[e]] : 0.0);
        t0 = ::fma(a[e], ga, t0);
        t1 = ::fma(s[e], ga, t1);
      }
      q0 = wave_sum(t0);
      q1 = wave_sum(t1);
    }
    // p' = -H' g = -(H g + s (a.g) + a (s.g)), H g = v (accepted) or -p
#pragma unroll
    for (int e = 0; e < EL; ++e)
      if (ok[e]) {
        const double hg = acc ? v[e] : -Pd[row[e]];
        Pd[row[e]] = -::fma(a[e], q1, ::fma(s[e], q0, hg));
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
      if (ok[e]) {
        sk[row[e]] = s[e];
        sk[D + row[e]] = a[e];
      }
    ++st.k;
  } else if (acc) {
#pragma unroll
    for (int e = 0; e < EL; ++e)
      if (ok[e])
        Pd[row[e]] = -v[e];
  }
  if (acc) {
    st.best = fnew;
#pragma unroll
    for (int e = 0; e < EL; ++e)
      if (ok[e]) {
        X[row[e]] = xt[e];
        G[row[e]] = gn[e];
      }
  }
  const double nrm = gnorm_now();
  if (!::isfinite(nrm) || !::isfinite(fnew)) {
    finish(kStatusNonFinite, fnew, nrm);
    return;
  }
  if (nrm < P.tol) {
    if (lane == 0)
      (void)__hip_atomic_fetch_add(&ctl->converged, 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT); // no wait for the old value
    finish(kStatusConverged, st.best, nrm);
    return;
  }
  ++st.iter;
  begin_iteration();
}

template <int D>
__global__ __launch_bounds__(kTileThreads) void
bfgs_tile_paired_kernel(TileArgs a)
{
  using S = TileShape<D>;
  constexpr int VS = PairLds<D>::VS;
  __shared__ PairLds<D> L;
  const int tid = threadIdx.x, lane = tid % kWave;
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int li = lane & 15, lk = lane >> 4;
  const int my_half = w / kPairWaves, lw = w % kPairWaves;
  auto slot_mem = [&](int j) {
    return a.slot_mem + (int64_t(blockIdx.x) * kTileStarts + j) * S::kSlotDoubles;
  };

  for (int64_t tile = blockIdx.x; tile < a.n_tiles; tile += gridDim.x) {
    const int pi = wave_uniform(a.tile_problem[tile]);
    if (tid == 0)
      L.prob = a.probs[pi];
    __syncthreads();
    const ProblemDesc& P = L.prob;
    const double* Cf = P.aux_frag;
    // slot w: its wavefront writes the slot (every lane the same values), its
    // start point and zero direction and gradient (padding rows included)
    {
      TileSlot& st = L.slot[w];
      const TileSlot init{-1, 0, 0, 0, 0, 0, 0, 0, 0.0, 0.0, 0.0};
      st = init;
      const int64_t sl = (tile - P.block0) * kTileStarts + w; // local start
      const bool has = sl < P.n_starts;
      double* X = L.x + w * VS;
      for (int d = lane; d < VS; d += kWave) {
        double v = 0.0;
        if (has && d < D)
          v = P.x0 ? P.x0[sl * D + d] : start_uniform_in(P, uint64_t(sl), d);
        X[d] = v;
        L.p[w * VS + d] = 0.0;
        L.g[w * VS + d] = 0.0;
      }
      if (has) {
        st.s = P.start0 + sl;
        st.phase = 1;
      }
      if (lane == 0) {
        L.alpha[w][0] = 0.0; // the first point: x + 0 p
        L.alpha[w][1] = 0.0;
        L.live[0][w] = L.live[1][w] = has;
      }
      if (tid == 0) {
        L.stop[0] = L.stop[1] = 0;
        L.next_unit[0] = L.next_unit[1] = 0;
      }
    }
    Control* ctl = a.ctl + pi;
    __syncthreads();

    for (int h = 0;; ++h) {
      const int ph = h & 1; // half-tile whose product runs in this half-step
      constexpr int kStopReader = kTileThreads - kWave;
      int conv = 0;
      if (tid == kStopReader)
        conv = __hip_atomic_load(&ctl->converged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (my_half != ph && h > 0) {
        // ---- logic of the other half-tile's start w on its last product ----
        if (L.slot[w].phase != 0) {
          static_assert(S::U <= kWave, "one unit per lane");
          const double* fpv = L.fpart[my_half];
          const int c0 = 2 * lw;
          const double fu0 = lane < S::U ? fpv[lane * kCsTile + c0] : 0.0;
          const double fu1 = lane < S::U ? fpv[lane * kCsTile + c0 + 1] : 0.0;
          const double f0 = 0.5 * wave_sum(fu0);
          const double f1 = 0.5 * wave_sum(fu1);
          pair_slot_step<D>(L, w, P, ctl, a.out, slot_mem(w), f0, f1, L.stop[(h + 1) & 1]);
        }
      }
      // ---- product of half-tile ph: every wavefront takes units from a
      // counter, the other half-tile's wavefronts once their logic is done
      // (a unit's result does not depend on which wavefront computes it) ----
      int live_ph = 0;
#pragma unroll
      for (int j = 0; j < kPairStarts; ++j)
        live_ph |= L.live[(h + 1) & 1][ph * kPairStarts + j];
      if (live_ph) {
        const int sj = ph * kPairStarts + li / 2; // start of column li
        const double al = L.alpha[sj][li & 1];
        const double* xs = L.x + sj * VS;
        const double* ps = L.p + sj * VS;
        double* gt = L.gt[ph];
        double* fp_out = L.fpart[ph];
        for (;;) {
          int u = 0;
          if (lane == 0)
            u = atomicAdd(&L.next_unit[h & 1], 1);
          u = __builtin_amdgcn_readfirstlane(u);
          if (u >= S::U)
            break;
          const int mb = u >> 1, kh = u & 1;
          const mfma_acc_t acc =
            kh ? pair_unit_range<D, S::KB1, S::KB>(Cf, xs, ps, al, mb, li, lk)
               : pair_unit_range<D, 0, S::KB1>(Cf, xs, ps, al, mb, li, lk);
          double* gp = gt + kh * D * S::XS;
          double fp = 0.0;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int rr = 16 * mb + 4 * q + lk;
            if (rr < D) {
              gp[rr * S::XS + li] = acc[q];
              fp = ::fma(::fma(al, ps[rr], xs[rr]), acc[q], fp);
            }
          }
          fp += __shfl_xor(fp, 16, kWave);
          fp += __shfl_xor(fp, 32, kWave);
          if (lk == 0)
            fp_out[u * kCsTile + li] = fp;
        }
      }
      if (tid == 0) // counter of half-step h + 1 (last used in h - 1)
        L.next_unit[(h + 1) & 1] = 0;
      // every wavefront publishes its own slot's state for this half-step
      if (lane == 0)
        L.live[h & 1][w] = L.slot[w].phase != 0;
      if (tid == kStopReader)
        L.stop[h & 1] = P.required > 0 && conv >= P.required;
      __syncthreads();
      int any = 0;
#pragma unroll
      for (int k = 0; k < kTileStarts; ++k)
        any |= L.live[h & 1][k];
      if (!any)
        break;
    }
    __syncthreads(); // LDS reuse by the next tile
  }
}

} // namespace zeus::dev
