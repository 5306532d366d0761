// bfgs_group.hpp -- per-workgroup BFGS: one workgroup owns one start at a
// time and walks a strided list of starts (persistent grid).
//
// Restates the reference's tile form (include/bfgs_parallel.cuh:67-285) and
// keeps its per-start semantics, but every lane works, not only lane 0:
//   * H lives in the row-chunk layout of block.hpp, in registers when it
//     fits and in a per-workgroup global slab otherwise;
//   * one read-modify-write pass over H per iteration: the rank-2 update of
//     iteration k is applied lazily inside iteration k+1's product
//     v = H' g_new, and since H is symmetric, u = H y = v + p and the next
//     direction p' = -H' g follow from v, u and two dot products;
//   * the objective is evaluated cooperatively (GroupEval below), the
//     gradient by a hand-written adjoint or by D dual passes spread over
//     the lanes (the reference's grad_ad_tile method, bfgs_parallel.cuh:17-54).
// Sums are workgroup reductions in a fixed order: results are deterministic
// but not bit-identical to the serial oracle (parity is to tolerance).
#pragma once

#include "batch.hpp"
#include "block.hpp"
#include "lane_ops.hpp"
#include "philox.hpp"
#include "prng.hpp"

namespace zeus::dev {

// LDS scratch handed to the evaluators.
struct GroupScratch {
  double* part;   // CH * D
  double* stash;  // D
  double* scalar; // 4
};

// ---------------------------------------------------------------------------
// Default evaluator: one lane evaluates (works for any functor); gradient by
// the functor's adjoint on one lane, or by dual passes spread over lanes.
template <class F, int D, int NT, Gradient G>
struct GroupEvalLane {
  GroupScratch sc;
  int par = 0;
  ZEUS_D void bind(const ProblemDesc&) {}
  ZEUS_D double value(const F& f, const double* x, Block<NT>&)
  {
    double* b = sc.scalar + par;
    par ^= 1;
    if (threadIdx.x == 0)
      *b = lane_value<F, D>(f, x);
    __syncthreads();
    return *b;
  }
  ZEUS_D void gradient(const F& f, const double* x, double* g, bool, Block<NT>&)
  {
    if constexpr (G == Gradient::Adjoint && HasAdjoint<F>) {
      if (threadIdx.x == 0)
        (void)f.value_and_gradient(x, g);
    } else {
      for (int i = threadIdx.x; i < D; i += NT)
        g[i] = lane_dual_pass<F, D>(f, x, i);
    }
  }
};

// Dual gradient shared by the cooperative evaluators: pass i on lane i.
template <class F, int D, int NT>
ZEUS_D ZEUS_FI void
group_dual_gradient(const F& f, const double* x, double* g)
{
  for (int i = threadIdx.x; i < D; i += NT)
    g[i] = lane_dual_pass<F, D>(f, x, i);
}

// ---------------------------------------------------------------------------
// Gaussian: Cs = (C + C^T) / 2 (prepared once per call, ProblemDesc::aux)
// held in the row-chunk layout; f = 0.5 x.(Cs x); g = Cs x is the
// by-product of the value, so an accepted line-search trial already carries
// the gradient.
template <int D, int NT, Gradient G>
struct GroupEvalGaussian {
  using F = obj::Gaussian<D>;
  using Lay = RowChunk<D, NT>;
  static constexpr bool kReg = Lay::kPerThread <= 50;
  DistMatrix<D, NT, kReg> C;
  const double* Cs = nullptr;
  GroupScratch sc;

  ZEUS_D void bind(const ProblemDesc& P)
  {
    Cs = P.aux;
    if constexpr (kReg)
      C.load(Cs);
  }
  // part <- chunk partials of Cs x
  ZEUS_D ZEUS_FI void partials(const double* x)
  {
    if constexpr (kReg) {
      C.pass(x, nullptr, nullptr, false, sc.part);
    } else {
      const int c = threadIdx.x % Lay::CH, r0 = threadIdx.x / Lay::CH;
      for (int k = 0; k < Lay::R; ++k) {
        const int row = r0 + k * Lay::RPP;
        if (row < D) {
          double acc = 0.0;
          for (int l = 0; l < Lay::L; ++l) {
            const int col = c * Lay::L + l;
            if (col < D) // Cs symmetric: read column-wise, lanes coalesce
              acc = ::fma(Cs[col * D + row], x[col], acc);
          }
          sc.part[c * D + row] = acc;
        }
      }
    }
  }
  ZEUS_D double value(const F&, const double* x, Block<NT>& blk)
  {
    partials(x);
    __syncthreads();
    double t = 0.0;
    for (int row = threadIdx.x; row < D; row += NT) {
      const double v = chunk_total<D, NT>(sc.part, row);
      sc.stash[row] = v;
      t = ::fma(x[row], v, t);
    }
    return 0.5 * blk.sum1(t);
  }
  ZEUS_D void gradient(const F& f, const double* x, double* g, bool at_last_value,
                       Block<NT>& blk)
  {
    if constexpr (G == Gradient::Dual) {
      group_dual_gradient<F, D, NT>(f, x, g);
    } else {
      if (!at_last_value) {
        partials(x);
        __syncthreads();
        for (int row = threadIdx.x; row < D; row += NT)
          sc.stash[row] = chunk_total<D, NT>(sc.part, row);
      }
      for (int row = threadIdx.x; row < D; row += NT)
        g[row] = sc.stash[row];
    }
    (void)blk;
  }
};

// ---------------------------------------------------------------------------
// Separable sums over the coordinates (Rosenbrock, Rastrigin, Ackley,
// sphere): lane i evaluates term i, the workgroup sums.
template <class F, int D, int NT, Gradient G>
struct GroupEvalSeparable {
  GroupScratch sc;
  ZEUS_D void bind(const ProblemDesc&) {}

  ZEUS_D double value(const F&, const double* x, Block<NT>& blk)
  {
    if constexpr (F::id == obj::kAckley) {
      double v[2] = {0.0, 0.0};
      for (int i = threadIdx.x; i < D; i += NT) {
        v[0] = ::fma(x[i], x[i], v[0]);
        v[1] += ::cos(2.0 * obj::kPi * x[i]);
      }
      blk.template sum<2>(v);
      const double term1 = -20.0 * ::exp(-0.2 * ::sqrt(v[0] / int(D)));
      const double term2 = -::exp(v[1] / int(D));
      return term1 + term2 + 20.0 + ::exp(1.0);
    } else {
      double t = 0.0;
      for (int i = threadIdx.x; i < D; i += NT)
        t += term(x, i);
      const double s = blk.sum1(t);
      if constexpr (F::id == obj::kRastrigin)
        return 10.0 * D + s;
      else
        return s;
    }
  }
  ZEUS_D ZEUS_FI double term(const double* x, int i) const
  {
    if constexpr (F::id == obj::kRosenbrock) {
      if (i >= D - 1)
        return 0.0;
      const double t1 = 1.0 - x[i];
      const double t2 = x[i + 1] - x[i] * x[i];
      return t1 * t1 + 100.0 * t2 * t2;
    } else if constexpr (F::id == obj::kRastrigin) {
      return x[i] * x[i] - 10.0 * ::cos(2.0 * obj::kPi * x[i]);
    } else {
      return x[i] * x[i];
    }
  }
  ZEUS_D void gradient(const F& f, const double* x, double* g, bool, Block<NT>& blk)
  {
    if constexpr (G == Gradient::Dual) {
      group_dual_gradient<F, D, NT>(f, x, g);
    } else if constexpr (F::id == obj::kRosenbrock) {
      for (int i = threadIdx.x; i < D; i += NT) {
        double gi = 0.0;
        if (i < D - 1) {
          const double t1 = 1.0 - x[i];
          const double t2 = x[i + 1] - x[i] * x[i];
          gi = -2.0 * t1 - (2.0 * x[i]) * (200.0 * t2);
        }
        if (i > 0) {
          const double t2p = x[i] - x[i - 1] * x[i - 1];
          gi += 200.0 * t2p;
        }
        g[i] = gi;
      }
    } else if constexpr (F::id == obj::kRastrigin) {
      const double k = 10.0 * (2.0 * obj::kPi);
      for (int i = threadIdx.x; i < D; i += NT)
        g[i] = 2.0 * x[i] + k * ::sin(2.0 * obj::kPi * x[i]);
    } else if constexpr (F::id == obj::kAckley) {
      double v[2] = {0.0, 0.0};
      for (int i = threadIdx.x; i < D; i += NT) {
        v[0] = ::fma(x[i], x[i], v[0]);
        v[1] += ::cos(2.0 * obj::kPi * x[i]);
      }
      blk.template sum<2>(v);
      const double r = ::sqrt(v[0] / int(D));
      const double e1 = ::exp(-0.2 * r);
      const double e2 = ::exp(v[1] / int(D));
      const double q = (4.0 * e1) / (double(int(D)) * r);
      const double c = (2.0 * obj::kPi / int(D)) * e2;
      for (int i = threadIdx.x; i < D; i += NT)
        g[i] = q * x[i] + c * ::sin(2.0 * obj::kPi * x[i]);
    } else {
      for (int i = threadIdx.x; i < D; i += NT)
        g[i] = 2.0 * x[i];
    }
  }
};

// ---------------------------------------------------------------------------
// Poisson NLL: bins spread over the lanes.
template <int NT, Gradient G>
struct GroupEvalLogLik {
  using F = obj::LogLikelihood;
  GroupScratch sc;
  ZEUS_D void bind(const ProblemDesc&) {}
  ZEUS_D double value(const F& f, const double* th, Block<NT>& blk)
  {
    const double t0 = th[0], t1 = th[1], t2 = th[2], t3 = th[3];
    const double tt[4] = {t0, t1, t2, t3};
    double t = 0.0;
    for (int i = threadIdx.x; i < f.n; i += NT)
      t += f.bin_value(tt, i);
    return blk.sum1(t);
  }
  ZEUS_D void gradient(const F& f, const double* th, double* g, bool, Block<NT>& blk)
  {
    if constexpr (G == Gradient::Dual) {
      group_dual_gradient<F, 4, NT>(f, th, g);
    } else {
      const double a = th[0], b = th[1], c = th[2], d = th[3];
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      for (int i = threadIdx.x; i < f.n; i += NT) {
        const double x = f.P[i];
        const double k = f.P[f.n + i];
        const double lx = ::log(x);
        const double pw = ::pow(1.0 - x, b);
        const double qw = ::pow(x, c + d * lx);
        const double mean = a * pw / qw;
        const double w = 1.0 - k / mean;
        const double wm = w * mean;
        v[0] += w * (pw / qw);
        v[1] += wm * ::log(1.0 - x);
        v[2] -= wm * lx;
        v[3] -= wm * (lx * lx);
      }
      blk.template sum<4>(v);
      if (threadIdx.x < 4)
        g[threadIdx.x] = v[threadIdx.x];
    }
  }
};

// ---------------------------------------------------------------------------
// MLP (1-16-16-1 tanh, least squares over m points): points spread over the
// lanes.  Value: each lane's squared residuals, one workgroup sum.  Adjoint
// gradient: each lane runs its point's forward and reverse sweep
// (activations recomputed, no tape), then the sums over points are formed
// in a fixed order: the 65 per-neuron terms (output weights and bias,
// hidden biases, first-layer weights) by wavefront sums and a per-wavefront
// fold, the 256 hidden-to-hidden weights as dot products over the points
// staged in LDS (lane k*16 + j owns weight (k, j)), half a workgroup's points
// at a time so that the staging fits two workgroups per CU.  The dual
// gradient keeps the reference's D forward passes spread over the lanes.
template <int NT, Gradient G>
struct GroupEvalMlp {
  using F = obj::Mlp;
  static constexpr int H = F::H, W2 = F::W2, W3 = F::W3;
  static constexpr int kCols = 4 * H + 1; // w3 (16), b3, b2 (16), w1 (16), b1 (16)
  static_assert(NT >= H * H, "one lane per hidden-to-hidden weight");
  GroupScratch sc;
  ZEUS_D void bind(const ProblemDesc&) {}

  // forward pass of point p: hidden activations and output
  ZEUS_D ZEUS_FI static double forward(const F& f, const double* th, int p, double (&h1)[H],
                                       double (&h2)[H])
  {
    const double xp = f.P[p];
#pragma unroll
    for (int j = 0; j < H; ++j) {
      double sum = 0.0;
      sum += th[2 * j] * xp;
      sum += th[2 * j + 1];
      h1[j] = ::tanh(sum);
    }
#pragma unroll 4
    for (int k = 0; k < H; ++k) {
      const double* w = th + W2 + k * (H + 1);
      double sum = 0.0;
#pragma unroll
      for (int j = 0; j < H; ++j)
        sum += h1[j] * w[j];
      sum += w[H];
      h2[k] = ::tanh(sum);
    }
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < H; ++k)
      sum += h2[k] * th[W3 + k];
    sum += th[W3 + H];
    return sum;
  }

  ZEUS_D double value(const F& f, const double* th, Block<NT>& blk)
  {
    double t = 0.0;
    for (int p = threadIdx.x; p < f.m; p += NT) {
      double h1[H], h2[H];
      const double dd = forward(f, th, p, h1, h2) - f.P[f.m + p];
      t = ::fma(dd, dd, t);
    }
    return blk.sum1(t) / double(f.m);
  }

  ZEUS_D void gradient(const F& f, const double* th, double* g, bool, Block<NT>& blk)
  {
    if constexpr (G == Gradient::Dual) {
      group_dual_gradient<F, int(F::arity), NT>(f, th, g);
    } else {
      constexpr int kStage = NT / 2; // points staged at a time
      __shared__ double h1s[kStage][H + 1]; // one point per lane, padded rows
      __shared__ double d2s[kStage][H + 1];
      __shared__ double colw[NT / kWave][kCols];
      const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
      const double scale = 2.0 / double(f.m);
      for (int c = lane; c < kCols; c += kWave)
        colw[w][c] = 0.0;
      // per-neuron term c of this lane's point, summed over the wavefront
      // and added to the wavefront's running total (chunks in order)
      auto col_add = [&](int c, double v) {
        const double t = wave_sum(v);
        if (lane == 0)
          colw[w][c] += t;
      };
      double gw = 0.0; // hidden-to-hidden weight (k, j) = (tid / H, tid % H)
      for (int c0 = 0; c0 < f.m; c0 += NT) {
        const int p = c0 + tid;
        const bool live = p < f.m;
        double h1[H], h2[H], d2[H];
        double dy = 0.0;
        if (live) {
          dy = scale * (forward(f, th, p, h1, h2) - f.P[f.m + p]);
        } else {
#pragma unroll
          for (int j = 0; j < H; ++j)
            h1[j] = h2[j] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < H; ++k) {
          col_add(k, dy * h2[k]);                                  // w3
          d2[k] = (dy * th[W3 + k]) * (1.0 - h2[k] * h2[k]);
          col_add(H + 1 + k, d2[k]);                               // b2
        }
        col_add(H, dy);                                            // b3
        const double xp = live ? f.P[p] : 0.0;
#pragma unroll 4
        for (int j = 0; j < H; ++j) {
          double d1 = 0.0;
#pragma unroll
          for (int k = 0; k < H; ++k)
            d1 += d2[k] * th[W2 + k * (H + 1) + j];
          const double dj = d1 * (1.0 - h1[j] * h1[j]);
          col_add(2 * H + 1 + j, dj * xp);                         // w1
          col_add(3 * H + 1 + j, dj);                              // b1
        }
        // points c0 .. c0 + NT - 1 in two halves, in point order
#pragma unroll 1
        for (int half = 0; half < 2; ++half) {
          if (tid / kStage == half) {
#pragma unroll
            for (int j = 0; j < H; ++j) {
              h1s[tid % kStage][j] = h1[j];
              d2s[tid % kStage][j] = d2[j];
            }
          }
          __syncthreads();
          if (tid < H * H) {
            const int k = tid / H, j = tid % H;
            const int left = f.m - c0 - half * kStage;
            const int n = left < kStage ? (left > 0 ? left : 0) : kStage;
            for (int q = 0; q < n; ++q)
              gw = ::fma(d2s[q][k], h1s[q][j], gw);
          }
          __syncthreads();
        }
      }
      for (int c = tid; c < kCols; c += NT) {
        double t = 0.0;
        for (int v = 0; v < NT / kWave; ++v)
          t += colw[v][c];
        if (c < H)
          g[W3 + c] = t;
        else if (c == H)
          g[W3 + H] = t;
        else if (c < 2 * H + 1)
          g[W2 + (c - H - 1) * (H + 1) + H] = t;
        else if (c < 3 * H + 1)
          g[2 * (c - 2 * H - 1)] = t;
        else
          g[2 * (c - 3 * H - 1) + 1] = t;
      }
      if (tid < H * H)
        g[W2 + (tid / H) * (H + 1) + tid % H] = gw;
    }
    (void)blk;
  }
};

// Evaluator per objective type.  Specialisations key on the objective type
// alone (its dimension parameter is a std::size_t; deducing it through an
// int parameter would fail and silently select the one-lane fallback).
template <class F, int NT, Gradient G>
struct GroupEvalSelect {
  using type = GroupEvalLane<F, int(F::arity), NT, G>;
};
template <std::size_t D, int NT, Gradient G>
struct GroupEvalSelect<obj::Gaussian<D>, NT, G> {
  using type = GroupEvalGaussian<int(D), NT, G>;
};
template <std::size_t D, int NT, Gradient G>
struct GroupEvalSelect<obj::Rosenbrock<D>, NT, G> {
  using type = GroupEvalSeparable<obj::Rosenbrock<D>, int(D), NT, G>;
};
template <std::size_t D, int NT, Gradient G>
struct GroupEvalSelect<obj::Rastrigin<D>, NT, G> {
  using type = GroupEvalSeparable<obj::Rastrigin<D>, int(D), NT, G>;
};
template <std::size_t D, int NT, Gradient G>
struct GroupEvalSelect<obj::Ackley<D>, NT, G> {
  using type = GroupEvalSeparable<obj::Ackley<D>, int(D), NT, G>;
};
template <std::size_t D, int NT, Gradient G>
struct GroupEvalSelect<obj::Sphere<D>, NT, G> {
  using type = GroupEvalSeparable<obj::Sphere<D>, int(D), NT, G>;
};
template <int NT, Gradient G>
struct GroupEvalSelect<obj::LogLikelihood, NT, G> {
  using type = GroupEvalLogLik<NT, G>;
};
template <int NT, Gradient G>
struct GroupEvalSelect<obj::Mlp, NT, G> {
  using type = GroupEvalMlp<NT, G>;
};

// ---------------------------------------------------------------------------
// LDS vectors of one workgroup.
template <int D, int NT>
struct GroupVectors {
  double x[D], g[D], p[D], xt[D], gn[D], v[D];
  double s0[D], a0[D], s1[D], a1[D]; // current / pending rank-2 pair
  double part[RowChunk<D, NT>::CH * D];
  double stash[D];
  double red[8 * (NT / kWave)];
  double scalar[4];
  int flag;
};

template <class F, int D, int NT, Gradient G, bool HREG, class Eval>
ZEUS_D void
group_bfgs(const F& f, GroupVectors<D, NT>& L, DistMatrix<D, NT, HREG>& H, Eval& ev,
           Block<NT>& blk, const ProblemDesc& P, Control* ctl, int64_t s,
           const double* starts, const StartOut& out)
{
  const int tid = threadIdx.x;
  const int64_t sl = s - P.start0;
  double* sc = L.s0; // s of this iteration
  double* ac = L.a0;
  double* sp = L.s1; // pending pair from the previous iteration
  double* ap = L.a1;
  bool pending = false;

  H.set_identity();
  for (int d = tid; d < D; d += NT) {
    if (starts)
      L.x[d] = starts[s * D + d];
    else if (P.x0)
      L.x[d] = P.x0[sl * D + d];
    else
      L.x[d] = start_uniform_in(P, uint64_t(sl), d);
  }
  __syncthreads();
  double bestVal = ev.value(f, L.x, blk);
  ev.gradient(f, L.x, L.g, true, blk);
  __syncthreads();
  for (int i = tid; i < D; i += NT)
    L.p[i] = -L.g[i];
  int grads = 1, ls_evals = 0;
  int status = kStatusUnset;
  double fval = 333777.0, gnorm = 69.0;
  int iter = 0;
  bool done = false;

  for (; iter < P.max_iter; ++iter) {
    // trajectory step (bfgs_parallel.cuh:155-166); uniform per workgroup
    const int64_t to = traj_slot(P.traj, sl, iter);
    if (to >= 0) {
      double gg = 0.0;
      for (int i = tid; i < D; i += NT)
        gg = ::fma(L.g[i], L.g[i], gg);
      const double nrm = ::sqrt(blk.sum1(gg));
      for (int d = tid; d < D; d += NT)
        P.traj.coords[(int64_t(iter) * D + d) * P.traj.starts + sl] = L.x[d];
      if (tid == 0) {
        P.traj.fval[to] = bestVal;
        P.traj.grad[to] = nrm;
        P.traj.status[to] = int8_t(kStatusUnset);
      }
    }
    if (tid == 0)
      L.flag = load_flag(&ctl->stop_flag);
    __syncthreads();
    if (L.flag != 0) {
      status = kStatusStopped;
      fval = bestVal;
      double gg = 0.0;
      for (int i = tid; i < D; i += NT)
        gg = ::fma(L.g[i], L.g[i], gg);
      gnorm = ::sqrt(blk.sum1(gg));
      done = true;
      break;
    }
    double dd = 0.0;
    for (int i = tid; i < D; i += NT)
      dd = ::fma(L.g[i], L.p[i], dd);
    const double ddir = blk.sum1(dd);

    double alpha = 1.0, f1 = 0.0;
    bool accepted = false;
    for (int t = 0; t < kLineSearchTrials; ++t) {
      for (int j = tid; j < D; j += NT)
        L.xt[j] = ::fma(alpha, L.p[j], L.x[j]);
      __syncthreads();
      f1 = ev.value(f, L.xt, blk);
      ++ls_evals;
      if (f1 <= ::fma(kArmijoC1 * alpha, ddir, bestVal) && f1 < bestVal) {
        accepted = true;
        break;
      }
      alpha *= 0.5;
    }
    if (tid == 0 && to >= 0 && P.traj.alpha) // bfgs_parallel.cuh:182-184
      P.traj.alpha[to] = alpha;
    double fnew = f1;
    if (!accepted) {
      for (int j = tid; j < D; j += NT)
        L.xt[j] = ::fma(alpha, L.p[j], L.x[j]);
      __syncthreads();
      fnew = ev.value(f, L.xt, blk);
      ++ls_evals;
    }
    ev.gradient(f, L.xt, L.gn, true, blk);
    ++grads;
    __syncthreads();

    // v = H' g_new while applying the pending update (one pass over H)
    H.pass(L.gn, sp, ap, pending, L.part);
    __syncthreads();
    double st[2] = {0.0, 0.0};
    for (int i = tid; i < D; i += NT) {
      const double v = chunk_total<D, NT>(L.part, i);
      L.v[i] = v;
      const double si = L.xt[i] - L.x[i];
      const double yi = L.gn[i] - L.g[i];
      sc[i] = si;
      st[0] = ::fma(si, yi, st[0]);
      st[1] = ::fma(yi, v + L.p[i], st[1]); // u = H y = v + p
    }
    blk.template sum<2>(st);
    const double sTy = st[0];
    const bool upd = !(::fabs(sTy) < kCurvatureSkip);
    const bool acc = fnew < bestVal;
    if (upd) {
      const double rho = 1.0 / sTy;
      const double beta = ::fma(rho * rho, st[1], rho);
      const double hb = 0.5 * beta;
      double q[2] = {0.0, 0.0};
      for (int i = tid; i < D; i += NT) {
        const double u = L.v[i] + L.p[i];
        const double ai = ::fma(hb, sc[i], -(rho * u));
        ac[i] = ai;
        const double ga = acc ? L.gn[i] : L.g[i];
        q[0] = ::fma(ai, ga, q[0]);
        q[1] = ::fma(sc[i], ga, q[1]);
      }
      blk.template sum<2>(q);
      for (int i = tid; i < D; i += NT) {
        const double hg = acc ? L.v[i] : -L.p[i];
        L.p[i] = -::fma(ac[i], q[1], ::fma(sc[i], q[0], hg));
      }
    } else {
      for (int i = tid; i < D; i += NT)
        L.p[i] = acc ? -L.v[i] : L.p[i];
    }
    if (acc) {
      bestVal = fnew;
      for (int i = tid; i < D; i += NT) {
        L.x[i] = L.xt[i];
        L.g[i] = L.gn[i];
      }
    }
    double gg = 0.0;
    for (int i = tid; i < D; i += NT)
      gg = ::fma(L.g[i], L.g[i], gg);
    const double nrm = ::sqrt(blk.sum1(gg));
    // the new pair becomes pending for the next pass
    pending = upd;
    double* t0 = sp;
    double* t1 = ap;
    sp = sc;
    ap = ac;
    sc = t0;
    ac = t1;

    if (!::isfinite(nrm) || !::isfinite(fnew)) {
      status = kStatusNonFinite;
      fval = fnew;
      gnorm = nrm;
      done = true;
      break;
    }
    if (nrm < P.tol) {
      if (tid == 0)
        count_converged(ctl, P.required);
      status = kStatusConverged;
      fval = bestVal;
      gnorm = nrm;
      done = true;
      break;
    }
  }
  if (!done) {
    status = kStatusSurrendered;
    fval = bestVal;
    double gg = 0.0;
    for (int i = tid; i < D; i += NT)
      gg = ::fma(L.g[i], L.g[i], gg);
    gnorm = ::sqrt(blk.sum1(gg));
  }
  double e2 = 0.0;
  if (status == kStatusConverged) {
    double t = 0.0;
    for (int d = tid; d < D; d += NT) {
      const double df = L.x[d] - known_min(P.min_kind, d);
      t = ::fma(df, df, t);
    }
    e2 = blk.sum1(t);
  }
  if (tid == 0) {
    traj_final_status(P.traj, sl, status, iter, P.max_iter);
    out.status[s] = status;
    out.iter[s] = iter;
    out.fval[s] = fval;
    out.gnorm[s] = gnorm;
    out.ls_evals[s] = ls_evals;
    out.grads[s] = grads;
    out.near[s] = status == kStatusConverged && ::sqrt(e2) < 0.5;
  }
  for (int d = tid; d < D; d += NT)
    out.x[s * D + d] = L.x[d];
  __syncthreads();
}

// Resident workgroups per CU the kernel is compiled for: the MLP with the
// adjoint gradient alternates evaluation phases (no memory traffic) with the
// streamed H pass (HBM bound), so two workgroups per CU overlap one's pass
// with the other's evaluation; the register budget is set for that.
template <class F, Gradient G>
constexpr int group_min_blocks()
{
  if constexpr (requires { F::id; }) // built-in objectives
    return F::id == obj::kMlp && G == Gradient::Adjoint ? 2 : 1;
  else
    return 1;
}

template <class F, int D, int NT, Gradient G, bool HREG, class Prov>
__global__ __launch_bounds__(NT, (group_min_blocks<F, G>())) void
bfgs_group_kernel(ClassArgs a, Prov prov)
{
  __shared__ GroupVectors<D, NT> L;
  Block<NT> blk{L.red};
  using Eval = typename GroupEvalSelect<F, NT, G>::type;
  static_assert(int(F::arity) == D, "objective arity and kernel dimension differ");
  Eval ev;
  ev.sc = GroupScratch{L.part, L.stash, L.scalar};
  DistMatrix<D, NT, HREG> H;
  if constexpr (!HREG)
    H.slab = a.h_slab + int64_t(blockIdx.x) * a.h_slab_stride;
  int bound = -1;
  for (int64_t s = blockIdx.x; s < a.total_starts; s += gridDim.x) {
    const int pi = find_problem(a.probs, a.n_probs, s);
    const ProblemDesc P = a.probs[pi];
    if (pi != bound) {
      ev.bind(P);
      bound = pi;
    }
    decltype(auto) f = prov.template get<F>(P);
    group_bfgs<F, D, NT, G, HREG>(f, L, H, ev, blk, P, a.ctl + pi, s, a.starts, a.out);
  }
}

} // namespace zeus::dev
