// bfgs_team.hpp -- team BFGS: TS lanes of a wavefront own one start, a
// wavefront advances 64 / TS starts at once (persistent grid of teams).
//
// For small dimensions a whole workgroup per start leaves most lanes idle
// (20 of 64 at D = 20) and one lane per start runs the D-term evaluations
// serially; a team splits the coordinates i = r + TS k over its lanes
// (r = rank in the team).  Per-lane state: the owned coordinates of x, g,
// p, the trial point and the new gradient, and the owned rows of H (D
// columns each) -- all in registers.  The objective is a separable sum
// (Rosenbrock, Rastrigin, Ackley, sphere): each lane forms its terms from the
// team's point in LDS (Rosenbrock's coupling reads the neighbour
// coordinate there) and the team sums with DPP exchanges inside its 4-, 8-
// or 16-lane group, which give every lane of the team the same bits.  Vectors the H pass needs in full (the new
// gradient and the pending rank-2 pair) are broadcast through the team's LDS
// slot.  Teams of a wavefront take different branches independently; every
// value a branch depends on is uniform over the team.
//
// Method: the per-workgroup form's (bfgs_group.hpp), i.e. the reference's
// parallel BFGS (include/bfgs_parallel.cuh:67-285) -- same line search,
// update, skip rule, statuses and counters.  The rank-2 update is applied to
// the register rows at the end of its iteration (the per-workgroup form
// defers it into the next product pass to save a pass over memory; the
// entries are the same either way).  Sums run in
// another fixed order than the serial oracle: parity is to tolerance.
#pragma once

#include "batch.hpp"
#include "block.hpp"
#include "lane_ops.hpp"
#include "philox.hpp"
#include "prng.hpp"

namespace zeus::dev {

constexpr int kTeamThreads = 256; // workgroup: four wavefronts of teams

// Team sizes offered for a dimension: H rows held per lane stay within 64
// doubles (ceil(D / TS) rows of D), and every lane owns a coordinate.
template <int D, int TS>
constexpr bool
team_fits()
{
  return D >= TS && ((D + TS - 1) / TS) * D <= 64;
}

// Sum over the TS lanes of a team (aligned groups of 4, 8 or 16 lanes in a
// row): lane xor 1, xor 2 (quad_perm), then the half-row and row mirrors.
// Each step adds the same two operands in both lanes of a pair, so every
// lane of the team ends with the same bits; no step reads outside the team.
template <int TS>
ZEUS_D ZEUS_FI double
team_sum(double v)
{
  static_assert(TS == 4 || TS == 8 || TS == 16, "team size");
  v += dpp_f64<0xB1>(v); // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v); // quad_perm [2,3,0,1]
  if constexpr (TS >= 8)
    v += dpp_f64<0x141>(v); // row_half_mirror
  if constexpr (TS >= 16)
    v += dpp_f64<0x140>(v); // row_mirror
  return v;
}

// LDS writes of a team read back by other lanes of the same wavefront: the
// wavefront's LDS operations complete in order; these fences keep the
// compiler from moving accesses across the exchange.
ZEUS_D ZEUS_FI void
team_exchange()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value and adjoint gradient of a separable objective for a team; xl is the
// team's point in LDS, lanes own i = r + TS k.
template <class F, int D, int TS>
struct TeamEval {
  static constexpr int KC = (D + TS - 1) / TS;

  ZEUS_D ZEUS_FI static double term(const double* x, int i)
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

  ZEUS_D static double value(const double* xl, int r)
  {
    if constexpr (F::id == obj::kAckley) {
      double v0 = 0.0, v1 = 0.0;
#pragma unroll 1
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        if (i < D) {
          v0 = ::fma(xl[i], xl[i], v0);
          v1 += ::cos(2.0 * obj::kPi * xl[i]);
        }
      }
      v0 = team_sum<TS>(v0);
      v1 = team_sum<TS>(v1);
      const double term1 = -20.0 * ::exp(-0.2 * ::sqrt(v0 / int(D)));
      const double term2 = -::exp(v1 / int(D));
      return term1 + term2 + 20.0 + ::exp(1.0);
    } else {
      double t = 0.0;
#pragma unroll 1
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        if (i < D)
          t += term(xl, i);
      }
      const double s = team_sum<TS>(t);
      if constexpr (F::id == obj::kRastrigin)
        return 10.0 * D + s;
      else
        return s;
    }
  }

  template <Gradient G>
  ZEUS_D static void gradient(const F& f, const double* xl, int r, double (&g)[KC])
  {
    if constexpr (G == Gradient::Dual) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        g[k] = i < D ? lane_dual_pass<F, D>(f, xl, i) : 0.0;
      }
    } else if constexpr (F::id == obj::kRosenbrock) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        double gi = 0.0;
        if (i < D) {
          if (i < D - 1) {
            const double t1 = 1.0 - xl[i];
            const double t2 = xl[i + 1] - xl[i] * xl[i];
            gi = -2.0 * t1 - (2.0 * xl[i]) * (200.0 * t2);
          }
          if (i > 0) {
            const double t2p = xl[i] - xl[i - 1] * xl[i - 1];
            gi += 200.0 * t2p;
          }
        }
        g[k] = gi;
      }
    } else if constexpr (F::id == obj::kRastrigin) {
      const double kk = 10.0 * (2.0 * obj::kPi);
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        g[k] = i < D ? 2.0 * xl[i] + kk * ::sin(2.0 * obj::kPi * xl[i]) : 0.0;
      }
    } else if constexpr (F::id == obj::kAckley) {
      double v0 = 0.0, v1 = 0.0;
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        if (i < D) {
          v0 = ::fma(xl[i], xl[i], v0);
          v1 += ::cos(2.0 * obj::kPi * xl[i]);
        }
      }
      v0 = team_sum<TS>(v0);
      v1 = team_sum<TS>(v1);
      const double rr = ::sqrt(v0 / int(D));
      const double e1 = ::exp(-0.2 * rr);
      const double e2 = ::exp(v1 / int(D));
      const double q = (4.0 * e1) / (double(int(D)) * rr);
      const double c = (2.0 * obj::kPi / int(D)) * e2;
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        g[k] = i < D ? q * xl[i] + c * ::sin(2.0 * obj::kPi * xl[i]) : 0.0;
      }
    } else {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        g[k] = i < D ? 2.0 * xl[i] : 0.0;
      }
    }
  }
};

template <class F, int D, int TS, Gradient G, class Prov>
__global__ __launch_bounds__(kTeamThreads, 2) void
bfgs_team_kernel(ClassArgs a, Prov prov)
{
  static_assert(team_fits<D, TS>(), "team size for this dimension");
  static_assert(kWave % TS == 0, "whole teams per wavefront");
  constexpr int KC = (D + TS - 1) / TS;
  constexpr int kTeams = kTeamThreads / TS;
  using E = TeamEval<F, D, TS>;
  // per team: the point being evaluated, and the new gradient and pending
  // pair broadcast for the H pass
  __shared__ double lds_x[kTeams][D];
  __shared__ double lds_b[kTeams][3][D];

  const int team = threadIdx.x / TS, r = threadIdx.x % TS;
  double* xl = lds_x[team];
  double* gb = lds_b[team][0];
  double* sb = lds_b[team][1];
  double* ab = lds_b[team][2];
  const int64_t team_id = int64_t(blockIdx.x) * kTeams + team;
  const int64_t n_teams = int64_t(gridDim.x) * kTeams;

  for (int64_t s = team_id; s < a.total_starts; s += n_teams) {
    const int pi = find_problem(a.probs, a.n_probs, s);
    const ProblemDesc& P = a.probs[pi];
    Control* ctl = a.ctl + pi;
    decltype(auto) f = prov.template get<F>(P);
    const int64_t sl = s - P.start0;

    double x[KC], g[KC], p[KC], gn[KC]; // the trial point lives in LDS (xl)
    double H[KC][D];
    // the rank from an opaque copy, so that the identity rows below are
    // formed here for each start instead of being kept (spilled) across
    // the start loop
    int ro = r;
    asm volatile("" : "+v"(ro));
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int i = ro + TS * k;
      double v = 0.0;
      if (i < D) {
        if (a.starts)
          v = a.starts[s * D + i];
        else if (P.x0)
          v = P.x0[sl * D + i];
        else
          v = start_uniform_in(P, uint64_t(sl), i);
      }
      x[k] = v;
#pragma unroll
      for (int c = 0; c < D; ++c)
        H[k][c] = (i == c) ? 1.0 : 0.0;
    }
    auto stage = [&](const double (&v)[KC], double* dst) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        if (i < D)
          dst[i] = v[k];
      }
      team_exchange();
    };
    // trial point x + alpha p into the team's LDS slot
    auto stage_trial = [&](double alpha) {
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        if (i < D)
          xl[i] = ::fma(alpha, p[k], x[k]);
      }
      team_exchange();
    };
    auto dot = [&](const double (&u)[KC], const double (&v)[KC]) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < KC; ++k)
        t = ::fma(u[k], v[k], t);
      return team_sum<TS>(t);
    };

    stage(x, xl);
    double bestVal = E::value(xl, r);
    E::template gradient<G>(f, xl, r, g);
#pragma unroll
    for (int k = 0; k < KC; ++k)
      p[k] = -g[k];
    int grads = 1, ls_evals = 0;
    int status = kStatusUnset;
    double fval = 333777.0, gnorm = 69.0;
    int iter = 0;
    bool done = false;

    for (; iter < P.max_iter; ++iter) {
      team_exchange(); // the previous evaluation's reads of xl are done
      if (load_flag(&ctl->stop_flag) != 0) {
        status = kStatusStopped;
        fval = bestVal;
        gnorm = ::sqrt(dot(g, g));
        done = true;
        break;
      }
      const double ddir = dot(g, p);
      double alpha = 1.0, f1 = 0.0;
      bool accepted = false;
      for (int t = 0; t < kLineSearchTrials; ++t) {
        stage_trial(alpha);
        f1 = E::value(xl, r);
        ++ls_evals;
        if (f1 <= ::fma(kArmijoC1 * alpha, ddir, bestVal) && f1 < bestVal) {
          accepted = true;
          break;
        }
        alpha *= 0.5;
        team_exchange();
      }
      double fnew = f1;
      if (!accepted) {
        stage_trial(alpha);
        fnew = E::value(xl, r);
        ++ls_evals;
      }
      E::template gradient<G>(f, xl, r, gn);
      ++grads;

      // v = H g_new
      stage(gn, gb);
      double v[KC];
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c)
          acc = ::fma(H[k][c], gb[c], acc);
        v[k] = acc;
      }
      // s = xt - x, y = gn - g, u = H y = v + p
      double sc[KC];
      double t0 = 0.0, t1 = 0.0;
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        sc[k] = i < D ? xl[i] - x[k] : 0.0;
        const double y = gn[k] - g[k];
        t0 = ::fma(sc[k], y, t0);
        t1 = ::fma(y, v[k] + p[k], t1);
      }
      const double sTy = team_sum<TS>(t0);
      const double yu = team_sum<TS>(t1);
      const bool upd = !(::fabs(sTy) < kCurvatureSkip);
      const bool acc = fnew < bestVal;
      double ac[KC];
      if (upd) {
        const double rho = 1.0 / sTy;
        const double beta = ::fma(rho * rho, yu, rho);
        const double hb = 0.5 * beta;
        double q0 = 0.0, q1 = 0.0;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          const double u = v[k] + p[k];
          ac[k] = ::fma(hb, sc[k], -(rho * u));
          const double ga = acc ? gn[k] : g[k];
          q0 = ::fma(ac[k], ga, q0);
          q1 = ::fma(sc[k], ga, q1);
        }
        q0 = team_sum<TS>(q0);
        q1 = team_sum<TS>(q1);
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          const double hg = acc ? v[k] : -p[k];
          p[k] = -::fma(ac[k], q1, ::fma(sc[k], q0, hg));
        }
      } else {
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          ac[k] = 0.0;
          p[k] = acc ? -v[k] : p[k];
        }
      }
      if (acc) {
        bestVal = fnew;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          x[k] = (r + TS * k) < D ? xl[r + TS * k] : 0.0;
          g[k] = gn[k];
        }
      }
      if (upd) { // H += s a^T + a s^T, each entry as s_r a_c + a_r s_c
        stage(sc, sb);
        stage(ac, ab);
#pragma unroll
        for (int k = 0; k < KC; ++k)
#pragma unroll
          for (int c = 0; c < D; ++c)
            H[k][c] = H[k][c] + (sc[k] * ab[c] + ac[k] * sb[c]);
      }
      const double nrm = ::sqrt(dot(g, g));
      if (!::isfinite(nrm) || !::isfinite(fnew)) {
        status = kStatusNonFinite;
        fval = fnew;
        gnorm = nrm;
        done = true;
        break;
      }
      if (nrm < P.tol) {
        if (r == 0)
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
      gnorm = ::sqrt(dot(g, g));
    }
    double e2 = 0.0;
    if (status == kStatusConverged) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int i = r + TS * k;
        if (i < D) {
          const double df = x[k] - known_min(P.min_kind, i);
          t = ::fma(df, df, t);
        }
      }
      e2 = team_sum<TS>(t);
    }
    if (r == 0) {
      a.out.status[s] = status;
      a.out.iter[s] = iter;
      a.out.fval[s] = fval;
      a.out.gnorm[s] = gnorm;
      a.out.ls_evals[s] = ls_evals;
      a.out.grads[s] = grads;
      a.out.near[s] = status == kStatusConverged && ::sqrt(e2) < 0.5;
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const int i = r + TS * k;
      if (i < D)
        a.out.x[s * D + i] = x[k];
    }
    team_exchange(); // LDS reuse by the team's next start
  }
}

} // namespace zeus::dev
