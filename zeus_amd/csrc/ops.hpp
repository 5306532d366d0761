// ops.hpp -- the launch table of one (objective, D) class, as a template that
// each instantiation unit expands for its dimensions.
#pragma once

#include "runtime.hpp"

namespace zeus::rt {

template <class F, int D, Gradient G>
__global__ void
eval_kernel(dev::ProblemDesc P, const double* x, int64_t n, double* f, double* g)
{
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const F fn = dev::BuiltinProvider{}.template get<F>(P);
  f[i] = dev::lane_value<F, D>(fn, x + i * D);
  if (g)
    dev::lane_gradient<F, D, G>(fn, x + i * D, g + i * D);
}

template <class F, int D>
struct Ops {
  static constexpr bool kMfmaPso = F::id == obj::kGaussian && D >= 16;
  static constexpr bool kThreadOk = F::id != obj::kMlp && D <= 64;
  // workgroup sizes of the group form (the MLP's cooperative evaluator puts
  // one point per lane and one hidden-to-hidden weight per lane: 256)
  static constexpr int kNGroup = F::id == obj::kMlp ? 1 : group_nt_count(D);
  static constexpr int nt_at(int k) { return F::id == obj::kMlp ? 256 : group_nt_at(D, k); }
  // tile form: Gaussian classes whose PSO already runs in tiles of 16
  static constexpr bool kTileOk = kMfmaPso;
  // PSO with a workgroup per particle: objectives whose one evaluation is
  // worth a workgroup (the MLP loss over its data points)
  static constexpr bool kGroupPso = F::id == obj::kMlp;
  using PsoEval = typename dev::GroupEvalSelect<F, 256, Gradient::Adjoint>::type;
  static constexpr int kPsoPpb = kMfmaPso ? dev::kPsoTile : kGroupPso ? dev::kPsoGroupPpb : 256;

  static hipError_t pso(const dev::PsoArgs& a, int nb, hipStream_t s)
  {
    if (fault_take_launch())
      return hipErrorLaunchFailure;
    if constexpr (kMfmaPso)
      dev::pso_gaussian_mfma_kernel<D><<<nb, dev::kPsoTileThreads, 0, s>>>(a);
    else if constexpr (kGroupPso)
      dev::pso_group_kernel<F, D, 256, kPsoPpb, PsoEval, dev::BuiltinProvider>
        <<<nb, 256, 0, s>>>(a, {});
    else
      dev::pso_lane_kernel<F, D, dev::BuiltinProvider><<<nb, 256, 0, s>>>(a, {});
    return hipGetLastError();
  }

  // gBest after each problem's last iteration (one workgroup per problem)
  static hipError_t pso_final(const dev::PsoArgs& a, int n_probs, hipStream_t s)
  {
    if (fault_take_launch())
      return hipErrorLaunchFailure;
    dev::pso_final_kernel<256, kPsoPpb, D><<<n_probs, 256, 0, s>>>(a);
    return hipGetLastError();
  }

  static hipError_t bfgs_thread(const dev::ClassArgs& a, int gradient, hipStream_t s)
  {
    if constexpr (!kThreadOk) {
      return hipErrorInvalidConfiguration;
    } else {
      if (fault_take_launch())
        return hipErrorLaunchFailure;
      const int nb = int((a.total_starts + 255) / 256);
      if (gradient == 1)
        dev::bfgs_thread_kernel<F, D, Gradient::Dual, dev::BuiltinProvider>
          <<<nb, 256, 0, s>>>(a, {});
      else
        dev::bfgs_thread_kernel<F, D, Gradient::Adjoint, dev::BuiltinProvider>
          <<<nb, 256, 0, s>>>(a, {});
      return hipGetLastError();
    }
  }

  template <int NT>
  static constexpr bool kHReg = dev::RowChunk<D, NT>::kPerThread <= 50;

  template <int NT, Gradient G>
  static auto group_kernel()
  {
    return &dev::bfgs_group_kernel<F, D, NT, G, kHReg<NT>, dev::BuiltinProvider>;
  }

  template <int NT>
  static hipError_t group_nt(const dev::ClassArgs& a, int gradient, int grid, hipStream_t s)
  {
    if (gradient == 1)
      group_kernel<NT, Gradient::Dual>()<<<grid, NT, 0, s>>>(a, {});
    else
      group_kernel<NT, Gradient::Adjoint>()<<<grid, NT, 0, s>>>(a, {});
    return hipGetLastError();
  }

  // team form: separable built-in objectives whose H rows fit a lane
  static constexpr bool kTeamObj = F::id == obj::kRosenbrock || F::id == obj::kRastrigin ||
                                   F::id == obj::kAckley || F::id == obj::kSphere;
  template <int TS>
  static constexpr bool kTeamOk = kTeamObj && dev::team_fits<D, TS>();

  template <int TS, Gradient G>
  static auto team_kernel()
  {
    return &dev::bfgs_team_kernel<F, D, TS, G, dev::BuiltinProvider>;
  }
  template <int TS>
  static hipError_t team_ts(const dev::ClassArgs& a, int gradient, int grid, hipStream_t s)
  {
    if constexpr (!kTeamOk<TS>) {
      return hipErrorInvalidConfiguration;
    } else {
      if (gradient == 1)
        team_kernel<TS, Gradient::Dual>()<<<grid, dev::kTeamThreads, 0, s>>>(a, {});
      else
        team_kernel<TS, Gradient::Adjoint>()<<<grid, dev::kTeamThreads, 0, s>>>(a, {});
      return hipGetLastError();
    }
  }
  static hipError_t bfgs_team(const dev::ClassArgs& a, int gradient, int ts, int grid,
                              hipStream_t s)
  {
    if (fault_take_launch())
      return hipErrorLaunchFailure;
    if (ts == 4)
      return team_ts<4>(a, gradient, grid, s);
    if (ts == 8)
      return team_ts<8>(a, gradient, grid, s);
    if (ts == 16)
      return team_ts<16>(a, gradient, grid, s);
    return hipErrorInvalidConfiguration;
  }
  template <int TS>
  static int team_occ(int gradient)
  {
    if constexpr (!kTeamOk<TS>) {
      return 1;
    } else {
      static int cached[2] = {0, 0}; // per gradient; queried once per process
      int& c = cached[gradient == 1];
      if (c == 0) {
        int n = 0;
        const hipError_t e =
          gradient == 1
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void*>(team_kernel<TS, Gradient::Dual>()),
                dev::kTeamThreads, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void*>(team_kernel<TS, Gradient::Adjoint>()),
                dev::kTeamThreads, 0);
        c = (e == hipSuccess && n > 0) ? n : 1;
      }
      return c;
    }
  }
  static int team_blocks_per_cu(int gradient, int ts)
  {
    return ts == 4 ? team_occ<4>(gradient) : ts == 8 ? team_occ<8>(gradient) : team_occ<16>(gradient);
  }

  static hipError_t bfgs_group(const dev::ClassArgs& a, int gradient, int nt, int grid,
                               hipStream_t s)
  {
    if (fault_take_launch())
      return hipErrorLaunchFailure;
    if (nt == nt_at(0))
      return group_nt<nt_at(0)>(a, gradient, grid, s);
    if constexpr (kNGroup > 1)
      if (nt == nt_at(1))
        return group_nt<nt_at(1)>(a, gradient, grid, s);
    if constexpr (kNGroup > 2)
      if (nt == nt_at(2))
        return group_nt<nt_at(2)>(a, gradient, grid, s);
    return hipErrorInvalidConfiguration;
  }

  template <int NT>
  static int occ(int gradient)
  {
    static int cached[2] = {0, 0}; // per gradient; queried once per process
    int& c = cached[gradient == 1];
    if (c == 0)
      c = occ_query<NT>(gradient);
    return c;
  }
  template <int NT>
  static int occ_query(int gradient)
  {
    int n = 0;
    hipError_t e;
    if (gradient == 1)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, reinterpret_cast<const void*>(group_kernel<NT, Gradient::Dual>()), NT, 0);
    else
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, reinterpret_cast<const void*>(group_kernel<NT, Gradient::Adjoint>()), NT, 0);
    return (e == hipSuccess && n > 0) ? n : 1;
  }

  static int group_blocks_per_cu(int gradient, int nt)
  {
    if (nt == nt_at(0))
      return occ<nt_at(0)>(gradient);
    if constexpr (kNGroup > 1)
      if (nt == nt_at(1))
        return occ<nt_at(1)>(gradient);
    if constexpr (kNGroup > 2)
      if (nt == nt_at(2))
        return occ<nt_at(2)>(gradient);
    return 1;
  }

  template <int NT>
  static constexpr int64_t slab_doubles()
  {
    return kHReg<NT> ? 0 : int64_t(dev::RowChunk<D, NT>::kPerThread) * NT;
  }
  static int64_t group_slab_per_block(int nt)
  {
    if (nt == nt_at(0))
      return slab_doubles<nt_at(0)>();
    if constexpr (kNGroup > 1)
      if (nt == nt_at(1))
        return slab_doubles<nt_at(1)>();
    if constexpr (kNGroup > 2)
      if (nt == nt_at(2))
        return slab_doubles<nt_at(2)>();
    return 0;
  }

  // tile form (bfgs_tile.hpp)
  static hipError_t bfgs_tile(const dev::TileArgs& a, int grid, hipStream_t s)
  {
    if constexpr (!kTileOk) {
      return hipErrorInvalidConfiguration;
    } else {
      if (fault_take_launch())
        return hipErrorLaunchFailure;
      dev::bfgs_tile_kernel<D><<<grid, dev::kTileThreads, 0, s>>>(a);
      return hipGetLastError();
    }
  }
  static int tile_blocks_per_cu()
  {
    if constexpr (!kTileOk) {
      return 1;
    } else {
      static int cached = 0;
      if (cached == 0) {
        int n = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &n, reinterpret_cast<const void*>(&dev::bfgs_tile_kernel<D>), dev::kTileThreads, 0);
        cached = (e == hipSuccess && n > 0) ? n : 1;
      }
      return cached;
    }
  }

  static hipError_t eval(const dev::ProblemDesc& P, const double* x, int64_t n, double* f,
                         double* g, int gradient, hipStream_t s)
  {
    if (fault_take_launch())
      return hipErrorLaunchFailure;
    const int nb = int((n + 127) / 128);
    if (nb == 0)
      return hipSuccess;
    if (gradient == 1)
      eval_kernel<F, D, Gradient::Dual><<<nb, 128, 0, s>>>(P, x, n, f, g);
    else
      eval_kernel<F, D, Gradient::Adjoint><<<nb, 128, 0, s>>>(P, x, n, f, g);
    return hipGetLastError();
  }

  static const ClassOps* table()
  {
    static const ClassOps t = [] {
      ClassOps o;
      o.objective = F::id;
      o.dim = D;
      o.pso_ppb = kPsoPpb;
      o.pso = &pso;
      o.pso_final = &pso_final;
      o.thread_ok = kThreadOk;
      o.bfgs_thread = &bfgs_thread;
      o.thread_slab_per_start = D <= dev::kThreadRegMaxD ? 0 : int64_t(D) * (D + 1) / 2;
      o.n_group = kNGroup;
      for (int k = 0; k < o.n_group; ++k)
        o.group_nt[k] = nt_at(k);
      o.bfgs_group = &bfgs_group;
      o.group_blocks_per_cu = &group_blocks_per_cu;
      o.group_slab_per_block = &group_slab_per_block;
      o.eval = &eval;
      if constexpr (kTeamOk<4>)
        o.team_ts[o.n_team++] = 4;
      if constexpr (kTeamOk<8>)
        o.team_ts[o.n_team++] = 8;
      if constexpr (kTeamOk<16>)
        o.team_ts[o.n_team++] = 16;
      o.bfgs_team = &bfgs_team;
      o.team_blocks_per_cu = &team_blocks_per_cu;
      o.tile_ok = kTileOk;
      o.bfgs_tile = &bfgs_tile;
      o.tile_blocks_per_cu = &tile_blocks_per_cu;
      o.tile_slot_doubles = kTileOk ? dev::TileShape<D>::kSlotDoubles : 0;
      return o;
    }();
    return &t;
  }
};

} // namespace zeus::rt
