// runtime.hpp -- host-side internals of libzeus_amd.so (not installed API).
#pragma once

#include "zeus/batch.hpp"
#include "zeus/bfgs_group.hpp"
#include "zeus/bfgs_team.hpp"
#include "zeus/bfgs_thread.hpp"
#include "zeus/bfgs_tile.hpp"
#include "zeus/prng.hpp"
#include "zeus/pso.hpp"
#include "zeus/summary.hpp"

#include <hip/hip_runtime.h>

namespace zeus::rt {

// Launch table of one class (objective type x compile-time dimension).
struct ClassOps {
  int objective = 0;
  int dim = 0;
  int pso_ppb = 256; // particles per PSO workgroup
  hipError_t (*pso)(const dev::PsoArgs&, int nblocks, hipStream_t) = nullptr;
  hipError_t (*pso_final)(const dev::PsoArgs&, int n_probs, hipStream_t) = nullptr;
  bool thread_ok = false;
  hipError_t (*bfgs_thread)(const dev::ClassArgs&, int gradient, hipStream_t) = nullptr;
  int64_t thread_slab_per_start = 0; // doubles of global H per start
  int n_group = 0;
  int group_nt[4] = {0, 0, 0, 0}; // offered workgroup sizes
  hipError_t (*bfgs_group)(const dev::ClassArgs&, int gradient, int nt, int grid,
                           hipStream_t) = nullptr;
  int (*group_blocks_per_cu)(int gradient, int nt) = nullptr;
  int64_t (*group_slab_per_block)(int nt) = nullptr; // doubles, 0 if resident
  // team form (several lanes per start, several starts per wavefront)
  int n_team = 0;
  int team_ts[3] = {0, 0, 0}; // offered team sizes
  hipError_t (*bfgs_team)(const dev::ClassArgs&, int gradient, int ts, int grid,
                          hipStream_t) = nullptr;
  int (*team_blocks_per_cu)(int gradient, int ts) = nullptr;
  // tile form (16 starts per workgroup, objective on MFMA; adjoint only)
  bool tile_ok = false;
  hipError_t (*bfgs_tile)(const dev::TileArgs&, int grid, hipStream_t) = nullptr;
  int (*tile_blocks_per_cu)() = nullptr;
  int64_t tile_slot_doubles = 0; // doubles of global memory per start slot
  hipError_t (*eval)(const dev::ProblemDesc&, const double* x, int64_t n, double* f,
                     double* g, int gradient, hipStream_t) = nullptr;
};

// Registry (one function per instantiation unit).
const ClassOps* find_class(int objective, int dim);
const ClassOps* ops_rosenbrock(int dim);
const ClassOps* ops_rastrigin(int dim);
const ClassOps* ops_ackley(int dim);
const ClassOps* ops_sphere(int dim);
const ClassOps* ops_gaussian(int dim);
const ClassOps* ops_loglik(int dim);
const ClassOps* ops_mlp(int dim);

// Test-only fault hooks (zeus_set_fault_hooks).
bool fault_take_launch();

// Workgroup sizes of the group form offered for a dimension.
constexpr int
group_nt_count(int D)
{
  return D <= 16 ? 1 : (D <= 64 ? 2 : 3);
}
constexpr int
group_nt_at(int D, int k)
{
  if (D <= 16)
    return 64;
  if (D <= 64)
    return k == 0 ? 64 : 256;
  return k == 0 ? 256 : (k == 1 ? 512 : 1024);
}

} // namespace zeus::rt
