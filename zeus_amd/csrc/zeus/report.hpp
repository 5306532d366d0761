// report.hpp -- the reference's host reporting helpers for C++ callers:
// util::OptimizationParams, calculateEuclideanError, appendResultsToTsv and
// writeTrajectoryData (reference include/util.hpp:57-149, src/util.cpp:
// 11-41, src/utils.cu:126-286).  The formatting lives in the library
// (zeus_c.h: zeus_append_results_tsv, zeus_write_trajectory); these wrappers
// keep the reference's names, argument lists and side effects.
#pragma once

#include "result.hpp"
#include "zeus_c.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <string_view>

namespace util {

struct OptimizationParams {
  double lower_bound = 0.0;
  double upper_bound = 0.0;
  int N = 0;
  int MAX_ITER = 0;
  int PSO_ITER = 0;
  int requiredConverged = 0;
  double tolerance = 0.0;
  int seed = 0;
  int run = 0;
  bool parallel = false;
  std::string fun_name;
};

inline double
calculateEuclideanError(const std::string fun_name, const double* coordinates, const int dim)
{
  return zeus_euclidean_error(fun_name.c_str(), coordinates, dim);
}

inline zeus_metrics
to_abi(const Metrics& m)
{
  return zeus_metrics{m.ms_per_call, m.calls_per_thread_mean, m.fraction_of_kernel,
                      m.block95,     m.serialized,            m.fraction_of_thread};
}

// Appends to zeus_<dim>d_results.tsv and prints the convergence counts, as
// the reference does after every call.
template <std::size_t DIM>
void
appendResultsToTsv(const int dim, const int N, const std::string fun_name, float ms_init,
                   float ms_pso, float ms_opt, float ms_rand, const int max_iter,
                   const int pso_iter, const double error, const int run,
                   const zeus::Result<DIM>& best)
{
  zeus_result r{};
  r.idx = best.idx;
  r.status = best.status;
  r.iter = best.iter;
  r.fval = best.fval;
  r.gradient_norm = best.gradientNorm;
  r.claimed = best.c.claimed;
  r.actual = best.c.actual;
  r.surrendered = best.c.surrendered;
  r.stopped = best.c.stopped;
  r.ad = to_abi(best.ad);
  r.bfgs = to_abi(best.bfgs);
  if (zeus_append_results_tsv(nullptr, dim, N, fun_name.c_str(), ms_init, ms_pso, ms_opt,
                              ms_rand, max_iter, pso_iter, error, run, &r,
                              best.coordinates) != ZEUS_OK) {
    std::fprintf(stderr, "Error opening file: zeus_%dd_results.tsv\n", dim);
    return;
  }
  std::printf("claimed: %d\tactual : %d\tsurrendered: %d\tstopped: %d\n", best.c.claimed,
              best.c.actual, best.c.surrendered, best.c.stopped);
}

// Trajectory arrays (host) to a .tsv file; hipSuccess or hipErrorUnknown
// (reference returns cudaSuccess / cudaErrorUnknown).
inline hipError_t
writeTrajectoryData(double* hostTrajectoryCoords, double* hostTrajectoryFval,
                    double* hostTrajectoryGrad, int8_t* hostStatus, double* hostAlpha,
                    const OptimizationParams& params, int DIM, std::string_view filename)
{
  const std::string f(filename);
  const int rc =
    zeus_write_trajectory(f.c_str(), hostTrajectoryCoords, hostTrajectoryFval,
                          hostTrajectoryGrad, hostStatus, hostAlpha, params.N, params.MAX_ITER,
                          DIM);
  if (rc == ZEUS_E_UNSUPPORTED)
    std::fprintf(stderr,
                 "Error: NetCDF4 support was disabled at build time. Cannot write %s\n",
                 f.c_str());
  else if (rc == ZEUS_E_ARG)
    std::fprintf(stderr,
                 "Error: unsupported trajectory file extension for %s. Use .tsv or .nc\n",
                 f.c_str());
  return rc == ZEUS_OK ? hipSuccess : hipErrorUnknown;
}

} // namespace util
