// cli.hpp -- the command-line contract of the reference's example drivers
// (examples/optimization_utils.hpp and .cc:10-124):
//
//   <prog> <lower_bound> <upper_bound> <max_iter> <pso_iters> <converged>
//          <num_optimizations> <tolerance> <seed> <run>
//          [--parallel] [--save-trajectories <filename>]
//          [--prng <xorwow|philox|sobol>] [--nzerosteps <n>]
//
// plus one option of this build, --adjoint (use the objective's hand-written
// gradient instead of dual numbers).  Output lines and their formats follow
// the reference's print_banner / print_params / print_result.
#pragma once

#include "zeus/traits.hpp"

#include <cstdio>
#include <cstdlib>
#include <exception>
#include <iostream>
#include <string>

namespace zeus_examples {

struct OptimizationParams {
  double lower_bound = 0.0;
  double upper_bound = 0.0;
  int max_iterations = 0;
  int pso_iterations = 0;
  int required_converged = 0;
  int num_optimizations = 0;
  double tolerance = 0.0;
  int seed = 0;
  int run_id = 0;
  bool parallel = false;
  zeus::PRNGType prng_type = zeus::PRNGType::XORWOW;
  std::string trajectory_file;
  int nzerosteps = 0;
  bool adjoint = false; // this build's option
};

inline bool
parse_args(int argc, char* argv[], OptimizationParams& p)
{
  if (argc < 10) {
    std::cerr << "Usage: " << argv[0]
              << " <lower_bound> <upper_bound> <max_iter> <pso_iters> "
                 "<converged> <num_optimizations> <tolerance> <seed> <run> "
                 "[--parallel] [--save-trajectories <filename>] [--prng <xorwow|philox|sobol>]"
                 " [--nzerosteps <n>] [--adjoint]\n";
    return false;
  }
  try {
    p.lower_bound = std::atof(argv[1]);
    p.upper_bound = std::atof(argv[2]);
    p.max_iterations = std::stoi(argv[3]);
    p.pso_iterations = std::stoi(argv[4]);
    p.required_converged = std::stoi(argv[5]);
    p.num_optimizations = std::stoi(argv[6]);
    p.tolerance = std::stod(argv[7]);
    p.seed = std::stoi(argv[8]);
    p.run_id = std::stoi(argv[9]);
    for (int i = 10; i < argc; ++i) {
      const std::string a = argv[i];
      if (a == "--parallel") {
        p.parallel = true;
      } else if (a == "--save-trajectories" && i + 1 < argc) {
        p.trajectory_file = argv[++i];
      } else if (a == "--prng" && i + 1 < argc) {
        const std::string v = argv[++i];
        if (v == "xorwow")
          p.prng_type = zeus::PRNGType::XORWOW;
        else if (v == "philox")
          p.prng_type = zeus::PRNGType::PHILOX;
        else if (v == "sobol")
          p.prng_type = zeus::PRNGType::SOBOL;
        else
          std::cerr << "Unknown PRNG type: " << v << ". Using default xorwow.\n";
      } else if (a == "--nzerosteps" && i + 1 < argc) {
        p.nzerosteps = std::stoi(argv[++i]);
      } else if (a == "--adjoint") {
        p.adjoint = true;
      } else {
        std::cerr << "Unknown or incomplete argument: " << a << "\n";
        return false;
      }
    }
  } catch (const std::exception& e) {
    std::cerr << "Error parsing arguments: " << e.what() << "\n";
    return false;
  }
  return true;
}

inline void
print_params(const OptimizationParams& p)
{
  std::printf("Bounds: [%.10g, %.10g]\n", p.lower_bound, p.upper_bound);
  std::printf("Max iterations: %d\n", p.max_iterations);
  std::printf("PSO iterations: %d\n", p.pso_iterations);
  std::printf("Required converged: %d\n", p.required_converged);
  std::printf("Number of optimizations: %d\n", p.num_optimizations);
  std::printf("Tolerance: %.10g\n", p.tolerance);
  std::printf("Seed: %d\n", p.seed);
  std::printf("Run ID: %d\n", p.run_id);
  std::printf("Algorithm: %s\n", p.parallel ? "parallel" : "sequential");
  if (!p.trajectory_file.empty())
    std::printf("Trajectory file: %s\n", p.trajectory_file.c_str());
  if (p.nzerosteps > 0)
    std::printf("Max consecutive zero steps: %d\n", p.nzerosteps);
}

inline void
print_banner(const std::string& title, int dim)
{
  std::printf("\n=== %s (%dD) ===\n\n", title.c_str(), dim);
}

inline void
print_result(int status, double fval, double gradientNorm, int iter, double ms_opt,
             const double* coords, int dim)
{
  static const char* what[] = {"(Surrendered - max iterations reached)",
                               "(Converged)",
                               "(Stopped by flag)",
                               "(Device memory allocation failure)",
                               "(Device runtime error)",
                               "(Non-finite values encountered)",
                               "(Consecutive zero-step limit reached)"};
  std::printf("\n=== Optimization Result ===\n");
  std::printf("Status: %d %s\n", status, (status >= 0 && status <= 6) ? what[status] : "");
  std::printf("Function value: %e\n", fval);
  std::printf("Gradient norm: %e\n", gradientNorm);
  std::printf("Iterations: %d\n", iter);
  std::printf("Optimization time: %g ms\n", ms_opt);
  std::printf("Best coordinates: (");
  for (int i = 0; i < dim; ++i)
    std::printf("%s%.10g", i ? ", " : "", coords[i]);
  std::printf(")\n\n");
  std::fflush(stdout);
}

} // namespace zeus_examples
