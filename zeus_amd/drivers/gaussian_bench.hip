// gaussian_bench.hip -- the 150-D Gaussian multistart through the C++ front
// end, as a reference user calls it: zeus::Zeus(Gaussian<150>{C}, -5, 5,
// 4096, 10000, 10, 100, "gaussian", 1e-8, 42, run) (reference
// include/zeus.cuh:341-375, examples/gaussian.hpp:8-34).  zeus::Gaussian is
// a built-in objective, so the call runs the library's MFMA PSO and tile
// BFGS kernels.
//
//   zeus_gaussian_bench [calls=20] [warmup=3] [philox|xorwow]
//
// Prints one JSON line: host wall time per call (steady_clock around each
// call, results table off), converged starts per call and the library's
// tuned form for the class.
#include "zeus/zeus.hpp"

#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

int
main(int argc, char** argv)
{
  constexpr std::size_t D = 150;
  const int calls = argc > 1 ? std::atoi(argv[1]) : 20;
  const int warmup = argc > 2 ? std::atoi(argv[2]) : 3;
  const bool xorwow = argc > 3 && std::strcmp(argv[3], "xorwow") == 0;
  if (calls < 1 || warmup < 0) {
    std::fprintf(stderr, "usage: zeus_gaussian_bench [calls>=1] [warmup>=0] [philox|xorwow]\n");
    return 2;
  }
  const zeus::PRNGType prng = xorwow ? zeus::PRNGType::XORWOW : zeus::PRNGType::PHILOX;
  // C_ii = 1 + (D - 1) 0.5, C_ij = 0.5
  auto C = std::make_unique<std::array<std::array<double, D>, D>>();
  for (std::size_t i = 0; i < D; ++i)
    for (std::size_t j = 0; j < D; ++j)
      (*C)[i][j] = i == j ? 1.0 + (D - 1) * 0.5 : 0.5;
  const zeus::Gaussian<D> f(*C);
  zeus::set_results_table(false);

  auto call = [&](int run) {
    return zeus::Zeus(f, -5.0, 5.0, 4096, 10000, 10, 100, "gaussian", 1e-8, 42, run, true, prng);
  };
  for (int i = 0; i < warmup; ++i)
    (void)call(i);
  long long converged = 0;
  int bad = 0;
  double best = 1e300;
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < calls; ++i) {
    const auto r = call(i);
    converged += r.c.claimed;
    bad += r.status != 1 || r.c.claimed < 100;
    best = r.fval < best ? r.fval : best;
  }
  const double s =
    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  int form = 0, threads = 0;
  (void)zeus_tuned_config(ZEUS_OBJ_GAUSSIAN, int(D), ZEUS_GRAD_ADJOINT, &form, &threads);
  std::printf("{\"ms_per_call\": %.6f, \"converged_per_s\": %.1f, \"converged_per_call\": %.1f, "
              "\"calls\": %d, \"prng\": \"%s\", \"tuned_form\": %d, \"best_fval\": %.6g, "
              "\"bad_calls\": %d}\n",
              1e3 * s / calls, double(converged) / s, double(converged) / calls, calls,
              xorwow ? "xorwow" : "philox", form, best, bad);
  return bad == 0 ? 0 : 1;
}
