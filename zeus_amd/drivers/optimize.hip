// optimize.hip -- the reference's example optimization drivers
// (examples/optimize_{rosenbrock,rastrigin,ackley,goldstein_price,
// himmelblau}.cu) as one program over the C++ front end.
//
//   zeus_optimize <function> <lower> <upper> <max_iter> <pso_iters> <converged>
//                 <num_optimizations> <tolerance> <seed> <run> [options]
//   optimize_<function> <lower> <upper> ...        (same program, by name)
//
// Each call prints the banner, the parameters and the result block of the
// reference drivers, and (through zeus::Zeus) appends its row to
// zeus_2d_results.tsv and writes the trajectory file when asked.
#include "cli.hpp"
#include "zeus/zeus.hpp"

#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

// examples/goldstein_price.hpp: defined for two dimensions
struct GoldsteinPrice {
  static constexpr std::size_t arity = 2;
  template <class T>
  __host__ __device__ T operator()(const std::array<T, 2>& x) const
  {
    using ::pow;
    using dual::pow;
    const T x1 = x[0], x2 = x[1];
    const T a = T(1.0) + pow(x1 + x2 + T(1.0), 2.0) *
                           (T(19.0) - T(14.0) * x1 + T(3.0) * pow(x1, 2.0) - T(14.0) * x2 +
                            T(6.0) * x1 * x2 + T(3.0) * pow(x2, 2.0));
    const T b = T(30.0) + pow(T(2.0) * x1 - T(3.0) * x2, 2.0) *
                            (T(18.0) - T(32.0) * x1 + T(12.0) * pow(x1, 2.0) + T(48.0) * x2 -
                             T(36.0) * x1 * x2 + T(27.0) * pow(x2, 2.0));
    return a * b;
  }
};

// examples/himmelblau.hpp
struct Himmelblau {
  static constexpr std::size_t arity = 2;
  template <class T>
  __host__ __device__ T operator()(const std::array<T, 2>& x) const
  {
    using ::pow;
    using dual::pow;
    const T x1 = x[0], x2 = x[1];
    return pow(x1 * x1 + x2 - T(11.0), 2.0) + pow(x1 + x2 * x2 - T(7.0), 2.0);
  }
};

template <class F>
int
drive(const F& f, const char* title, const char* fun_name, int argc, char** argv)
{
  zeus_examples::OptimizationParams p;
  if (!zeus_examples::parse_args(argc, argv, p))
    return 1;
  zeus_examples::print_banner(title, 2);
  zeus_examples::print_params(p);
  util::setStackSize();
  const auto g = p.adjoint ? zeus::Gradient::Adjoint : zeus::Gradient::Dual;
  auto r = zeus::Zeus(f, p.lower_bound, p.upper_bound, std::size_t(p.num_optimizations),
                      p.max_iterations, p.pso_iterations, p.required_converged, fun_name,
                      p.tolerance, p.seed, p.run_id, p.parallel, p.prng_type, p.trajectory_file,
                      p.nzerosteps, g);
  zeus_examples::print_result(r.status, r.fval, r.gradientNorm, r.iter, r.ms_opt, r.coordinates,
                              2);
  return 0;
}

} // namespace

int
main(int argc, char** argv)
{
  std::string prog = argv[0];
  const size_t slash = prog.find_last_of('/');
  if (slash != std::string::npos)
    prog = prog.substr(slash + 1);
  std::string fn;
  if (prog.rfind("optimize_", 0) == 0) {
    fn = prog.substr(9);
  } else if (argc > 1) {
    fn = argv[1];
    argv[1] = argv[0];
    ++argv;
    --argc;
  }
  if (fn == "rosenbrock")
    return drive(zeus::obj::Rosenbrock<2>{}, "Rosenbrock Function Optimization", "rosenbrock",
                 argc, argv);
  if (fn == "rastrigin")
    return drive(zeus::obj::Rastrigin<2>{}, "Rastrigin Function Optimization", "rastrigin",
                 argc, argv);
  if (fn == "ackley")
    return drive(zeus::obj::Ackley<2>{}, "Ackley Function Optimization", "ackley", argc, argv);
  if (fn == "goldstein_price")
    return drive(GoldsteinPrice{}, "Goldstein-Price Function Optimization", "goldstein_price",
                 argc, argv);
  if (fn == "himmelblau")
    return drive(Himmelblau{}, "Himmelblau Function Optimization", "himmelblau", argc, argv);
  std::fprintf(stderr,
               "Usage: %s <rosenbrock|rastrigin|ackley|goldstein_price|himmelblau> <lower_bound> "
               "<upper_bound> <max_iter> <pso_iters> <converged> <num_optimizations> "
               "<tolerance> <seed> <run> [options]\n",
               argv[0]);
  return 1;
}
