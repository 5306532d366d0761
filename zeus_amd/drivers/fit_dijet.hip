// fit_dijet.hip -- the reference's dijet-spectrum fit driver
// (examples/fit_dijet_spectrum.cu:19-221): load a three-column TSV
// (bin_low bin_high counts, one header line), convert bin centres to
// x = m / sqrt_s, fit the four-parameter Poisson likelihood over [0, 10]^4
// with 10 PSO iterations, 100 required convergences, tol 1e-8, seed 42, and
// report the post-fit chi2/ndf and deviance/ndf; per-bin predictions go to
// dijet_fit_vs_data.tsv.
//
//   zeus_fit_dijet <num_optimizations> <max_bfgs_iter> <run_id> [--parallel]
//                  [--save-trajectories <filename>] [--prng <xorwow|philox|sobol>]
//                  [--nzerosteps <n>]
//                  [--data <path>] [--sqrt-s <units>] [--adjoint]
//
// --data (default: the reference's ../examples/simulation/dijet_spectrum.tsv,
// which the reference does not ship), --sqrt-s (default 13.6, the
// reference's value) and --adjoint are this build's options.
#include "zeus/loglikelihood.hpp"
#include "zeus/zeus.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

void
load_dijet_spectrum_as_x(const std::string& path, std::vector<double>& bin_low,
                         std::vector<double>& bin_high, std::vector<int>& counts,
                         std::vector<double>& x_centers, double sqrt_s_units)
{
  bin_low.clear();
  bin_high.clear();
  counts.clear();
  x_centers.clear();
  std::ifstream in(path);
  if (!in)
    throw std::runtime_error("couldn't open: " + path);
  std::string header;
  std::getline(in, header);
  double lo, hi;
  int k;
  constexpr double eps = 1e-12;
  while (in >> lo >> hi >> k) {
    const double x = std::min(std::max(0.5 * (lo + hi) / sqrt_s_units, eps), 1.0 - eps);
    bin_low.push_back(lo);
    bin_high.push_back(hi);
    counts.push_back(k);
    x_centers.push_back(x);
  }
  if (counts.empty())
    throw std::runtime_error("empty or mismatched data in: " + path);
  const auto [mn, mx] = std::minmax_element(x_centers.begin(), x_centers.end());
  std::cout << "Loaded " << counts.size() << " bins; x in [" << *mn << ", " << *mx << "]\n";
}

void
write_pred_vs_true_tsv(const std::string& path, const std::vector<double>& bin_low,
                       const std::vector<double>& bin_high, const std::vector<int>& counts,
                       const std::vector<double>& x, const std::vector<double>& mu,
                       const std::vector<double>& residual, const std::vector<double>& pull)
{
  FILE* out = std::fopen(path.c_str(), "w");
  if (!out)
    throw std::runtime_error("could not open: " + path);
  std::fputs("bin_low\tbin_high\tx_center\ttrue_counts\tpred_mu\tpred_residual\tpred_pull\n",
             out);
  for (std::size_t i = 0; i < counts.size(); ++i)
    std::fprintf(out, "%.17g\t%.17g\t%.17g\t%d\t%.17g\t%.17g\t%.17g\n", bin_low[i], bin_high[i],
                 x[i], counts[i], mu[i], residual[i], pull[i]);
  std::fclose(out);
}

int
usage(const char* prog)
{
  std::cerr << "Usage: " << prog
            << " <num_optimizations> <max_bfgs_iter> <run_id> "
               "[--parallel] [--save-trajectories <filename>] [--prng <xorwow|philox|sobol>]"
               " [--nzerosteps <n>] [--data <path>] [--sqrt-s <units>] [--adjoint]\n";
  return 1;
}

} // namespace

int
main(int argc, char* argv[])
{
  if (argc < 4)
    return usage(argv[0]);
  std::size_t N = 0;
  int bfgs = 0, run = 0;
  try {
    N = std::stoul(argv[1]);
    bfgs = std::stoi(argv[2]);
    run = std::stoi(argv[3]);
  } catch (const std::exception&) {
    return usage(argv[0]);
  }
  std::string trajectory_file, data = "../examples/simulation/dijet_spectrum.tsv";
  double sqrt_s = 13.6;
  zeus::PRNGType prng = zeus::PRNGType::XORWOW;
  bool parallel = false, adjoint = false;
  int nzerosteps = 0;
  for (int i = 4; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--parallel") {
      parallel = true;
    } else if (a == "--save-trajectories" && i + 1 < argc) {
      trajectory_file = argv[++i];
    } else if (a == "--prng" && i + 1 < argc) {
      const std::string v = argv[++i];
      if (v == "xorwow")
        prng = zeus::PRNGType::XORWOW;
      else if (v == "philox")
        prng = zeus::PRNGType::PHILOX;
      else if (v == "sobol")
        prng = zeus::PRNGType::SOBOL;
      else
        std::cerr << "Unknown PRNG type: " << v << ". Using default xorwow.\n";
    } else if (a == "--nzerosteps" && i + 1 < argc) {
      nzerosteps = std::stoi(argv[++i]);
    } else if (a == "--data" && i + 1 < argc) {
      data = argv[++i];
    } else if (a == "--sqrt-s" && i + 1 < argc) {
      sqrt_s = std::stod(argv[++i]);
    } else if (a == "--adjoint") {
      adjoint = true;
    } else {
      std::cerr << "Unknown argument: " << a << "\n";
      return 1;
    }
  }
  util::setStackSize();

  std::vector<double> bin_low, bin_high, centers;
  std::vector<int> counts;
  try {
    load_dijet_spectrum_as_x(data, bin_low, bin_high, counts, centers, sqrt_s);
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
  std::cout << "\n#bins = " << counts.size() << "\n";
  long long total = 0;
  for (int k : counts)
    total += k;
  std::cout << "total events in the spectrum = " << total << "\n\n";

  LogLikelihood ll(counts, centers);
  const auto g = adjoint ? zeus::Gradient::Adjoint : zeus::Gradient::Dual;
  auto res = zeus::Zeus(ll, 0.00, 10.00, N, bfgs, 10, 100, "poisson", 1e-8, 42, run, parallel,
                        prng, trajectory_file, nzerosteps, g);
  std::cout << "best NLL: " << res.fval << "\n";
  const double a = res.coordinates[0], b = res.coordinates[1], c = res.coordinates[2],
               d = res.coordinates[3];
  std::cout << "best-fit parameters: a=" << a << "  b=" << b << "  c=" << c << "  d=" << d
            << "\n\n";

  // post-fit diagnostics (fit_dijet_spectrum.cu:189-214)
  constexpr double eps = 1e-12;
  const std::size_t n = centers.size();
  std::vector<double> mu(n), residual(n), pull(n);
  for (std::size_t i = 0; i < n; ++i) {
    const double x = std::min(std::max(centers[i], eps), 1.0 - eps);
    mu[i] = a * std::pow(1.0 - x, b) / std::pow(x, c + d * std::log(x));
    residual[i] = double(counts[i]) - mu[i];
    pull[i] = residual[i] / std::sqrt(std::max(mu[i], eps));
  }
  double chi2 = 0.0, dev = 0.0;
  for (std::size_t i = 0; i < n; ++i) {
    const double k = double(counts[i]);
    const double m = std::max(mu[i], eps);
    chi2 += (k - m) * (k - m) / m;
    dev += (k > 0) ? 2.0 * (k * std::log(k / m) - (k - m)) : 2.0 * m;
  }
  const int ndf = std::max<int>(1, int(n) - 4);
  std::cout << "chi2/ndf=" << chi2 / ndf << "  dev/ndf=" << dev / ndf << "\n";
  try {
    write_pred_vs_true_tsv("dijet_fit_vs_data.tsv", bin_low, bin_high, counts, centers, mu,
                           residual, pull);
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
  std::cout << "predictions written to dijet_fit_vs_data.tsv\n";
  return 0;
}
