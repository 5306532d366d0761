// fit_dijet.hip -- dijet-spectrum fit program.
//
// Fits the four-parameter dijet shape mu(x) = a (1 - x)^b / x^(c + d ln x)
// to a binned mass spectrum by minimising the Poisson negative
// log-likelihood with zeus::Zeus over the box [0, 10]^4 (10 PSO iterations,
// 100 required convergences, tol 1e-8, seed 42), then reports the fit
// quality and writes the per-bin prediction table.  Command line, printed
// lines and the table's columns are those of the reference program
// (examples/fit_dijet_spectrum.cu:19-221), which scripts parse:
//
//   zeus_fit_dijet <num_optimizations> <max_bfgs_iter> <run_id> [--parallel]
//                  [--save-trajectories <filename>] [--prng <xorwow|philox|sobol>]
//                  [--nzerosteps <n>] [--data <path>] [--sqrt-s <units>] [--adjoint]
//
// Input: a text table with one header line and rows "bin_low bin_high
// counts"; a bin is placed at x = (bin_low + bin_high) / (2 sqrt_s), kept
// inside (0, 1) by 1e-12.  --data (default ../examples/simulation/
// dijet_spectrum.tsv), --sqrt-s (default 13.6) and --adjoint (hand-written
// adjoint gradient instead of dual numbers) are this program's options.
// LogLikelihood is a built-in objective: with --parallel and no trajectory
// the fit runs on the library's kernels for it.
#include "zeus/loglikelihood.hpp"
#include "zeus/zeus.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

constexpr double kEdge = 1e-12; // distance of x from 0 and 1, floor of mu

// A binned spectrum: edges, counts and the position x of each bin.
struct Spectrum {
  std::vector<double> lo, hi, x;
  std::vector<int> k;

  long long events() const
  {
    long long t = 0;
    for (int v : k)
      t += v;
    return t;
  }
};

// Reads the table at `path`; on failure returns false with `why` set.
bool
read_spectrum(const std::string& path, double sqrt_s, Spectrum& sp, std::string& why)
{
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) {
    why = "couldn't open: " + path;
    return false;
  }
  int c;
  while ((c = std::fgetc(f)) != EOF && c != '\n') // header line
    ;
  double lo, hi;
  int k;
  while (std::fscanf(f, "%lf %lf %d", &lo, &hi, &k) == 3) {
    sp.lo.push_back(lo);
    sp.hi.push_back(hi);
    sp.k.push_back(k);
    sp.x.push_back(std::clamp(0.5 * (lo + hi) / sqrt_s, kEdge, 1.0 - kEdge));
  }
  std::fclose(f);
  if (sp.k.empty()) {
    why = "empty or mismatched data in: " + path;
    return false;
  }
  return true;
}

// The fitted model at every bin and the two goodness-of-fit sums.
struct Prediction {
  std::vector<double> mu, residual, pull;
  double chi2 = 0.0, deviance = 0.0;
};

Prediction
predict(const Spectrum& sp, const double* p)
{
  Prediction pr;
  const std::size_t n = sp.k.size();
  pr.mu.resize(n);
  pr.residual.resize(n);
  pr.pull.resize(n);
  for (std::size_t i = 0; i < n; ++i) {
    const double x = sp.x[i], k = double(sp.k[i]);
    const double mu = p[0] * std::pow(1.0 - x, p[1]) / std::pow(x, p[2] + p[3] * std::log(x));
    const double m = std::max(mu, kEdge);
    pr.mu[i] = mu;
    pr.residual[i] = k - mu;
    pr.pull[i] = pr.residual[i] / std::sqrt(m);
    pr.chi2 += (k - m) * (k - m) / m;
    pr.deviance += k > 0 ? 2.0 * (k * std::log(k / m) - (k - m)) : 2.0 * m;
  }
  return pr;
}

bool
write_table(const char* path, const Spectrum& sp, const Prediction& pr)
{
  FILE* f = std::fopen(path, "w");
  if (!f)
    return false;
  std::fputs("bin_low\tbin_high\tx_center\ttrue_counts\tpred_mu\tpred_residual\tpred_pull\n", f);
  for (std::size_t i = 0; i < sp.k.size(); ++i)
    std::fprintf(f, "%.17g\t%.17g\t%.17g\t%d\t%.17g\t%.17g\t%.17g\n", sp.lo[i], sp.hi[i],
                 sp.x[i], sp.k[i], pr.mu[i], pr.residual[i], pr.pull[i]);
  return std::fclose(f) == 0;
}

struct Cli {
  std::size_t starts = 0;
  int max_iter = 0, run = 0, nzerosteps = 0;
  bool parallel = false, adjoint = false;
  zeus::PRNGType prng = zeus::PRNGType::XORWOW;
  std::string trajectory, data = "../examples/simulation/dijet_spectrum.tsv";
  double sqrt_s = 13.6;
};

// 0: parsed; 1: bad command line (the message is printed here).
int
parse(int argc, char** argv, Cli& o)
{
  const char* usage = " <num_optimizations> <max_bfgs_iter> <run_id> [--parallel] "
                      "[--save-trajectories <filename>] [--prng <xorwow|philox|sobol>] "
                      "[--nzerosteps <n>] [--data <path>] [--sqrt-s <units>] [--adjoint]\n";
  // the three positional numbers: whole decimal integers, N >= 0
  long long num[3] = {0, 0, 0};
  bool ok = argc >= 4;
  for (int j = 0; ok && j < 3; ++j) {
    char* end = nullptr;
    num[j] = std::strtoll(argv[1 + j], &end, 10);
    ok = end != argv[1 + j] && *end == '\0';
  }
  if (!ok || num[0] < 0) {
    std::fprintf(stderr, "Usage: %s%s", argv[0], usage);
    return 1;
  }
  o.starts = std::size_t(num[0]);
  o.max_iter = int(num[1]);
  o.run = int(num[2]);
  for (int i = 4; i < argc; ++i) {
    const std::string a = argv[i];
    const bool more = i + 1 < argc;
    if (a == "--parallel") {
      o.parallel = true;
    } else if (a == "--adjoint") {
      o.adjoint = true;
    } else if (a == "--save-trajectories" && more) {
      o.trajectory = argv[++i];
    } else if (a == "--nzerosteps" && more) {
      o.nzerosteps = std::atoi(argv[++i]);
    } else if (a == "--data" && more) {
      o.data = argv[++i];
    } else if (a == "--sqrt-s" && more) {
      o.sqrt_s = std::atof(argv[++i]);
    } else if (a == "--prng" && more) {
      const std::string v = argv[++i];
      if (v == "philox")
        o.prng = zeus::PRNGType::PHILOX;
      else if (v == "sobol")
        o.prng = zeus::PRNGType::SOBOL;
      else if (v != "xorwow")
        std::fprintf(stderr, "Unknown PRNG type: %s. Using default xorwow.\n", v.c_str());
    } else {
      std::fprintf(stderr, "Unknown argument: %s\n", a.c_str());
      return 1;
    }
  }
  return 0;
}

} // namespace

int
main(int argc, char** argv)
{
  Cli o;
  if (parse(argc, argv, o) != 0)
    return 1;
  util::setStackSize();

  Spectrum sp;
  std::string why;
  if (!read_spectrum(o.data, o.sqrt_s, sp, why)) {
    std::fprintf(stderr, "%s\n", why.c_str());
    return 1;
  }
  const auto [xmin, xmax] = std::minmax_element(sp.x.begin(), sp.x.end());
  std::printf("Loaded %zu bins; x in [%g, %g]\n", sp.k.size(), *xmin, *xmax);
  std::printf("\n#bins = %zu\ntotal events in the spectrum = %lld\n\n", sp.k.size(),
              sp.events());

  const LogLikelihood nll(sp.k, sp.x);
  const auto r = zeus::Zeus(nll, 0.0, 10.0, o.starts, o.max_iter, 10, 100, "poisson", 1e-8, 42,
                            o.run, o.parallel, o.prng, o.trajectory, o.nzerosteps,
                            o.adjoint ? zeus::Gradient::Adjoint : zeus::Gradient::Dual);
  const double* p = r.coordinates;
  std::printf("best NLL: %g\n", r.fval);
  std::printf("best-fit parameters: a=%g  b=%g  c=%g  d=%g\n\n", p[0], p[1], p[2], p[3]);

  const Prediction pr = predict(sp, p);
  const int ndf = std::max(1, int(sp.k.size()) - 4);
  std::printf("chi2/ndf=%g  dev/ndf=%g\n", pr.chi2 / ndf, pr.deviance / ndf);
  if (!write_table("dijet_fit_vs_data.tsv", sp, pr)) {
    std::fprintf(stderr, "could not open: dijet_fit_vs_data.tsv\n");
    return 1;
  }
  std::printf("predictions written to dijet_fit_vs_data.tsv\n");
  return 0;
}
