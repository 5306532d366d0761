// datagen.hip -- synthetic inputs of the benchmark configurations (host code).
//
// Every draw comes from the counter-based Philox stream of philox.hpp with a
// domain-separated key word 1, so the CPU oracle, the GPU path and any rerun
// see identical data:
//   dijet counts  key (seed, 0x80000000 | experiment), start = bin, draws 0..
//   MLP data      key (seed, 0x90000000), start = point, draws 0 (x), 1-2 (noise)
//   SPD matrices  key (seed, 0xA0000000 ^ stream), start = row, draws 2j, 2j+1
#include "zeus_c.h"

#include "zeus/philox.hpp"

#include <cmath>
#include <vector>

namespace {

using zeus::rng::uniform;

constexpr double kPi = 3.14159265358979323846;

// Poisson(mu) from one start's stream; *draw advances.  Inversion by
// sequential search for mu < 10; above, Hoermann's transformed rejection
// with squeeze (PTRS), W. Hoermann, "The transformed rejection method for
// generating Poisson random variables", Insurance: Mathematics and
// Economics 12 (1993) 39-45.
double
poisson(double mu, uint32_t seed, uint32_t key1, uint64_t start, uint64_t* draw)
{
  if (!(mu > 0.0))
    return 0.0;
  if (mu < 10.0) {
    const double u = uniform(seed, key1, start, (*draw)++);
    double p = std::exp(-mu), F = p;
    int k = 0;
    while (u > F && k < 1000) {
      ++k;
      p *= mu / k;
      F += p;
    }
    return double(k);
  }
  const double smu = std::sqrt(mu);
  const double b = 0.931 + 2.53 * smu;
  const double a = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (int guard = 0; guard < 100000; ++guard) {
    const double U = uniform(seed, key1, start, (*draw)++) - 0.5;
    const double V = uniform(seed, key1, start, (*draw)++);
    const double us = 0.5 - std::fabs(U);
    const double k = std::floor((2.0 * a / us + b) * U + mu + 0.43);
    if (us >= 0.07 && V <= vr)
      return k;
    if (k < 0.0 || (us < 0.013 && V > us))
      continue;
    if (std::log(V) + std::log(inv_alpha) - std::log(a / (us * us) + b) <=
        -mu + k * std::log(mu) - std::lgamma(k + 1.0))
      return k;
  }
  return std::floor(mu);
}

double
gauss(uint32_t seed, uint32_t key1, uint64_t start, uint64_t d0)
{
  // Box-Muller on two (0,1] uniforms
  const double u1 = uniform(seed, key1, start, d0);
  const double u2 = uniform(seed, key1, start, d0 + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

} // namespace

extern "C" {

// 100 bins of m in [1000, 7000] GeV, x = m / 13000 at the bin centre;
// mu_i = p0 (1 - x)^10 / x^(5 + 0.1 ln x) with p0 chosen for 1e6 expected
// events in total; k_i ~ Poisson(mu_i).
zeus_status_t
zeus_make_dijet(uint32_t seed, uint32_t experiment, int32_t n_bins, double* out)
{
  if (n_bins <= 0 || !out)
    return ZEUS_E_ARG;
  const double w = 6000.0 / n_bins;
  std::vector<double> shape(n_bins);
  double total = 0.0;
  for (int i = 0; i < n_bins; ++i) {
    const double m = 1000.0 + (i + 0.5) * w;
    const double x = m / 13000.0;
    out[i] = x;
    shape[i] = std::pow(1.0 - x, 10.0) / std::pow(x, 5.0 + 0.1 * std::log(x));
    total += shape[i];
  }
  const double p0 = 1e6 / total;
  const uint32_t key1 = 0x80000000u | experiment;
  for (int i = 0; i < n_bins; ++i) {
    uint64_t draw = 0;
    const double k = poisson(p0 * shape[i], seed, key1, uint64_t(i), &draw);
    out[n_bins + i] = k;
    out[2 * n_bins + i] = std::lgamma(k + 1.0);
  }
  return ZEUS_OK;
}

// m points: x ~ U(-1, 1], y = sin(3x) + 0.1 z, z ~ N(0, 1).
zeus_status_t
zeus_make_mlp_data(uint32_t seed, int32_t m, double* out)
{
  if (m <= 0 || !out)
    return ZEUS_E_ARG;
  const uint32_t key1 = 0x90000000u;
  for (int p = 0; p < m; ++p) {
    const double x = -1.0 + 2.0 * uniform(seed, key1, uint64_t(p), 0);
    out[p] = x;
    out[m + p] = std::sin(3.0 * x) + 0.1 * gauss(seed, key1, uint64_t(p), 1);
  }
  return ZEUS_OK;
}

// C = A A^T / D + I, A_ij ~ N(0, 1): symmetric positive definite.
zeus_status_t
zeus_make_spd(uint32_t seed, uint32_t stream, int32_t dim, double* c)
{
  if (dim <= 0 || !c)
    return ZEUS_E_ARG;
  const uint32_t key1 = 0xA0000000u ^ stream;
  std::vector<double> A(size_t(dim) * dim);
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j)
      A[size_t(i) * dim + j] = gauss(seed, key1, uint64_t(i), 2u * uint64_t(j));
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j) {
      double s = 0.0;
      for (int k = 0; k < dim; ++k)
        s += A[size_t(i) * dim + k] * A[size_t(j) * dim + k];
      c[size_t(i) * dim + j] = s / dim + (i == j ? 1.0 : 0.0);
    }
  return ZEUS_OK;
}

} // extern "C"
