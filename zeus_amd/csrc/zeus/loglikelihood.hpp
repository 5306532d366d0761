// loglikelihood.hpp -- the reference's LogLikelihood objective class
// (examples/simulation/fitting/LogLikelihood.h:11-82) for C++ callers:
// constructed from bin counts and bin centres, evaluated on the host and the
// device through operator() over std::array<T, 4> (double or dual), plus the
// hand-written adjoint of the built-in objective (obj::LogLikelihood).
//
// Poisson mean of bin i: a (1 - x_i)^b / x_i^(c + d ln x_i); the value is
// -sum_i [k_i ln mu_i - mu_i - lgamma(k_i + 1)].  The lgamma(k_i + 1) terms
// do not depend on the parameters and are computed once on the host here
// (the reference evaluates them in every call).
#pragma once

#include "matrix.hpp"
#include "objectives.hpp"
#include "zeus_c.h"

#include <array>
#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <vector>

class LogLikelihood {
public:
  static constexpr std::size_t arity = 4;
  // a built-in objective (zeus/builtins.hpp): its parameter block is the
  // library's ZEUS_OBJ_LOGLIK layout [x (n) | k (n) | lgamma(k + 1) (n)]
  static constexpr int zeus_objective = ZEUS_OBJ_LOGLIK;
  const double* zeus_params() const { return P_.deviceData(); }
  long long zeus_n_params() const { return static_cast<long long>(3 * n_); }

  LogLikelihood(std::vector<int> const& bin_counts, std::vector<double> const& bin_centers)
    : n_(bin_counts.size()), P_(3 * bin_counts.size(), 1)
  {
    if (bin_counts.size() != bin_centers.size() || bin_counts.empty())
      throw std::invalid_argument("LogLikelihood: counts and centers must be equal, non-empty");
    std::vector<double> p(3 * n_);
    for (std::size_t i = 0; i < n_; ++i) {
      p[i] = bin_centers[i];
      p[n_ + i] = double(bin_counts[i]);
      p[2 * n_ + i] = std::lgamma(double(bin_counts[i]) + 1.0);
    }
    P_.set(p.data(), p.size()); // one host -> device copy
  }

  template <typename T>
  __host__ __device__ T operator()(std::array<T, 4> const& theta) const
  {
    return view()(theta);
  }

  __host__ __device__ double value_and_gradient(const double* theta, double* g) const
  {
    return view().value_and_gradient(theta, g);
  }

  std::size_t size() const noexcept { return n_; }

private:
  __host__ __device__ zeus::obj::LogLikelihood view() const
  {
    zeus::obj::LogLikelihood v;
    v.P = P_.data(); // device copy on the device, host copy on the host
    v.n = int(n_);
    return v;
  }

  std::size_t n_;
  zeus::Matrix<double> P_;
};
