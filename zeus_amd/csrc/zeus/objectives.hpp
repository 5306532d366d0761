// objectives.hpp -- the built-in objectives, as reference-compatible functors.
//
// Each objective is usable exactly like the reference's callables
// (operator() over std::array<T, D>, T = double or dual::DualNumber) and adds
//   * eval<T>(x): the same function over any indexable x (a plain pointer, an
//     array, or a seeded-dual view of a vector in LDS), so dual passes need
//     no per-lane std::array copy;
//   * value_and_gradient(x, g): a hand-written reverse-mode adjoint (value
//     plus full gradient for about the cost of two to three evaluations)
//     whose forward sweep repeats the value path operation for operation, so
//     the value it returns is bit-identical to eval<double>.
// Where the reference's double and dual overloads associate sums
// differently, both associations are kept (if constexpr on T).
//
// Parameter layouts (device pointers, row-major, FP64):
//   Gaussian      C (D x D)                          f = 0.5 x^T C x
//   LogLikelihood x (n), k (n), lgamma(k+1) (n)      Poisson NLL, D = 4
//   Mlp           x (P), y (P)                       1-16-16-1 tanh, MSE
#pragma once

#include "dual.hpp"
#include "traits.hpp"

#include <array>
#include <cmath>
#include <type_traits>

namespace zeus::obj {

using dual::DualNumber;

// Objective ids shared with the C ABI (zeus_c.h).
enum Id : int {
  kRosenbrock = 1,
  kRastrigin = 2,
  kAckley = 3,
  kGaussian = 4,
  kLogLikelihood = 5,
  kMlp = 6,
  kSphere = 7,
};

constexpr double kPi = 3.14159265358979323846;

// --- index views -----------------------------------------------------------
struct PtrView {
  const double* p;
  ZEUS_HD double operator[](int i) const { return p[i]; }
};
// x with dimension `seed` carrying derivative 1 (one forward-mode pass).
struct SeededView {
  const double* p;
  int seed;
  ZEUS_HD DualNumber operator[](int i) const
  {
    return DualNumber(p[i], i == seed ? 1.0 : 0.0);
  }
};
template <class T, std::size_t D>
struct ArrayView {
  const std::array<T, D>& a;
  ZEUS_HD T operator[](int i) const { return a[i]; }
};

template <class T>
inline constexpr bool kIsDouble = std::is_same_v<T, double>;

// --- Rosenbrock: examples/rosenbrock.hpp:6-41 --------------------------------
template <std::size_t D>
struct Rosenbrock {
  static constexpr std::size_t arity = D;
  static constexpr int id = kRosenbrock;
  // run by zeus::Zeus on the library's built-in kernels (no parameters)
  static constexpr int zeus_objective = kRosenbrock;
  const double* zeus_params() const { return nullptr; }
  long long zeus_n_params() const { return 0; }

  template <class T, class X>
  ZEUS_HD T eval(const X& x) const
  {
    T sum = T(0.0);
    for (int i = 0; i < int(D) - 1; ++i) {
      const T xi = x[i];
      const T t1 = T(1.0) - xi;
      const T t2 = x[i + 1] - xi * xi;
      if constexpr (kIsDouble<T>)
        sum += t1 * t1 + 100.0 * t2 * t2; // rosenbrock.hpp:27
      else
        sum = sum + t1 * t1 + T(100.0) * t2 * t2; // rosenbrock.hpp:14
    }
    return sum;
  }
  template <class T>
  ZEUS_HD T operator()(const std::array<T, D>& x) const
  {
    return eval<T>(ArrayView<T, D>{x});
  }
  ZEUS_HD double value_and_gradient(const double* x, double* g) const
  {
    for (int i = 0; i < int(D); ++i)
      g[i] = 0.0;
    double sum = 0.0;
    for (int i = 0; i < int(D) - 1; ++i) {
      const double t1 = 1.0 - x[i];
      const double t2 = x[i + 1] - x[i] * x[i];
      sum += t1 * t1 + 100.0 * t2 * t2;
      const double a = 200.0 * t2; // d/dt2 of 100 t2^2
      g[i] = (g[i] - 2.0 * t1) - (2.0 * x[i]) * a;
      g[i + 1] = g[i + 1] + a;
    }
    return sum;
  }
};

// --- Rastrigin: examples/rastrigin.hpp:6-38 ---------------------------------
template <std::size_t D>
struct Rastrigin {
  static constexpr std::size_t arity = D;
  static constexpr int id = kRastrigin;
  // run by zeus::Zeus on the library's built-in kernels (no parameters)
  static constexpr int zeus_objective = kRastrigin;
  const double* zeus_params() const { return nullptr; }
  long long zeus_n_params() const { return 0; }

  template <class T, class X>
  ZEUS_HD T eval(const X& x) const
  {
    using ::cos;
    using dual::cos;
    T sum = T(10.0 * D);
    for (int i = 0; i < int(D); ++i) {
      const T xi = x[i];
      sum = sum + (xi * xi - T(10.0) * cos(xi * T(2.0 * kPi)));
    }
    return sum;
  }
  template <class T>
  ZEUS_HD T operator()(const std::array<T, D>& x) const
  {
    return eval<T>(ArrayView<T, D>{x});
  }
  ZEUS_HD double value_and_gradient(const double* x, double* g) const
  {
    const double k = 10.0 * (2.0 * kPi);
    double sum = 10.0 * D;
    for (int i = 0; i < int(D); ++i) {
      const double w = 2.0 * kPi * x[i];
      sum += x[i] * x[i] - 10.0 * ::cos(w);
      g[i] = 2.0 * x[i] + k * ::sin(w);
    }
    return sum;
  }
};

// --- Ackley: examples/ackley.hpp:10-50 ----------------------------------------
template <std::size_t D>
struct Ackley {
  static constexpr std::size_t arity = D;
  static constexpr int id = kAckley;
  // run by zeus::Zeus on the library's built-in kernels (no parameters)
  static constexpr int zeus_objective = kAckley;
  const double* zeus_params() const { return nullptr; }
  long long zeus_n_params() const { return 0; }

  template <class T, class X>
  ZEUS_HD T eval(const X& x) const
  {
    if constexpr (kIsDouble<T>) { // ackley.hpp:25-38
      double sum_sq = 0.0, sum_cos = 0.0;
      for (int i = 0; i < int(D); ++i) {
        const double xi = x[i];
        sum_sq += xi * xi;
        sum_cos += ::cos(2.0 * kPi * xi);
      }
      const double term1 = -20.0 * ::exp(-0.2 * ::sqrt(sum_sq / int(D)));
      const double term2 = -::exp(sum_cos / int(D));
      return term1 + term2 + 20.0 + ::exp(1.0);
    } else { // ackley.hpp:10-23
      DualNumber sum_sq = 0.0, sum_cos = 0.0;
      for (int i = 0; i < int(D); ++i) {
        const DualNumber xi = x[i];
        sum_sq += dual::pow(xi, 2.0);
        sum_cos += dual::cos(2.0 * kPi * xi);
      }
      const DualNumber Dd = DualNumber(static_cast<double>(int(D)));
      const DualNumber term1 =
        DualNumber(-20.0) * dual::exp(-0.2 * dual::sqrt(sum_sq / Dd));
      const DualNumber term2 = DualNumber(0.0) - dual::exp(sum_cos / Dd);
      return term1 + term2 + DualNumber(20.0) + dual::exp(DualNumber(1.0));
    }
  }
  template <class T>
  ZEUS_HD T operator()(const std::array<T, D>& x) const
  {
    return eval<T>(ArrayView<T, D>{x});
  }
  // df/dx_i = 4 e1 x_i / (D r) + (2 pi / D) e2 sin(2 pi x_i),
  // r = sqrt(sum x^2 / D), e1 = exp(-0.2 r), e2 = exp(sum cos(2 pi x) / D).
  ZEUS_HD double value_and_gradient(const double* x, double* g) const
  {
    double sum_sq = 0.0, sum_cos = 0.0;
    for (int i = 0; i < int(D); ++i) {
      sum_sq += x[i] * x[i];
      sum_cos += ::cos(2.0 * kPi * x[i]);
    }
    const double r = ::sqrt(sum_sq / int(D));
    const double e1 = ::exp(-0.2 * r);
    const double e2 = ::exp(sum_cos / int(D));
    const double f = -20.0 * e1 + -e2 + 20.0 + ::exp(1.0);
    const double q = (4.0 * e1) / (double(int(D)) * r);
    const double c = (2.0 * kPi / int(D)) * e2;
    for (int i = 0; i < int(D); ++i)
      g[i] = q * x[i] + c * ::sin(2.0 * kPi * x[i]);
    return f;
  }
};

// --- Gaussian: examples/gaussian.hpp:19-33 (C held as a device pointer) ------
template <std::size_t D>
struct Gaussian {
  static constexpr std::size_t arity = D;
  static constexpr int id = kGaussian;
  const double* C = nullptr;

  template <class T, class X>
  ZEUS_HD T eval(const X& x) const
  {
    T q = T(0.0);
    for (int i = 0; i < int(D); ++i) {
      const T xi = x[i];
      for (int j = 0; j < int(D); ++j)
        q += xi * T(C[i * int(D) + j]) * x[j];
    }
    return T(0.5) * q;
  }
  template <class T>
  ZEUS_HD T operator()(const std::array<T, D>& x) const
  {
    return eval<T>(ArrayView<T, D>{x});
  }
  // df/dx_k = 0.5 (sum_j C_kj x_j + sum_i C_ik x_i)
  ZEUS_HD double value_and_gradient(const double* x, double* g) const
  {
    const double f = eval<double>(PtrView{x});
    for (int k = 0; k < int(D); ++k) {
      double r = 0.0, c = 0.0;
      for (int j = 0; j < int(D); ++j) {
        r = ::fma(C[k * int(D) + j], x[j], r);
        c = ::fma(C[j * int(D) + k], x[j], c);
      }
      g[k] = 0.5 * (r + c);
    }
    return f;
  }
};

// --- Poisson negative log-likelihood: simulation/fitting/LogLikelihood.h -----
// mu_i = a (1 - x_i)^b / x_i^(c + d ln x_i);  f = -sum_i [k_i ln mu_i - mu_i
// - lgamma(k_i + 1)].  The lgamma(k+1) constants come precomputed in the
// parameter block (they do not depend on the parameters being fitted).
struct LogLikelihood {
  static constexpr std::size_t arity = 4;
  static constexpr int id = kLogLikelihood;
  const double* P = nullptr; // [x (n) | k (n) | lgamma(k+1) (n)]
  int n = 0;

  template <class T, class X>
  ZEUS_HD T eval(const X& th) const
  {
    using ::log;
    using ::pow;
    using dual::log;
    using dual::pow;
    const T a = th[0], b = th[1], c = th[2], d = th[3];
    T nll = T(0.0);
    for (int i = 0; i < n; ++i) {
      const T x = T(P[i]);
      const T k = T(P[n + i]);
      const T mean = a * pow(T(1.0) - x, b) / pow(x, c + d * log(x));
      const T log_p = k * log(mean) - mean - T(P[2 * n + i]);
      nll -= log_p;
    }
    return nll;
  }
  template <class T>
  ZEUS_HD T operator()(const std::array<T, 4>& x) const
  {
    return eval<T>(ArrayView<T, 4>{x});
  }
  ZEUS_HD double value_and_gradient(const double* th, double* g) const
  {
    const double a = th[0], b = th[1], c = th[2], d = th[3];
    double nll = 0.0, ga = 0.0, gb = 0.0, gc = 0.0, gd = 0.0;
    for (int i = 0; i < n; ++i) {
      const double x = P[i];
      const double k = P[n + i];
      const double lx = ::log(x);
      const double pw = ::pow(1.0 - x, b);
      const double qw = ::pow(x, c + d * lx);
      const double mean = a * pw / qw;
      const double log_p = k * ::log(mean) - mean - P[2 * n + i];
      nll -= log_p;
      const double w = 1.0 - k / mean; // d nll / d mean
      const double wm = w * mean;
      ga += w * (pw / qw);
      gb += wm * ::log(1.0 - x);
      gc -= wm * lx;
      gd -= wm * (lx * lx);
    }
    g[0] = ga;
    g[1] = gb;
    g[2] = gc;
    g[3] = gd;
    return nll;
  }
  // One data bin's contribution (used by the cooperative evaluator).
  ZEUS_HD double bin_value(const double* th, int i) const
  {
    const double x = P[i];
    const double k = P[n + i];
    const double mean = th[0] * ::pow(1.0 - x, th[1]) / ::pow(x, th[2] + th[3] * ::log(x));
    return -(k * ::log(mean) - mean - P[2 * n + i]);
  }
};

// --- 1-16-16-1 tanh MLP, mean squared error ---------------------------------
// Packing per neuron, weights then bias, layer by layer (the order of
// examples/nn.hpp:52-68): [w1_j, b1_j]_{j<16}, [w2_kj (j<16), b2_k]_{k<16},
// [w3_k]_{k<16}, b3.  Hidden layers tanh, output linear.
struct Mlp {
  static constexpr int H = 16;
  static constexpr int W2 = 2 * H;                 // start of layer 2
  static constexpr int W3 = W2 + H * (H + 1);      // start of output layer
  static constexpr std::size_t arity = W3 + H + 1; // 321
  static constexpr int id = kMlp;
  const double* P = nullptr; // [x (m) | y (m)]
  int m = 0;

  template <class T, class X>
  ZEUS_HD T point_output(const X& th, double xp) const
  {
    using ::tanh;
    using dual::tanh;
    int idx = 0;
    T h1[H], h2[H];
    for (int j = 0; j < H; ++j) {
      T sum = T(0.0);
      sum += th[idx] * T(xp);
      ++idx;
      sum += th[idx];
      ++idx;
      h1[j] = tanh(sum);
    }
    for (int k = 0; k < H; ++k) {
      T sum = T(0.0);
      for (int j = 0; j < H; ++j) {
        sum += h1[j] * th[idx];
        ++idx;
      }
      sum += th[idx];
      ++idx;
      h2[k] = tanh(sum);
    }
    T sum = T(0.0);
    for (int k = 0; k < H; ++k) {
      sum += h2[k] * th[idx];
      ++idx;
    }
    sum += th[idx];
    return sum;
  }
  template <class T, class X>
  ZEUS_HD T eval(const X& th) const
  {
    T loss = T(0.0);
    for (int p = 0; p < m; ++p) {
      const T dd = point_output<T>(th, P[p]) - T(P[m + p]);
      loss += dd * dd;
    }
    return loss / T(double(m));
  }
  template <class T>
  ZEUS_HD T operator()(const std::array<T, arity>& x) const
  {
    return eval<T>(ArrayView<T, arity>{x});
  }
  // Reverse sweep per point; activations are recomputed per point (no tape
  // across points).  g must hold 321 entries.
  ZEUS_HD double value_and_gradient(const double* th, double* g) const
  {
    for (int i = 0; i < int(arity); ++i)
      g[i] = 0.0;
    double loss = 0.0;
    const double scale = 2.0 / double(m);
    for (int p = 0; p < m; ++p)
      loss = point_adjoint(th, g, p, loss, scale);
    return loss / double(m);
  }
  // Adds point p's gradient (scaled) into g; returns loss + this point's
  // squared residual.  Shared by the lane-level and cooperative evaluators.
  ZEUS_HD double point_adjoint(const double* th, double* g, int p, double loss,
                               double scale) const
  {
    const double xp = P[p];
    double h1[H], h2[H];
    int idx = 0;
    for (int j = 0; j < H; ++j) {
      double sum = 0.0;
      sum += th[idx++] * xp;
      sum += th[idx++];
      h1[j] = ::tanh(sum);
    }
    for (int k = 0; k < H; ++k) {
      double sum = 0.0;
      for (int j = 0; j < H; ++j)
        sum += h1[j] * th[idx++];
      sum += th[idx++];
      h2[k] = ::tanh(sum);
    }
    double sum = 0.0;
    for (int k = 0; k < H; ++k)
      sum += h2[k] * th[idx++];
    sum += th[idx++];
    const double dd = sum - P[m + p];
    loss += dd * dd;
    const double dy = scale * dd;
    double d2[H];
    for (int k = 0; k < H; ++k) {
      g[W3 + k] += dy * h2[k];
      d2[k] = (dy * th[W3 + k]) * (1.0 - h2[k] * h2[k]);
    }
    g[W3 + H] += dy;
    double d1[H];
    for (int j = 0; j < H; ++j)
      d1[j] = 0.0;
    for (int k = 0; k < H; ++k) {
      const int base = W2 + k * (H + 1);
      for (int j = 0; j < H; ++j) {
        g[base + j] += d2[k] * h1[j];
        d1[j] += d2[k] * th[base + j];
      }
      g[base + H] += d2[k];
    }
    for (int j = 0; j < H; ++j) {
      const double dj = d1[j] * (1.0 - h1[j] * h1[j]);
      g[2 * j] += dj * xp;
      g[2 * j + 1] += dj;
    }
    return loss;
  }
};

// --- sphere: sum x_i^2 (reference test objectives Quad / GoodObjective) -------
template <std::size_t D>
struct Sphere {
  static constexpr std::size_t arity = D;
  static constexpr int id = kSphere;
  // run by zeus::Zeus on the library's built-in kernels (no parameters)
  static constexpr int zeus_objective = kSphere;
  const double* zeus_params() const { return nullptr; }
  long long zeus_n_params() const { return 0; }
  template <class T, class X>
  ZEUS_HD T eval(const X& x) const
  {
    T s = T(0.0);
    for (int i = 0; i < int(D); ++i) {
      const T xi = x[i];
      s += xi * xi;
    }
    return s;
  }
  template <class T>
  ZEUS_HD T operator()(const std::array<T, D>& x) const
  {
    return eval<T>(ArrayView<T, D>{x});
  }
  ZEUS_HD double value_and_gradient(const double* x, double* g) const
  {
    double s = 0.0;
    for (int i = 0; i < int(D); ++i) {
      s += x[i] * x[i];
      g[i] = 2.0 * x[i];
    }
    return s;
  }
};

} // namespace zeus::obj
