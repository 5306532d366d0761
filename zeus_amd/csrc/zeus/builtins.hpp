// builtins.hpp -- objectives that zeus::Zeus runs on the library's built-in
// kernels.
//
// A functor F is a built-in when it declares
//   static constexpr int zeus_objective;      // ZEUS_OBJ_*
//   const double* zeus_params() const;        // device data (or nullptr)
//   long long zeus_n_params() const;
// zeus::Zeus then hands the call to zeus_run_batch (the C ABI), whose
// kernels for that objective are compiled into libzeus_amd.so: the MFMA PSO
// and the tile forms for the Gaussian, the team form for the small separable
// objectives, picked by the library's autotuner.  A dimension the library
// does not instantiate falls back to the functor path (the kernel templates
// instantiated in the caller's translation unit), as do the per-thread form
// (parallel = false, which times its gradients and updates) and trajectory
// recording.  F stays an ordinary reference-style callable either way.
//
// The parameter-free objectives of zeus/objectives.hpp (Rosenbrock,
// Rastrigin, Ackley, Sphere) are built-ins; zeus::Gaussian<N> below is the
// reference's example Gaussian (examples/gaussian.hpp:8-34: a Matrix member
// filled from a host std::array, f = 0.5 x^T C x) as a built-in.
#pragma once

#include "matrix.hpp"
#include "objectives.hpp"
#include "traits.hpp"
#include "zeus_c.h"

#include <array>
#include <cstddef>
#include <stdexcept>
#include <type_traits>
#include <utility>

namespace zeus {

template <typename F, typename = void>
struct IsBuiltinT : std::false_type {};
template <typename F>
struct IsBuiltinT<F, VoidT<decltype(F::zeus_objective), decltype(std::declval<const F&>().zeus_params()),
                           decltype(std::declval<const F&>().zeus_n_params())>>
  : std::true_type {};
template <typename F>
inline constexpr bool IsBuiltin = IsBuiltinT<F>::value;

template <std::size_t N>
class Gaussian {
public:
  static constexpr std::size_t arity = N;
  static constexpr int zeus_objective = ZEUS_OBJ_GAUSSIAN;

  // one bulk copy of the host matrix and one upload (examples/gaussian.hpp:14-18)
  explicit Gaussian(const std::array<std::array<double, N>, N>& C_host) : C_(N, N)
  {
    C_.set(&C_host[0][0], N * N);
  }
  explicit Gaussian(const double* C_host_row_major) : C_(N, N)
  {
    C_.set(C_host_row_major, N * N);
  }

  template <class T>
  ZEUS_HD T operator()(const std::array<T, N>& x) const
  {
    return view()(x);
  }
  // hand-written adjoint (used by the functor path)
  ZEUS_HD double value_and_gradient(const double* x, double* g) const
  {
    return view().value_and_gradient(x, g);
  }

  const double* zeus_params() const { return C_.deviceData(); }
  long long zeus_n_params() const { return static_cast<long long>(N * N); }
  const Matrix<double>& matrix() const { return C_; }

private:
  Matrix<double> C_;
  ZEUS_HD obj::Gaussian<N> view() const { return obj::Gaussian<N>{C_.data()}; }
};

} // namespace zeus
