// traits.hpp -- the objective contract (reference include/traits.hpp:1-71).
//
// An objective F is callable with std::array<double, D> (returning something
// convertible to double) and with std::array<dual::DualNumber, D> (returning
// something convertible to DualNumber).  D comes from F::arity, from a
// Functor<N> template parameter, or from a user FnTraits specialisation.
//
// New in this implementation: an objective may also provide a hand-written
// reverse-mode adjoint,
//     double value_and_gradient(const double* x, double* g) const;  // device
// which the BFGS uses instead of the D dual passes when the adjoint gradient
// is requested.  Without it the dual passes are used.
#pragma once

#include "dual.hpp"

#include <array>
#include <concepts>
#include <cstddef>
#include <type_traits>
#include <utility>

namespace zeus {

template <typename...>
using VoidT = void;

template <typename F, typename = void>
struct FnTraits {
  static constexpr std::size_t arity = 0;
};

template <typename F>
struct FnTraits<F, VoidT<decltype(F::arity)>> {
  static constexpr std::size_t arity = F::arity;
};

template <template <std::size_t> class Functor, std::size_t N>
struct FnTraits<Functor<N>> {
  static constexpr std::size_t arity = N;
};

template <typename F, typename T, std::size_t D>
concept CallableWithArray = requires(F f, std::array<T, D> a) {
  { f(a) } -> std::convertible_to<T>;
};

template <typename F, std::size_t D>
concept ZeusObjective =
  CallableWithArray<F, double, D> && CallableWithArray<F, dual::DualNumber, D>;

// Detection by SFINAE rather than a requires-expression: clang's HIP mode
// evaluates a concept whose expression calls a __host__ __device__ member as
// unsatisfied when it is first checked inside device code, which would turn
// the adjoint path silently into dual passes.
template <typename F, typename = void>
struct HasAdjointT : std::false_type {};
template <typename F>
struct HasAdjointT<F, VoidT<decltype(std::declval<const F&>().value_and_gradient(
                        std::declval<const double*>(), std::declval<double*>()))>>
  : std::true_type {};
template <typename F>
inline constexpr bool HasAdjoint = HasAdjointT<F>::value;

// PRNG algorithms (reference traits.hpp:67-71): Marsaglia's xorwow with a
// state per start, counter-based Philox4x32-10, and dimension 1 of Sobol's
// sequence read from the start's index on (prng.hpp).  Seeding and
// conversions are this library's own: cuRAND's streams are not reproduced.
enum class PRNGType { XORWOW, PHILOX, SOBOL };

// Gradient method of the BFGS.
enum class Gradient { Adjoint, Dual };

} // namespace zeus
