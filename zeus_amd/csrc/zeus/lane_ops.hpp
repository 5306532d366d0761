// lane_ops.hpp -- evaluation of an objective by a single lane.
//
// Works for the built-in objectives (eval<T>(view)) and for any
// reference-style user functor (operator() over std::array<T, D>).
#pragma once

#include "dual.hpp"
#include "objectives.hpp"
#include "traits.hpp"

#include <array>

namespace zeus::dev {

// eval<T>(view) detection (SFINAE, see HasAdjointT for why not a concept).
template <class F, class T, class X, class = void>
struct HasEvalT : std::false_type {};
template <class F, class T, class X>
struct HasEvalT<F, T, X,
                VoidT<decltype(std::declval<const F&>().template eval<T>(
                  std::declval<const X&>()))>> : std::true_type {};
template <class F, class T, class X>
inline constexpr bool HasEval = HasEvalT<F, T, X>::value;

template <class F, int D>
ZEUS_D ZEUS_FI double
lane_value(const F& f, const double* x)
{
  if constexpr (HasEval<F, double, obj::PtrView>) {
    return f.template eval<double>(obj::PtrView{x});
  } else {
    std::array<double, D> a;
#pragma unroll
    for (int i = 0; i < D; ++i)
      a[i] = x[i];
    return double(f(a));
  }
}

// One forward-mode pass seeded at dimension i: returns df/dx_i.
template <class F, int D>
ZEUS_D ZEUS_FI double
lane_dual_pass(const F& f, const double* x, int i)
{
  if constexpr (HasEval<F, dual::DualNumber, obj::SeededView>) {
    return f.template eval<dual::DualNumber>(obj::SeededView{x, i}).dual;
  } else {
    std::array<dual::DualNumber, D> a;
    for (int j = 0; j < D; ++j)
      a[j] = dual::DualNumber(x[j], j == i ? 1.0 : 0.0);
    return dual::DualNumber(f(a)).dual;
  }
}

// Full gradient by one lane: the adjoint when requested and available,
// otherwise D dual passes (reference include/duals.cuh:242-266).
template <class F, int D, Gradient G>
ZEUS_D ZEUS_FI void
lane_gradient(const F& f, const double* x, double* g)
{
  if constexpr (G == Gradient::Adjoint && HasAdjoint<F>) {
    (void)f.value_and_gradient(x, g);
  } else {
    for (int i = 0; i < D; ++i)
      g[i] = lane_dual_pass<F, D>(f, x, i);
  }
}

} // namespace zeus::dev
