// dual.hpp -- forward-mode dual numbers, host and device.
//
// Same arithmetic rules as the reference's dual::DualNumber
// (include/duals.cuh:14-239, digamma src/duals.cu:8-39), so that objectives
// written against the reference (templated on T = double or T = DualNumber)
// compile unchanged and produce the same derivatives:
//   (a,a')(b,b') = (ab, a'b + ab')       (a,a')/(b,b') = (a/b, (a'b - ab')/b^2)
//   abs at 0 -> derivative NaN           pow(b, e) with b == 0 -> e b^(e-1) b'
// tanh is added (the reference has none; the MLP objective needs it).
#pragma once

#include "config.hpp"

#include <array>
#include <cmath>
#include <limits>

namespace dual {

struct DualNumber {
  double real;
  double dual;

  ZEUS_HD constexpr DualNumber(double r = 0.0, double d = 0.0) : real(r), dual(d) {}

  ZEUS_HD constexpr DualNumber operator-() const { return {-real, -dual}; }
  ZEUS_HD constexpr DualNumber& operator+=(const DualNumber& o)
  {
    real += o.real;
    dual += o.dual;
    return *this;
  }
  ZEUS_HD constexpr DualNumber& operator-=(const DualNumber& o)
  {
    real -= o.real;
    dual -= o.dual;
    return *this;
  }
  ZEUS_HD constexpr DualNumber operator+(const DualNumber& o) const
  {
    return {real + o.real, dual + o.dual};
  }
  ZEUS_HD constexpr DualNumber operator-(const DualNumber& o) const
  {
    return {real - o.real, dual - o.dual};
  }
  ZEUS_HD constexpr DualNumber operator*(const DualNumber& o) const
  {
    return {real * o.real, dual * o.real + real * o.dual};
  }
  ZEUS_HD constexpr DualNumber operator/(const DualNumber& o) const
  {
    const double den = o.real * o.real;
    return {real / o.real, (dual * o.real - real * o.dual) / den};
  }
  ZEUS_HD constexpr DualNumber& operator*=(const DualNumber& o)
  {
    *this = *this * o;
    return *this;
  }
  ZEUS_HD constexpr DualNumber& operator/=(const DualNumber& o)
  {
    *this = *this / o;
    return *this;
  }
};

ZEUS_HD constexpr inline DualNumber
operator-(double l, const DualNumber& r)
{
  return {l - r.real, -r.dual};
}
ZEUS_HD constexpr inline DualNumber
operator*(double l, const DualNumber& r)
{
  return {l * r.real, l * r.dual};
}
ZEUS_HD constexpr inline bool
operator==(const DualNumber& a, const DualNumber& b)
{
  return a.real == b.real && a.dual == b.dual;
}
ZEUS_HD constexpr inline bool
operator!=(const DualNumber& a, const DualNumber& b)
{
  return !(a == b);
}
// Ordering compares the value part (used by objectives with branches).
ZEUS_HD constexpr inline bool
operator<(const DualNumber& a, const DualNumber& b)
{
  return a.real < b.real;
}
ZEUS_HD constexpr inline bool
operator>(const DualNumber& a, const DualNumber& b)
{
  return a.real > b.real;
}

ZEUS_HD constexpr inline double
pi()
{
  return 3.14159265358979323846;
}

// Digamma by reflection (x < 1/2), upward recurrence to x >= 8 and the
// asymptotic series through the B12 term; +inf at the non-positive integers.
ZEUS_HD inline double
digamma(double x)
{
  if (x <= 0.0 && x == ::floor(x))
    return std::numeric_limits<double>::infinity();
  double shift = 0.0;
  if (x < 0.5) {
    // psi(x) = psi(1 - x) - pi cot(pi x)
    shift = -pi() / ::tan(pi() * x);
    x = 1.0 - x;
  }
  double acc = 0.0;
  while (x < 8.0) {
    acc -= 1.0 / x;
    x += 1.0;
  }
  const double r = 1.0 / x;
  const double r2 = r * r;
  // -sum_k B_2k / (2k x^2k), k = 1..6
  const double tail =
    r2 * (-1.0 / 12.0 +
          r2 * (1.0 / 120.0 +
                r2 * (-1.0 / 252.0 +
                      r2 * (1.0 / 240.0 + r2 * (-1.0 / 132.0 + r2 * (691.0 / 32760.0))))));
  return (acc + ::log(x) - 0.5 * r + tail) + shift;
}

ZEUS_HD inline DualNumber
abs(const DualNumber& a)
{
  if (a.real == 0.0)
    return {0.0, std::numeric_limits<double>::quiet_NaN()};
  return a.real < 0.0 ? DualNumber{-a.real, -a.dual} : a;
}
ZEUS_HD inline DualNumber
sin(const DualNumber& a)
{
  return {::sin(a.real), a.dual * ::cos(a.real)};
}
ZEUS_HD inline DualNumber
cos(const DualNumber& a)
{
  return {::cos(a.real), -a.dual * ::sin(a.real)};
}
ZEUS_HD inline DualNumber
exp(const DualNumber& a)
{
  const double e = ::exp(a.real);
  return {e, a.dual * e};
}
ZEUS_HD inline DualNumber
sqrt(const DualNumber& a)
{
  const double r = ::sqrt(a.real);
  return {r, a.dual / (2.0 * r)};
}
ZEUS_HD inline DualNumber
log(const DualNumber& a)
{
  return {::log(a.real), a.dual / a.real};
}
ZEUS_HD inline DualNumber
atan2(const DualNumber& y, const DualNumber& x)
{
  const double den = x.real * x.real + y.real * y.real;
  return {::atan2(y.real, x.real), (x.real * y.dual - y.real * x.dual) / den};
}
ZEUS_HD inline DualNumber
pow(const DualNumber& b, const DualNumber& e)
{
  const double v = ::pow(b.real, e.real);
  return {v, v * (e.dual * ::log(b.real) + e.real * b.dual / b.real)};
}
ZEUS_HD inline DualNumber
pow(const DualNumber& b, double e)
{
  const double v = ::pow(b.real, e);
  if (b.real == 0.0)
    return {v, e * ::pow(b.real, e - 1.0) * b.dual};
  return {v, e * v / b.real * b.dual};
}
ZEUS_HD inline DualNumber
lgamma(const DualNumber& a)
{
  return {::lgamma(a.real), a.dual * digamma(a.real)};
}
ZEUS_HD inline DualNumber
tanh(const DualNumber& a)
{
  const double t = ::tanh(a.real);
  return {t, a.dual * (1.0 - t * t)};
}

// Gradient by D forward passes, pass i seeded with e_i
// (reference include/duals.cuh:242-266).
template <class F, std::size_t D>
ZEUS_HD inline void
calculateGradientUsingAD(const F& f, const std::array<double, D>& x,
                         std::array<double, D>& g)
{
  std::array<DualNumber, D> xd;
  for (std::size_t i = 0; i < D; ++i)
    xd[i] = DualNumber(x[i], 0.0);
  for (std::size_t i = 0; i < D; ++i) {
    xd[i].dual = 1.0;
    g[i] = DualNumber(f(xd)).dual;
    xd[i].dual = 0.0;
  }
}

} // namespace dual
