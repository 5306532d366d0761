// prng.hpp -- the reference's selectable generators (include/traits.hpp:
// 67-71: XORWOW, PHILOX, SOBOL), identical on host and device.
//
// Draws are uniforms in (0, 1], taken in the order of the draw-index
// layout of philox.hpp (PSO init, PSO iterations, BFGS random start).
//   PHILOX: a pure function of (seed, stream, start, draw index).
//   XORWOW: Marsaglia, "Xorshift RNGs", J. Stat. Softw. 8(14), 2003, the
//           xorwow recurrence (five-word xorshift plus a Weyl sequence with
//           increment 362437).  Every (start, coordinate) owns one state,
//           seeded from two Philox blocks that no draw uses, and gives that
//           coordinate's draws in order: PSO init (position, velocity), each
//           PSO iteration (r1, r2), or a BFGS random start (one draw).  The
//           reference keeps one state per start because one thread owns a
//           start there; here one thread owns a coordinate, so the state
//           travels with it and no thread waits on another's stream.  A
//           uniform takes two outputs (top 32 + next 21 bits -> 53-bit m,
//           value (m + 1) 2^-53).
//   SOBOL:  dimension 1 of Sobol's sequence, direction numbers 2^(31-j)
//           (base-2 van der Corput in Gray-code order: x_n =
//           bitreverse32(n ^ (n >> 1))); draw k of start i is point i + k,
//           so neighbouring starts' streams overlap by one draw as the
//           reference's per-thread offsets do; value (x + 1) 2^-32.
// None of the three reproduces cuRAND's streams; seeding and conversions are
// this library's own.
#pragma once

#include "batch.hpp"
#include "philox.hpp"

namespace zeus::rng {

enum Kind : int32_t { kXorwow = 0, kPhilox = 1, kSobol = 2 };

struct Xorwow {
  uint32_t x, y, z, w, v, d;
};

// State of (start, coordinate): counter blocks (2c, 2^32 - 1) and
// (2c + 1, 2^32 - 1) of the start, above every draw block (< 2^63).
ZEUS_HD ZEUS_FI Xorwow
xorwow_seed(uint32_t seed, uint32_t stream, uint64_t start, uint32_t coord)
{
  const U32x4 a = philox4x32_10(2u * coord, 0xFFFFFFFFu, uint32_t(start),
                                uint32_t(start >> 32), seed, stream);
  const U32x4 b = philox4x32_10(2u * coord + 1u, 0xFFFFFFFFu, uint32_t(start),
                                uint32_t(start >> 32), seed, stream);
  Xorwow s{a.v[0], a.v[1], a.v[2], a.v[3], b.v[0], b.v[1]};
  if ((s.x | s.y | s.z | s.w | s.v) == 0u)
    s.v = 0x2545F491u; // the xorshift part must not be all zero
  return s;
}

ZEUS_HD ZEUS_FI uint32_t
xorwow_next(Xorwow& s)
{
  const uint32_t t = s.x ^ (s.x >> 2);
  s.x = s.y;
  s.y = s.z;
  s.z = s.w;
  s.w = s.v;
  s.v = (s.v ^ (s.v << 4)) ^ (t ^ (t << 1));
  s.d += 362437u;
  return s.v + s.d;
}

ZEUS_HD ZEUS_FI double
xorwow_uniform(Xorwow& s)
{
  const uint64_t hi = xorwow_next(s);
  const uint64_t lo = xorwow_next(s);
  const uint64_t m = (hi << 21) | (lo >> 11);
  return double(m + 1u) * 0x1p-53;
}

ZEUS_HD ZEUS_FI uint32_t
bitreverse32(uint32_t v)
{
  v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
  v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
  v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
  v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
  return (v >> 16) | (v << 16);
}

// Draw `draw` of start `start` in SOBOL mode: point start + draw.
ZEUS_HD ZEUS_FI double
sobol_uniform(uint64_t start, uint64_t draw)
{
  const uint32_t n = uint32_t(start + draw);
  const uint32_t x = bitreverse32(n ^ (n >> 1));
  return (double(x) + 1.0) * 0x1p-32;
}

} // namespace zeus::rng

namespace zeus::dev {

// XORWOW states of a class: word k of (start s, coordinate d) at
// st[k * stride + s * D + d] (stride = starts x D; coalesced over d).
ZEUS_D ZEUS_FI rng::Xorwow
xw_load(const uint32_t* st, int64_t stride, int64_t o)
{
  return rng::Xorwow{st[o], st[stride + o], st[2 * stride + o], st[3 * stride + o],
                     st[4 * stride + o], st[5 * stride + o]};
}
ZEUS_D ZEUS_FI void
xw_store(uint32_t* st, int64_t stride, int64_t o, const rng::Xorwow& x)
{
  st[o] = x.x;
  st[stride + o] = x.y;
  st[2 * stride + o] = x.z;
  st[3 * stride + o] = x.w;
  st[4 * stride + o] = x.v;
  st[5 * stride + o] = x.d;
}

// Draw of coordinate d of a BFGS random start (problem-local start sl):
// draw index d, or the first draw of the (start, coordinate) XORWOW stream;
// lower + (upper + (-lower)) u as reference include/utils.cuh:353-358.
ZEUS_HD ZEUS_FI double
start_uniform_in(const ProblemDesc& P, uint64_t sl, int d)
{
  double u;
  if (P.prng == rng::kXorwow) {
    rng::Xorwow st = rng::xorwow_seed(P.seed, P.stream, sl, uint32_t(d));
    u = rng::xorwow_uniform(st);
  } else if (P.prng == rng::kSobol) {
    u = rng::sobol_uniform(sl, uint64_t(d));
  } else {
    u = rng::uniform(P.seed, P.stream, sl, uint64_t(d));
  }
  return P.lower + (P.upper + (-P.lower)) * u;
}

} // namespace zeus::dev
