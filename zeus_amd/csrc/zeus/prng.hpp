// prng.hpp -- the reference's selectable generators (include/traits.hpp:
// 67-71: XORWOW, PHILOX, SOBOL), identical on host and device.
//
// Every start owns one stream of uniforms in (0, 1], consumed in the order
// of the draw-index layout of philox.hpp (PSO init, PSO iterations, BFGS
// random start).  PHILOX draws are a pure function of the index.  The other
// two follow the reference's state-per-start model (bfgs_common.cuh:15-81)
// from published definitions, not cuRAND's seeding or conversions:
//   XORWOW: Marsaglia, "Xorshift RNGs", J. Stat. Softw. 8(14), 2003, the
//           xorwow recurrence (five-word xorshift plus a Weyl sequence with
//           increment 362437).  The state of start i is seeded from Philox
//           blocks that no draw uses; a uniform takes two outputs
//           (top 32 + next 21 bits -> 53-bit m, value (m + 1) 2^-53).
//   SOBOL:  dimension 1 of Sobol's sequence, direction numbers 2^(31-j)
//           (base-2 van der Corput in Gray-code order: x_n =
//           bitreverse32(n ^ (n >> 1))); start i reads the sequence from
//           index i on, so neighbouring starts' streams overlap by one draw
//           as the reference's per-thread offsets do; value (x + 1) 2^-32.
// Stateful streams reach the kernels through a draw buffer filled once per
// PSO iteration (and once for BFGS random starts), so the PSO kernels read
// one layout whatever the generator.
#pragma once

#include "batch.hpp"
#include "philox.hpp"

namespace zeus::rng {

enum Kind : int32_t { kXorwow = 0, kPhilox = 1, kSobol = 2 };

struct Xorwow {
  uint32_t x, y, z, w, v, d;
};

ZEUS_HD ZEUS_FI Xorwow
xorwow_seed(uint32_t seed, uint32_t stream, uint64_t start)
{
  // counter blocks 2^64 - 1 and 2^64 - 2 of the start: beyond every draw
  const U32x4 a = philox4x32_10(0xFFFFFFFFu, 0xFFFFFFFFu, uint32_t(start),
                                uint32_t(start >> 32), seed, stream);
  const U32x4 b = philox4x32_10(0xFFFFFFFEu, 0xFFFFFFFFu, uint32_t(start),
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

// One word-packed XORWOW state per start: 6 uint32 at state[6 s].
ZEUS_D ZEUS_FI rng::Xorwow
load_state(const uint32_t* p)
{
  return rng::Xorwow{p[0], p[1], p[2], p[3], p[4], p[5]};
}
ZEUS_D ZEUS_FI void
store_state(uint32_t* p, const rng::Xorwow& s)
{
  p[0] = s.x;
  p[1] = s.y;
  p[2] = s.z;
  p[3] = s.w;
  p[4] = s.v;
  p[5] = s.d;
}

// Seed the XORWOW states of a class's starts (other generators: nothing).
static __global__ void
rng_init_kernel(const ProblemDesc* probs, int n_probs, int64_t total, uint32_t* state)
{
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= total)
    return;
  const ProblemDesc P = probs[find_problem(probs, n_probs, s)];
  if (P.prng == rng::kXorwow)
    store_state(state + 6 * s, rng::xorwow_seed(P.seed, P.stream, uint64_t(s - P.start0)));
}

// The 2D uniforms of PSO iteration `iter` (0 = init) of every start still
// iterating: draws[k * total + s], k = 2d + j in draw-index order.
static __global__ void
rng_pso_draws_kernel(const ProblemDesc* probs, int n_probs, int64_t total, int iter,
                     uint32_t* state, double* draws)
{
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= total)
    return;
  const ProblemDesc P = probs[find_problem(probs, n_probs, s)];
  if (iter > P.pso_iter)
    return;
  const uint64_t sl = uint64_t(s - P.start0);
  const int n = 2 * P.dim;
  if (P.prng == rng::kXorwow) {
    rng::Xorwow st = load_state(state + 6 * s);
    for (int k = 0; k < n; ++k)
      draws[int64_t(k) * total + s] = rng::xorwow_uniform(st);
    store_state(state + 6 * s, st);
  } else {
    const uint64_t k0 = rng::pso_draw(P.dim, iter, 0, 0);
    for (int k = 0; k < n; ++k)
      draws[int64_t(k) * total + s] = P.prng == rng::kSobol
                                        ? rng::sobol_uniform(sl, k0 + k)
                                        : rng::uniform(P.seed, P.stream, sl, k0 + k);
  }
}

// BFGS random starts of the problems without PSO and without caller start
// points, for generators other than PHILOX: x[s * D + d] = lo + (hi - lo) u.
static __global__ void
rng_starts_kernel(const ProblemDesc* probs, int n_probs, int64_t total, uint32_t* state,
                  double* x)
{
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= total)
    return;
  const ProblemDesc P = probs[find_problem(probs, n_probs, s)];
  if (P.pso_iter >= 0 || P.prng == rng::kPhilox || P.x0 != x + P.start0 * P.dim)
    return;
  const uint64_t sl = uint64_t(s - P.start0);
  rng::Xorwow st{};
  if (P.prng == rng::kXorwow)
    st = load_state(state + 6 * s);
  for (int d = 0; d < P.dim; ++d) {
    const double u = P.prng == rng::kXorwow ? rng::xorwow_uniform(st)
                                            : rng::sobol_uniform(sl, uint64_t(d));
    x[s * P.dim + d] = P.lower + (P.upper + (-P.lower)) * u;
  }
  if (P.prng == rng::kXorwow)
    store_state(state + 6 * s, st);
}

} // namespace zeus::dev
