// philox.hpp -- counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), identical on host
// and device.  Replaces the per-thread cuRAND states of the reference
// (src/utils.cu:17-50, include/bfgs_common.cuh:15-81): every draw is a pure
// function of (seed, stream, start, draw index), so no state-init kernel is
// needed and a sharded run draws the same numbers as a single-device run.
#pragma once

#include "config.hpp"

namespace zeus::rng {

struct U32x4 {
  uint32_t v[4];
};

ZEUS_HD ZEUS_FI U32x4
philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
              uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u; // Weyl key increments of the paper
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = uint64_t(0xD2511F53u) * c0; // round multipliers
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
    const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
    c1 = uint32_t(p1);
    c3 = uint32_t(p0);
    c0 = n0;
    c2 = n2;
  }
  return U32x4{{c0, c1, c2, c3}};
}

// Draw `draw` of start `start` in stream (seed, stream): block draw/2 of the
// counter (draw/2, start), words (2j, 2j+1) with j = draw & 1, top 53 bits,
// mapped to (m + 1) * 2^-53 in (0, 1] -- the range of the reference's
// curand_uniform_double.
ZEUS_HD ZEUS_FI double
uniform(uint32_t seed, uint32_t stream, uint64_t start, uint64_t draw)
{
  const uint64_t blk = draw >> 1;
  const U32x4 o = philox4x32_10(uint32_t(blk), uint32_t(blk >> 32),
                                uint32_t(start), uint32_t(start >> 32), seed,
                                stream);
  const int j = int(draw & 1u) * 2;
  const uint64_t m = ((uint64_t(o.v[j]) << 32) | o.v[j + 1]) >> 11;
  return double(m + 1u) * 0x1p-53;
}

// Draws 2b and 2b + 1 of a start (both halves of counter block b).
ZEUS_HD ZEUS_FI void
uniform_pair(uint32_t seed, uint32_t stream, uint64_t start, uint64_t blk, double& u0,
             double& u1)
{
  const U32x4 o = philox4x32_10(uint32_t(blk), uint32_t(blk >> 32), uint32_t(start),
                                uint32_t(start >> 32), seed, stream);
  const uint64_t m0 = ((uint64_t(o.v[0]) << 32) | o.v[1]) >> 11;
  const uint64_t m1 = ((uint64_t(o.v[2]) << 32) | o.v[3]) >> 11;
  u0 = double(m0 + 1u) * 0x1p-53;
  u1 = double(m1 + 1u) * 0x1p-53;
}

// reference include/utils.cuh:353-358: lower + (upper + (-lower)) * u
ZEUS_HD ZEUS_FI double
uniform_in(uint32_t seed, uint32_t stream, uint64_t start, uint64_t draw,
           double lo, double hi)
{
  const double u = uniform(seed, stream, start, draw);
  return lo + (hi + (-lo)) * u;
}

// Draw-index layout of one start's stream:
//   PSO init     : 2d (position), 2d+1 (velocity)            d < D
//   PSO iter t>=1: 2Dt + 2d (r1), 2Dt + 2d + 1 (r2)           d < D
//   BFGS random start (PSO_ITER < 0): d                        d < D
ZEUS_HD ZEUS_FI uint64_t
pso_draw(int dim, int iter, int d, int which)
{
  return 2ull * uint64_t(dim) * uint64_t(iter) + 2ull * uint64_t(d) +
         uint64_t(which);
}

} // namespace zeus::rng
