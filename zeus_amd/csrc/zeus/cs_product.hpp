// cs_product.hpp -- G = Cs Xt for a tile of 16 points on FP64 MFMA, shared
// by the Gaussian PSO and the tile form of the BFGS.
//
// Cs = (C + C^T) / 2 in MFMA A-operand order (fragment (mb, kb) = 64
// consecutive doubles, lane l holding Cs[16 mb + l % 16][4 kb + l / 16], zero
// outside D x D; prepared once per call, L2-resident); Xt holds 16 points
// as columns in LDS, row-major with a padded stride of 17 doubles (rows D ..
// KP-1 zero).  v_mfma_f64_16x16x4f64 operand layout: A[i][k] on lane
// i + 16k, B[k][j] on lane j + 16k, result D[i][j] in register i % 4 of lane
// j + 16 (i / 4).  The product is cut into units (16-row block of Cs, half
// of the k range); a unit loads all its Cs fragments (one coalesced 512-byte
// read each) before its MFMA chain starts, so the L2 latency is paid once
// per unit.
#pragma once

#include "config.hpp"

namespace zeus::dev {

typedef double mfma_acc_t __attribute__((ext_vector_type(4)));

constexpr int kCsTile = 16; // points per tile = MFMA N dimension

template <int D>
struct CsShape {
  static constexpr int MB = (D + 15) / 16;   // 16-row blocks of Cs
  static constexpr int KB = (D + 3) / 4;     // 4-deep k steps
  static constexpr int KP = KB * 4;          // padded depth (rows of Xt)
  static constexpr int KB1 = (KB + 1) / 2;   // k steps of the first half
  static constexpr int U = 2 * MB;           // (row block, k half) units
  static constexpr int XS = kCsTile + 1;     // padded LDS row stride of Xt / G
};

template <int D, int KB0, int KB1>
ZEUS_D ZEUS_FI mfma_acc_t
cs_unit_range(const double* Cf, const double* xt, int mb, int li, int lk)
{
  using S = CsShape<D>;
  const double* frag = Cf + (int64_t(mb) * S::KB << 6) + (li + 16 * lk);
  double av[KB1 - KB0];
#pragma unroll
  for (int kb = KB0; kb < KB1; ++kb)
    av[kb - KB0] = frag[kb << 6];
  mfma_acc_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kb = KB0; kb < KB1; ++kb)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kb - KB0], xt[(4 * kb + lk) * S::XS + li], acc,
                                               0, 0, 0);
  return acc;
}

// The same product split in two steps, so that a caller can issue a unit's
// fragment loads early (before the points are staged) and run the MFMA
// chain later: CsFrag<D>::load, then CsFrag<D>::mma.
template <int D>
struct CsFrag {
  using S = CsShape<D>;
  static constexpr int N = S::KB - S::KB1 > S::KB1 ? S::KB - S::KB1 : S::KB1;
  double av[N];

  ZEUS_D ZEUS_FI void load(const double* Cf, int u, int li, int lk)
  {
    const int kb0 = (u & 1) ? S::KB1 : 0, n = (u & 1) ? S::KB - S::KB1 : S::KB1;
    const double* frag = Cf + (int64_t(u >> 1) * S::KB << 6) + (li + 16 * lk) + (kb0 << 6);
#pragma unroll
    for (int k = 0; k < N; ++k)
      av[k] = k < n ? frag[k << 6] : 0.0;
  }

  ZEUS_D ZEUS_FI mfma_acc_t mma(const double* xt, int u, int li, int lk) const
  {
    const int kb0 = (u & 1) ? S::KB1 : 0, n = (u & 1) ? S::KB - S::KB1 : S::KB1;
    mfma_acc_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < N; ++k)
      if (k < n)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[k], xt[(4 * (kb0 + k) + lk) * S::XS + li],
                                                   acc, 0, 0, 0);
    return acc;
  }
};

// Unit u: rows 16 (u / 2) .. + 15 of Cs, k half u % 2.  Lane (li, lk) of the
// result holds G[16 (u / 2) + 4 q + lk][li] in component q (partial over
// the unit's k half).
template <int D>
ZEUS_D ZEUS_FI mfma_acc_t
cs_unit(const double* Cf, const double* xt, int u, int li, int lk)
{
  using S = CsShape<D>;
  return (u & 1) ? cs_unit_range<D, S::KB1, S::KB>(Cf, xt, u >> 1, li, lk)
                 : cs_unit_range<D, 0, S::KB1>(Cf, xt, u >> 1, li, lk);
}

} // namespace zeus::dev
