// cs_product.hpp -- G = Cs Xt for a tile of 16 points on FP64 MFMA, shared
// by the Gaussian PSO and the tile form of the BFGS.
//
// Cs = (C + C^T) / 2 in MFMA A-operand order (fragment (mb, kb) = 64
// consecutive doubles, lane l holding Cs[16 mb + l % 16][4 kb + l / 16], zero
// outside D x D; prepared once per call, L2-resident); Xt holds 16 points
// as columns in LDS, row-major with a padded stride of 17 doubles (rows D ..
// KP-1 zero).  v_mfma_f64_16x16x4f64 operand layout: A[i][k] on lane
// i + 16k, B[k][j] on lane j + 16k, result D[i][j] in register i % 4 of lane
// j + 16 (i / 4).  Both products below cut their MFMA steps into contiguous
// segments, one per wavefront, whose fragments (one coalesced 512-byte read
// each) are loaded before the segment's MFMA chain starts.
#pragma once

#include <utility>

#include "block.hpp"
#include "config.hpp"

namespace zeus::dev {

typedef double mfma_acc_t __attribute__((ext_vector_type(4)));

constexpr int kCsTile = 16; // points per tile = MFMA N dimension

template <int D>
struct CsShape {
  static constexpr int MB = (D + 15) / 16;   // 16-row blocks of Cs
  static constexpr int KB = (D + 3) / 4;     // 4-deep k steps
  static constexpr int KP = KB * 4;          // padded depth (rows of Xt)
  static constexpr int XS = kCsTile + 1;     // padded LDS row stride of Xt / G
};

// A device-memory pointer as a global-address-space pointer: loads through
// it are global loads, counted by vmcnt alone.  (Through a generic pointer
// they are flat loads, which also count in lgkmcnt, so every wait for an
// LDS read would wait for the fragment loads in flight as well.)
typedef const double __attribute__((address_space(1)))* gdouble_ptr;
ZEUS_D ZEUS_FI gdouble_ptr
as_global(const double* p)
{
  return reinterpret_cast<gdouble_ptr>(reinterpret_cast<uint64_t>(p));
}

// ---- values only: the upper block triangle ----------------------------------
//
// Where only f = 0.5 x^T Cs x is needed (the PSO), symmetry halves the work.
// With Cs cut into 16 x 16 blocks, f = sum_b x_b . (0.5 Cs_bb x_b +
// sum_{c > b} Cs_bc x_c).  The upper fragments hold, for each 16-row block
// mb, the k steps 4 mb .. KB - 1 (the diagonal block and every block right of
// it, in the fragment layout above), the diagonal block multiplied by 0.5
// (exact); they follow the full fragments in the preparation buffer.  Row
// block mb contributes x_mb . acc_mb with no further factor.  At D = 150 a
// tile takes 200 MFMA steps instead of 380.
//
// The steps m = 0 .. T-1 (row block by row block) are cut into NA contiguous
// segments of near-equal length, segment w run by wavefront w, which loads
// all of its fragments (one contiguous range) before its first MFMA.  A
// segment covers parts of up to NRB row blocks; each part's partial value
// goes to its own slot w NRB + r, unused slots are written as zero, and f is
// the sum of the NA NRB slots in slot order.  Every quantity of a segment is
// a compile-time constant of (D, NA, w): the code is specialised per
// wavefront.
namespace upper_detail {
template <int D>
constexpr int len(int mb)
{
  return CsShape<D>::KB - 4 * mb;
}
template <int D>
constexpr int off(int mb)
{
  int o = 0;
  for (int b = 0; b < mb; ++b)
    o += len<D>(b);
  return o;
}
template <int D>
constexpr int mb_of(int m)
{
  int b = 0;
  while (b + 1 < CsShape<D>::MB && off<D>(b + 1) <= m)
    ++b;
  return b;
}
template <int D, int NA>
constexpr int lo(int w)
{
  return int(int64_t(off<D>(CsShape<D>::MB)) * w / NA);
}
template <int D, int NA>
constexpr int row_blocks()
{
  int r = 0;
  for (int w = 0; w < NA; ++w) {
    const int q = mb_of<D>(lo<D, NA>(w + 1) - 1) - mb_of<D>(lo<D, NA>(w)) + 1;
    r = q > r ? q : r;
  }
  return r;
}
} // namespace upper_detail

template <int D>
struct CsUpper {
  using S = CsShape<D>;
  static_assert(4 * (S::MB - 1) < S::KB, "every row block has a step");
  static constexpr int T = upper_detail::off<D>(S::MB); // MFMA steps per tile
  // offset of the upper fragments behind the full ones, in doubles
  static constexpr int64_t kFragOffset = int64_t(S::MB) * S::KB * 64;
  static constexpr int64_t kDoubles = int64_t(T) * 64;
  // segments: at least four steps each, at most one per wavefront
  template <int NW>
  static constexpr int na() { return T / 4 >= NW ? NW : (T / 4 > 0 ? T / 4 : 1); }
};

// A wavefront's fragments, loaded early and consumed later.
template <int D, int NA>
struct CsUpperSeg {
  using S = CsShape<D>;
  using U = CsUpper<D>;
  static constexpr int NRB = upper_detail::row_blocks<D, NA>();
  static constexpr int NF = NA * NRB; // value slots
  static constexpr int LMAX = (U::T + NA - 1) / NA;
  double av[LMAX];

  // Fragments of segment w (wave-uniform): one code path for every
  // wavefront, the segment's start a scalar.  Global loads (the buffer is
  // device memory).
  ZEUS_D ZEUS_FI void load(int w, const double* Cu, int li, int lk)
  {
    w = __builtin_amdgcn_readfirstlane(w);
    if (w >= NA)
      return;
    const int lo = int(int64_t(U::T) * w / NA), len = int(int64_t(U::T) * (w + 1) / NA) - lo;
    const gdouble_ptr frag = as_global(Cu) + (int64_t(lo) << 6) + (li + 16 * lk);
    // unconditional loads (a branch per load would split them into basic
    // blocks): past the segment's end the last fragment is loaded again, and
    // never used
#pragma unroll
    for (int i = 0; i < LMAX; ++i)
      av[i] = frag[(i < len ? i : len - 1) << 6];
  }

  template <int W>
  ZEUS_D ZEUS_FI void mma_w(const double* xt, double* fslot, int li, int lk) const
  {
    constexpr int LO = upper_detail::lo<D, NA>(W), LEN = upper_detail::lo<D, NA>(W + 1) - LO;
    constexpr int MB0 = upper_detail::mb_of<D>(LO);
    constexpr int NR = upper_detail::mb_of<D>(LO + LEN - 1) - MB0 + 1;
    mfma_acc_t acc = {0.0, 0.0, 0.0, 0.0};
    [&]<int... I>(std::integer_sequence<int, I...>) {
      (
        [&] {
          constexpr int m = LO + I, mb = upper_detail::mb_of<D>(m);
          constexpr int kb = 4 * mb + (m - upper_detail::off<D>(mb));
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[I], xt[(4 * kb + lk) * S::XS + li], acc,
                                                     0, 0, 0);
          if constexpr (I == LEN - 1 || m + 1 == upper_detail::off<D>(mb + 1)) {
            // the block's x rows read together (clamped into Xt; rows past
            // D are not used), then the partial value in row order, summed
            // across rows with the lane swaps (as in CsFullSeg)
            constexpr bool full = 16 * mb + 16 <= D;
            double xr[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int rr = 16 * mb + 4 * q + lk;
              xr[q] = xt[(full || rr < D ? rr : D - 1) * S::XS + li];
            }
            double fp = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int rr = 16 * mb + 4 * q + lk;
              if (full || rr < D)
                fp = ::fma(xr[q], acc[q], fp);
            }
            fp = xor_sum<16>(fp);
            fp = xor_sum<32>(fp);
            if (lk == 0)
              fslot[(W * NRB + (mb - MB0)) * kCsTile + li] = fp;
            acc = mfma_acc_t{0.0, 0.0, 0.0, 0.0};
          }
        }(),
        ...);
    }(std::make_integer_sequence<int, LEN>{});
    if (lk == 0)
#pragma unroll
      for (int r = NR; r < NRB; ++r)
        fslot[(W * NRB + r) * kCsTile + li] = 0.0;
  }

  // wavefront w (wave-uniform); wavefronts w >= NA do nothing.  Plain
  // force-inlined recursion, so that av never leaves registers.
  template <int W = 0>
  ZEUS_D ZEUS_FI void mma(int w, const double* xt, double* fslot, int li, int lk) const
  {
    if constexpr (W < NA) {
      if (w == W)
        mma_w<W>(xt, fslot, li, lk);
      else
        mma<W + 1>(w, xt, fslot, li, lk);
    }
  }
};

// ---- the full product in balanced segments ----------------------------------
//
// The MB x KB steps (row block by row block, k ascending: the fragment order
// of the buffer) cut into NA contiguous segments of near-equal length,
// segment w run by wavefront w.  A segment covers parts of one or two row
// blocks; each part leaves its 16 x 16 partial block of G in an LDS slot
// (16 rows, stride XS) and its partial value x . G in a value slot.  Row
// d of G is the sum of the slots of the segments covering its row block, in
// segment order (cover(): at most kCover of them); f is the sum of the value
// slots in slot order.  At D = 150: 380 steps, 23-24 per wavefront, 24 parts.
// A segment's first kKeep fragments can be held in registers across rounds
// (Cs does not change), so a round's chain starts on them while the rest
// arrive from L2.
namespace full_detail {
template <int D>
constexpr int steps()
{
  return CsShape<D>::MB * CsShape<D>::KB;
}
template <int D, int NA>
constexpr int lo(int w)
{
  return int(int64_t(steps<D>()) * w / NA);
}
template <int D>
constexpr int mb_of(int m)
{
  return m / CsShape<D>::KB;
}
template <int D, int NA>
constexpr int parts_in(int w)
{
  return lo<D, NA>(w + 1) > lo<D, NA>(w)
           ? mb_of<D>(lo<D, NA>(w + 1) - 1) - mb_of<D>(lo<D, NA>(w)) + 1
           : 0;
}
template <int D, int NA>
constexpr int part_base(int w)
{
  int p = 0;
  for (int v = 0; v < w; ++v)
    p += parts_in<D, NA>(v);
  return p;
}
// segments covering row block mb: first and last (inclusive)
template <int D, int NA>
constexpr int seg_first(int mb)
{
  int w = 0;
  while (lo<D, NA>(w + 1) <= mb * CsShape<D>::KB)
    ++w;
  return w;
}
template <int D, int NA>
constexpr int seg_last(int mb)
{
  int w = 0;
  while (lo<D, NA>(w + 1) <= (mb + 1) * CsShape<D>::KB - 1)
    ++w;
  return w;
}
template <int D, int NA>
constexpr int max_cover()
{
  int c = 0;
  for (int mb = 0; mb < CsShape<D>::MB; ++mb) {
    const int n = seg_last<D, NA>(mb) - seg_first<D, NA>(mb) + 1;
    c = n > c ? n : c;
  }
  return c;
}
} // namespace full_detail

template <int D, int NA, int KEEP = 0>
struct CsFullSeg {
  using S = CsShape<D>;
  static constexpr int T = full_detail::steps<D>();
  static constexpr int LMAX = (T + NA - 1) / NA;
  static constexpr int kKeep = KEEP < LMAX ? KEEP : LMAX;
  static constexpr int NP = full_detail::part_base<D, NA>(NA); // partial blocks
  static constexpr int kSlot = 16 * S::XS;                     // doubles per block
  static constexpr int kCover = full_detail::max_cover<D, NA>();
  double av[LMAX];

  // Part index of cover entry c of row block mb (-1: none).  Compile-time
  // only: the kernels use the packed first part and count below.
  static constexpr int cover(int mb, int c)
  {
    const int w = full_detail::seg_first<D, NA>(mb) + c;
    if (w > full_detail::seg_last<D, NA>(mb))
      return -1;
    return full_detail::part_base<D, NA>(w) + mb -
           full_detail::mb_of<D>(full_detail::lo<D, NA>(w));
  }
  static constexpr int n_parts(int mb)
  {
    return full_detail::seg_last<D, NA>(mb) - full_detail::seg_first<D, NA>(mb) + 1;
  }
  // The parts of a row block are consecutive (a segment after the first
  // that covers the block starts inside it), so a block's parts are its
  // first part and a count: 6 and 4 bits per row block, packed.
  static constexpr bool parts_consecutive()
  {
    for (int mb = 0; mb < S::MB; ++mb)
      for (int c = 0; c < n_parts(mb); ++c)
        if (cover(mb, c) != cover(mb, 0) + c)
          return false;
    return true;
  }
  static constexpr uint64_t pack(int bits, bool count)
  {
    uint64_t v = 0;
    for (int mb = 0; mb < S::MB; ++mb)
      v |= uint64_t(count ? n_parts(mb) : cover(mb, 0)) << (bits * mb);
    return v;
  }
  static constexpr uint64_t kFirstBits = pack(6, false);
  static constexpr uint64_t kCountBits = pack(4, true);
  static_assert(parts_consecutive(), "a row block's parts are consecutive");
  static_assert(S::MB * 6 <= 64 && NP <= 64 && kCover < 16, "packed row-block parts");

  // Fragments i0 .. i1 - 1 of segment w (wave-uniform); elements past the
  // segment's length are zero and never used.
  ZEUS_D ZEUS_FI void load(int w, const double* Cf, int li, int lk, int i0 = 0, int i1 = LMAX)
  {
    w = __builtin_amdgcn_readfirstlane(w);
    if (w >= NA)
      return;
    const int lo = int(int64_t(T) * w / NA), len = int(int64_t(T) * (w + 1) / NA) - lo;
    const gdouble_ptr frag = as_global(Cf) + (int64_t(lo) << 6) + (li + 16 * lk);
    // unconditional loads (a branch per load would split them into basic
    // blocks, each opened by a wait): past the segment's end the last
    // fragment is loaded again, and never used
#pragma unroll
    for (int i = 0; i < LMAX; ++i)
      if (i >= i0 && i < i1)
        av[i] = frag[(i < len ? i : len - 1) << 6];
  }

  // `hook` runs once, a third of the way through the segment's MFMA chain
  // (a load issued there completes behind the rest of the chain)
  template <int W, class Hook>
  ZEUS_D ZEUS_FI void mma_w(const double* xt, double* gpart, double* fpart, int li, int lk,
                            const Hook& hook) const
  {
    constexpr int LO = full_detail::lo<D, NA>(W), LEN = full_detail::lo<D, NA>(W + 1) - LO;
    constexpr int MB0 = full_detail::mb_of<D>(LO);
    constexpr int PB = full_detail::part_base<D, NA>(W);
    if constexpr (LEN == 0) // fewer steps than segments: the hook still runs
      hook();
    // B operands (rows 4 kb + lk of Xt) read kAhead steps before their MFMA,
    // in a ring of registers, so that the LDS latency overlaps the chain
    constexpr int kAhead = LEN < 6 ? LEN : 6;
    auto b_at = [&](int m) {
      const int kb = m - full_detail::mb_of<D>(m) * S::KB;
      return xt[(4 * kb + lk) * S::XS + li];
    };
    double bq[kAhead > 0 ? kAhead : 1];
#pragma unroll
    for (int i = 0; i < kAhead; ++i)
      bq[i] = b_at(LO + i);
    mfma_acc_t acc = {0.0, 0.0, 0.0, 0.0};
    [&]<int... I>(std::integer_sequence<int, I...>) {
      (
        [&] {
          constexpr int m = LO + I, mb = full_detail::mb_of<D>(m), kb = m - mb * S::KB;
          if constexpr (I == LEN / 3) {
            __builtin_amdgcn_sched_barrier(0);
            hook();
            __builtin_amdgcn_sched_barrier(0);
          }
          const double b = bq[I % kAhead];
          if constexpr (I + kAhead < LEN)
            bq[I % kAhead] = b_at(m + kAhead);
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[I], b, acc, 0, 0, 0);
          // the read stays kAhead steps early (the scheduler would sink it)
          __builtin_amdgcn_sched_barrier(0);
          if constexpr (I == LEN - 1 || kb == S::KB - 1) {
            constexpr int part = PB + mb - MB0;
            constexpr bool full = 16 * mb + 16 <= D; // no row of the block past D
            double* gp = gpart + part * kSlot;
            // the block's x rows read together (clamped into Xt; rows past D
            // are not used), then the partial value in row order
            double xr[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int rr = 16 * mb + 4 * q + lk;
              xr[q] = xt[(full || rr < D ? rr : D - 1) * S::XS + li];
            }
            double fp = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int r = 4 * q + lk, rr = 16 * mb + r;
              gp[r * S::XS + li] = acc[q];
              if (full || rr < D)
                fp = ::fma(xr[q], acc[q], fp);
            }
            fp = xor_sum<16>(fp);
            fp = xor_sum<32>(fp);
            if (lk == 0)
              fpart[part * kCsTile + li] = fp;
            acc = mfma_acc_t{0.0, 0.0, 0.0, 0.0};
          }
        }(),
        ...);
    }(std::make_integer_sequence<int, LEN>{});
  }

  template <int W = 0, class Hook>
  ZEUS_D ZEUS_FI void mma(int w, const double* xt, double* gpart, double* fpart, int li, int lk,
                          const Hook& hook) const
  {
    if constexpr (W < NA) {
      if (w == W)
        mma_w<W>(xt, gpart, fpart, li, lk, hook);
      else
        mma<W + 1>(w, xt, gpart, fpart, li, lk, hook);
    }
  }
};

} // namespace zeus::dev
