// block.hpp -- workgroup-level primitives for the per-workgroup BFGS.
#pragma once

#include "config.hpp"

namespace zeus::dev {

// One DPP lane exchange of a double (two 32-bit moves).  Every pattern used
// here reads an active lane of the same row, so no lane reads out of range.
template <int kCtrl>
ZEUS_D ZEUS_FI double
dpp_f64(double v)
{
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(uint32_t(b)), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(uint32_t(b >> 32)), kCtrl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

ZEUS_D ZEUS_FI double
readlane_f64(double v, int lane)
{
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(int(uint32_t(b)), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(int(uint32_t(b >> 32)), lane);
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// One DPP lane exchange of any 64-bit value (two 32-bit moves).
template <int kCtrl, class T>
ZEUS_D ZEUS_FI T
dpp_b64(T v)
{
  static_assert(sizeof(T) == 8, "64-bit values");
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(uint32_t(b)), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(uint32_t(b >> 32)), kCtrl, 0xF, 0xF, false);
  return __builtin_bit_cast(T, (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

// The gfx950 row swaps of a 64-bit value with itself (kXor = 16:
// v_permlane16_swap, odd rows of the first operand with even rows of the
// second; kXor = 32: v_permlane32_swap, upper half with lower half).  Lane
// by lane, {first, second} hold v and v[lane ^ kXor] in some order.  All
// lanes must be active.
template <int kXor, class T>
ZEUS_D ZEUS_FI void
row_swap_b64(T v, T& first, T& second)
{
  static_assert(sizeof(T) == 8, "64-bit values");
  static_assert(kXor == 16 || kXor == 32, "row swaps only");
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const unsigned lo = unsigned(b), hi = unsigned(b >> 32);
  auto swap = [](unsigned x) {
    if constexpr (kXor == 16)
      return __builtin_amdgcn_permlane16_swap(x, x, false, false);
    else
      return __builtin_amdgcn_permlane32_swap(x, x, false, false);
  };
  const auto l = swap(lo);
  const auto h = swap(hi);
  first = __builtin_bit_cast(T, (uint64_t(h[0]) << 32) | l[0]);
  second = __builtin_bit_cast(T, (uint64_t(h[1]) << 32) | l[1]);
}

// v + v[lane ^ kXor] in every lane, kXor = 16 or 32, by the gfx950 lane
// swaps (v_permlane16_swap: odd rows of the first operand with even rows of
// the second; v_permlane32_swap: upper half with lower half): register moves,
// no LDS.  With both operands v, the two results hold, lane by lane, v and
// its partner's v in some order; IEEE addition is commutative, so every lane
// gets the bits of v + v[lane ^ kXor].  All lanes must be active.
template <int kXor>
ZEUS_D ZEUS_FI double
xor_sum(double v)
{
  double a, c;
  row_swap_b64<kXor>(v, a, c);
  return a + c;
}

// Sum of a 64-bit integer over the 64 lanes, every lane gets it (integer
// sums are exact in any order): DPP exchanges inside each row, then the row
// swaps.  All lanes must be active.
ZEUS_D ZEUS_FI int64_t
wave_sum_i64(int64_t v)
{
  v += dpp_b64<0xB1>(v);  // quad_perm [1,0,3,2]
  v += dpp_b64<0x4E>(v);  // quad_perm [2,3,0,1]
  v += dpp_b64<0x141>(v); // row_half_mirror
  v += dpp_b64<0x140>(v); // row_mirror
  int64_t a, c;
  row_swap_b64<16>(v, a, c);
  v = a + c;
  row_swap_b64<32>(v, a, c);
  return a + c;
}

// A value every lane of the wavefront holds identically, re-read from the
// first active lane so that the compiler treats it as wave-uniform (scalar
// registers) instead of one copy per lane.
template <class T>
ZEUS_D ZEUS_FI T
wave_uniform(const T& v)
{
  static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
  struct Words {
    uint32_t w[sizeof(T) / 4];
  };
  Words b = __builtin_bit_cast(Words, v);
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; ++i)
    b.w[i] = __builtin_amdgcn_readfirstlane(int(b.w[i]));
  return __builtin_bit_cast(T, b);
}

// Sum over the 64 lanes of a wavefront, the same bits in every lane (a
// wave-uniform value).  Within each 16-lane row: DPP exchanges with lane
// xor 1, xor 2 (quad_perm), then the half-row and row mirrors; each step
// adds the same two operands in both lanes, so the row sum is bitwise equal
// across the row.  The four row sums are then combined in the fixed order
// (r0 + r1) + (r2 + r3).  All lanes of the wavefront must be active.
ZEUS_D ZEUS_FI double
wave_sum(double v)
{
  v += dpp_f64<0xB1>(v);  // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E>(v);  // quad_perm [2,3,0,1]
  v += dpp_f64<0x141>(v); // row_half_mirror
  v += dpp_f64<0x140>(v); // row_mirror
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) +
         (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Deterministic workgroup sums of K values at a time.  Two LDS slots are
// used alternately so that back-to-back reductions need one barrier each.
template <int NT>
struct Block {
  static constexpr int kWaves = NT / kWave;
  static_assert(NT % kWave == 0, "workgroup must be whole wavefronts");
  double* red;   // LDS, 2 * 4 * kWaves doubles
  int parity = 0;

  template <int K>
  ZEUS_D ZEUS_FI void sum(double (&v)[K])
  {
    static_assert(K <= 4, "at most 4 values per reduction");
#pragma unroll
    for (int k = 0; k < K; ++k)
      v[k] = wave_sum(v[k]);
    double* r = red + parity * (4 * kWaves);
    parity ^= 1;
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        r[k * kWaves + w] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < kWaves; ++i)
        t += r[k * kWaves + i];
      v[k] = t;
    }
  }
  ZEUS_D ZEUS_FI double sum1(double v)
  {
    double a[1] = {v};
    sum<1>(a);
    return a[0];
  }
};

// Row-chunk layout of a D x D matrix over NT threads: thread t owns chunk
// c = t % CH of the rows r0 + k * RPP (r0 = t / CH, k < R); a chunk is L
// consecutive columns.  Matrix-vector products reduce the CH chunk partial
// sums of a row in a fixed order, so results do not depend on timing.
template <int D, int NT>
struct RowChunk {
  static constexpr int CH = (NT / D) >= 1 ? (NT / D) : 1;
  static constexpr int L = (D + CH - 1) / CH;
  static constexpr int RPP = NT / CH;
  static constexpr int R = (D + RPP - 1) / RPP;
  static constexpr int kPerThread = R * L;
};

// A D x D matrix distributed over the workgroup in the row-chunk layout,
// resident in registers (REG) or in a per-workgroup global slab whose
// element e of thread t sits at slab[e * NT + t] (coalesced).
template <int D, int NT, bool REG>
struct DistMatrix {
  using Lay = RowChunk<D, NT>;
  double m[REG ? Lay::kPerThread : 1];
  double* slab = nullptr;

  ZEUS_D ZEUS_FI double& at(int e)
  {
    if constexpr (REG)
      return m[e];
    else
      return slab[int64_t(e) * NT + threadIdx.x];
  }

  ZEUS_D ZEUS_FI void set_identity()
  {
    const int c = threadIdx.x % Lay::CH, r0 = threadIdx.x / Lay::CH;
#pragma unroll(REG ? 64 : 1)
    for (int k = 0; k < Lay::R; ++k) {
      const int row = r0 + k * Lay::RPP;
#pragma unroll(REG ? 64 : 1)
      for (int l = 0; l < Lay::L; ++l) {
        const int col = c * Lay::L + l;
        at(k * Lay::L + l) = (row == col && row < D) ? 1.0 : 0.0;
      }
    }
  }

  // Load a row-major D x D matrix.
  ZEUS_D ZEUS_FI void load(const double* src)
  {
    const int c = threadIdx.x % Lay::CH, r0 = threadIdx.x / Lay::CH;
#pragma unroll(REG ? 64 : 1)
    for (int k = 0; k < Lay::R; ++k) {
      const int row = r0 + k * Lay::RPP;
#pragma unroll(REG ? 64 : 1)
      for (int l = 0; l < Lay::L; ++l) {
        const int col = c * Lay::L + l;
        at(k * Lay::L + l) = (row < D && col < D) ? src[row * D + col] : 0.0;
      }
    }
  }

  // part[c * D + row] = sum over the chunk of M[row][col] * x[col].
  // With `pending`, first applies M += s a^T + a s^T (formed as
  // s_r a_c + a_r s_c, which is symmetric bit for bit) and stores it back:
  // one read-modify-write pass per BFGS iteration.
  ZEUS_D ZEUS_FI void pass(const double* x, const double* s, const double* a,
                           bool pending, double* part)
  {
    if constexpr (!REG && D > 256) {
      if (pending)
        pass_slab<true>(x, s, a, part);
      else
        pass_slab<false>(x, s, a, part);
    } else {
      pass_regs(x, s, a, pending, part);
    }
  }

  // Register form, and slab form for rows up to 256 elements.
  ZEUS_D ZEUS_FI void pass_regs(const double* x, const double* s, const double* a,
                                bool pending, double* part)
  {
    const int c = threadIdx.x % Lay::CH, r0 = threadIdx.x / Lay::CH;
#pragma unroll(REG ? 64 : 1)
    for (int k = 0; k < Lay::R; ++k) {
      const int row = r0 + k * Lay::RPP;
      if (row < D) {
        const double sr = pending ? s[row] : 0.0;
        const double ar = pending ? a[row] : 0.0;
        double acc = 0.0;
#pragma unroll(REG ? 64 : 1)
        for (int l = 0; l < Lay::L; ++l) {
          const int col = c * Lay::L + l;
          if (col < D) {
            double h = at(k * Lay::L + l);
            if (pending) {
              h = h + (sr * a[col] + ar * s[col]);
              at(k * Lay::L + l) = h;
            }
            acc = ::fma(h, x[col], acc);
          }
        }
        part[c * D + row] = acc;
      }
    }
  }

  // Slab form of pass() for long rows (D > 256): the same arithmetic in the
  // same order, with the columns taken in batches of kBatch whose slab loads
  // are all issued before any of them is used (a branch per element would
  // wait out the memory latency element by element, and at these sizes one
  // wave per SIMD has no other wave to hide it; measured 2x on the 321-D
  // MLP).  Shorter rows keep the element loop of pass_regs, which leaves
  // registers for more resident waves (measured faster at D = 150).  Padding
  // columns (col >= D) are loaded but neither updated nor summed.
  template <bool PENDING>
  ZEUS_D ZEUS_FI void pass_slab(const double* x, const double* s, const double* a, double* part)
  {
    constexpr int kBatch = 16;
    const int c = threadIdx.x % Lay::CH, r0 = threadIdx.x / Lay::CH;
    double* base = slab + threadIdx.x;
    for (int k = 0; k < Lay::R; ++k) {
      const int row = r0 + k * Lay::RPP;
      if (row >= D)
        continue;
      const double sr = PENDING ? s[row] : 0.0;
      const double ar = PENDING ? a[row] : 0.0;
      double* hrow = base + int64_t(k) * Lay::L * NT;
      double acc = 0.0;
      int l0 = 0;
      // one batch per step: unrolled further, the loads of every batch are
      // hoisted together and the register count (and occupancy) suffers
#pragma unroll 1
      for (; l0 + kBatch <= Lay::L; l0 += kBatch) {
        double h[kBatch];
#pragma unroll
        for (int i = 0; i < kBatch; ++i)
          h[i] = hrow[int64_t(l0 + i) * NT];
#pragma unroll
        for (int i = 0; i < kBatch; ++i) {
          const int col = c * Lay::L + l0 + i;
          const bool ok = col < D;
          const int cc = ok ? col : D - 1;
          if constexpr (PENDING) {
            const double hn = h[i] + (sr * a[cc] + ar * s[cc]);
            h[i] = ok ? hn : h[i];
          }
          const double t = ::fma(h[i], x[cc], acc);
          acc = ok ? t : acc;
        }
        if constexpr (PENDING) {
#pragma unroll
          for (int i = 0; i < kBatch; ++i)
            hrow[int64_t(l0 + i) * NT] = h[i];
        }
      }
#pragma unroll 1
      for (int l = l0; l < Lay::L; ++l) {
        const int col = c * Lay::L + l;
        if (col < D) {
          double h = hrow[int64_t(l) * NT];
          if constexpr (PENDING) {
            h = h + (sr * a[col] + ar * s[col]);
            hrow[int64_t(l) * NT] = h;
          }
          acc = ::fma(h, x[col], acc);
        }
      }
      part[c * D + row] = acc;
    }
  }
};

// v[row] = sum_c part[c * D + row], rows owned i = tid, tid + NT, ...
template <int D, int NT>
ZEUS_D ZEUS_FI double
chunk_total(const double* part, int row)
{
  using Lay = RowChunk<D, NT>;
  double t = 0.0;
#pragma unroll
  for (int c = 0; c < Lay::CH; ++c)
    t += part[c * D + row];
  return t;
}

} // namespace zeus::dev
