// hip_buffer.hpp -- owning device arrays (the role of CudaBuffer<T>,
// DoubleBuffer and ResultBuffer<D>, reference include/cuda_buffer.hpp:14-209
// and include/util.hpp:204-207).
//
// Contract kept from the reference: size 0 owns no memory; a failed
// allocation throws HipError; copy is a deep device-to-device copy; move
// leaves the source empty; copyToHost(vector&) returns 0 or 4 and
// copyToHost(T*, n) returns 1 on a size mismatch; == compares pointer and
// size.  Memory comes from the library's C ABI (zeus_alloc / zeus_free).
#pragma once

#include "exception.hpp"
#include "result.hpp"
#include "zeus_c.h"

#include <array>
#include <cstddef>
#include <utility>
#include <vector>

namespace zeus {

template <typename T>
struct HipBuffer {
  T* d = nullptr;
  std::size_t sz = 0;

  HipBuffer() noexcept = default;
  explicit HipBuffer(std::size_t n) : sz(n)
  {
    if (sz > 0) {
      void* p = nullptr;
      const int rc = zeus_alloc(&p, sz * sizeof(T));
      if (rc != ZEUS_OK) {
        sz = 0;
        throw HipError(hip_error_of(rc), "zeus_alloc failed");
      }
      d = static_cast<T*>(p);
    }
  }
  template <std::size_t N>
  explicit HipBuffer(const std::array<T, N>& host) : HipBuffer(N)
  {
    if (N > 0 && zeus_memcpy_h2d(d, host.data(), N * sizeof(T)) != ZEUS_OK) {
      release();
      throw HipError(hipErrorInvalidValue, "host-to-device copy failed");
    }
  }
  HipBuffer(const HipBuffer& o) : HipBuffer(o.sz)
  {
    if (sz > 0 && zeus_memcpy_d2d(d, o.d, sz * sizeof(T)) != ZEUS_OK) {
      release();
      throw HipError(hipErrorInvalidValue, "device-to-device copy failed");
    }
  }
  HipBuffer& operator=(const HipBuffer& o)
  {
    if (this != &o) {
      HipBuffer t(o);
      swap(t);
    }
    return *this;
  }
  HipBuffer(HipBuffer&& o) noexcept : d(o.d), sz(o.sz)
  {
    o.d = nullptr;
    o.sz = 0;
  }
  HipBuffer& operator=(HipBuffer&& o) noexcept
  {
    if (this != &o) {
      release();
      d = o.d;
      sz = o.sz;
      o.d = nullptr;
      o.sz = 0;
    }
    return *this;
  }
  ~HipBuffer() { release(); }

  void swap(HipBuffer& o) noexcept
  {
    std::swap(d, o.d);
    std::swap(sz, o.sz);
  }
  T* data() const noexcept { return d; }
  operator T*() const noexcept { return d; }
  std::size_t size() const noexcept { return sz; }

  std::vector<T> copyToHost() const
  {
    std::vector<T> out(sz);
    if (sz > 0 && zeus_memcpy_d2h(out.data(), d, sz * sizeof(T)) != ZEUS_OK)
      throw HipError(hipErrorInvalidValue, "device-to-host copy failed");
    return out;
  }
  int copyToHost(std::vector<T>& out) const
  {
    out.resize(sz);
    if (sz > 0 && zeus_memcpy_d2h(out.data(), d, sz * sizeof(T)) != ZEUS_OK)
      return 4;
    return 0;
  }
  int copyToHost(T* out, std::size_t n) const
  {
    if (n != sz)
      return 1;
    if (sz > 0 && zeus_memcpy_d2h(out, d, sz * sizeof(T)) != ZEUS_OK)
      return 4;
    return 0;
  }

private:
  void release() noexcept
  {
    if (d)
      (void)zeus_free(d);
    d = nullptr;
    sz = 0;
  }
};

template <typename T>
bool
operator==(const HipBuffer<T>& a, const HipBuffer<T>& b)
{
  return a.d == b.d && a.sz == b.sz;
}

using DoubleBuffer = HipBuffer<double>;
template <std::size_t D>
using ResultBuffer = HipBuffer<Result<D>>;

} // namespace zeus
