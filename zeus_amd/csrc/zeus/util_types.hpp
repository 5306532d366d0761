// util_types.hpp -- NonNull, DeviceMatrix and setStackSize: API-parity types
// of the reference (include/utils.cuh:20-66, include/device_matrix.cuh:8-118,
// src/utils.cu:52-99).  None of them is used on the optimizer's hot path:
// this implementation keeps H in registers / preallocated slabs and needs no
// enlarged device stack or heap.
#pragma once

#include <hip/hip_runtime.h>

#include <concepts>
#include <cstddef>
#include <stdexcept>

namespace util {

// Pointer that must not be null: host construction from nullptr throws
// std::invalid_argument; converts implicitly to the raw pointer.
template <typename T>
struct NonNull {
  T ptr;
  __host__ __device__ explicit NonNull(const T p) : ptr(p)
  {
#ifndef __HIP_DEVICE_COMPILE__
    if (p == nullptr)
      throw std::invalid_argument("util::NonNull: construction from nullptr");
#endif
  }
  template <typename U>
    requires std::convertible_to<U, T>
  __host__ __device__ NonNull(const NonNull<U>& o) : ptr(o.get())
  {}
  NonNull(std::nullptr_t) = delete;
  NonNull& operator=(std::nullptr_t) = delete;
  __host__ __device__ T get() const { return ptr; }
  __host__ __device__ T operator->() const { return ptr; }
  __host__ __device__ auto& operator*() const { return *ptr; }
  __host__ __device__ operator T() const { return ptr; }
};

// Compatible no-op: the reference enlarges the device stack (64 KiB) and the
// malloc heap for its per-thread H; neither is needed here.  The stack limit
// is still raised so that user functors with deep recursion behave alike.
inline void
setStackSize()
{
  (void)hipDeviceSetLimit(hipLimitStackSize, 64 * 1024);
}

} // namespace util

// Device-heap matrix (device code only), kept for API parity.
template <typename T>
class DeviceMatrix {
public:
  __device__ DeviceMatrix(int rows, int cols) : rows_(rows), cols_(cols)
  {
    data_ = static_cast<T*>(malloc(static_cast<std::size_t>(rows_) * cols_ * sizeof(T)));
    if (!data_)
      rows_ = cols_ = 0; // callers check valid(); no trap
  }
  DeviceMatrix(const DeviceMatrix&) = delete;
  DeviceMatrix& operator=(const DeviceMatrix&) = delete;
  DeviceMatrix(DeviceMatrix&&) = delete; // the storage is owned by one object
  DeviceMatrix& operator=(DeviceMatrix&&) = delete;
  __device__ ~DeviceMatrix() { release(); }
  __device__ void release()
  {
    if (data_) {
      free(data_);
      data_ = nullptr;
      rows_ = cols_ = 0;
    }
  }
  __device__ bool valid() const { return data_ != nullptr; }
  __device__ T* data() { return data_; }
  __device__ const T* data() const { return data_; }
  __device__ int rows() const { return rows_; }
  __device__ int cols() const { return cols_; }
  __device__ T& operator()(int i, int j) { return data_[static_cast<std::size_t>(i) * cols_ + j]; }
  __device__ const T& operator()(int i, int j) const
  {
    return data_[static_cast<std::size_t>(i) * cols_ + j];
  }
  __device__ void set(int i, int j, const T& x) { (*this)(i, j) = x; }

private:
  T* data_ = nullptr;
  int rows_ = 0;
  int cols_ = 0;
};
