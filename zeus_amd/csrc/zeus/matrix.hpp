// matrix.hpp -- host matrix with a device mirror, the upload path objectives
// use for their data (the role of Matrix<T>, reference include/matrix.hpp:20-258).
//
// Kept contract: zero dimensions throw std::invalid_argument; set(i, j, x),
// set(std::array) and set(ptr, n) write the host copy and push it to the
// device; operator() is the bounds-checked host accessor; data() is the
// device pointer in device code and the host pointer in host code; a host
// copy is deep (host and device), a device copy is shallow.
// Kernels of this library do not copy functors per launch: the functor's
// bytes are uploaded once per call and read through a pointer, so the deep
// copy only runs when user code copies a Matrix on the host.
#pragma once

#include "exception.hpp"
#include "zeus_c.h"

#include <hip/hip_runtime.h>

#include <array>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <utility>

namespace zeus {

template <typename T>
class Matrix {
public:
  Matrix() noexcept = default;

  Matrix(std::size_t rows, std::size_t cols) : rows_(rows), cols_(cols)
  {
    if (!rows_ || !cols_)
      throw std::invalid_argument("Matrix dimensions must be > 0");
    alloc_both();
  }
  template <std::size_t N>
  Matrix(std::size_t rows, std::size_t cols, const std::array<T, N>& init)
    : Matrix(rows, cols)
  {
    set(init);
  }

  __host__ __device__ Matrix(const Matrix& o) : rows_(o.rows_), cols_(o.cols_)
  {
#ifdef __HIP_DEVICE_COMPILE__
    host_ = o.host_;
    dev_ = o.dev_;
#else
    if (!rows_ || !cols_)
      return;
    alloc_both();
    std::memcpy(host_, o.host_, bytes());
    push();
#endif
  }
  Matrix& operator=(const Matrix& o)
  {
    if (this != &o) {
      Matrix t(o);
      swap(t);
    }
    return *this;
  }
  Matrix(Matrix&& o) noexcept
    : host_(o.host_), dev_(o.dev_), rows_(o.rows_), cols_(o.cols_)
  {
    o.host_ = nullptr;
    o.dev_ = nullptr;
    o.rows_ = o.cols_ = 0;
  }
  Matrix& operator=(Matrix&& o) noexcept
  {
    if (this != &o) {
      Matrix t(std::move(o));
      swap(t);
    }
    return *this;
  }
  __host__ __device__ ~Matrix()
  {
#ifndef __HIP_DEVICE_COMPILE__
    if (host_)
      std::free(host_);
    if (dev_)
      (void)zeus_free(dev_);
#endif
  }

  void set(std::size_t i, std::size_t j, const T& x)
  {
    check(i, j);
    const std::size_t k = i * cols_ + j;
    host_[k] = x;
    if (zeus_memcpy_h2d(dev_ + k, &x, sizeof(T)) != ZEUS_OK)
      throw HipError(hipErrorInvalidValue, "Matrix::set upload failed");
  }
  template <std::size_t N>
  void set(const std::array<T, N>& a)
  {
    if (N != rows_ * cols_)
      throw std::invalid_argument("std::array size must equal rows*cols");
    std::memcpy(host_, a.data(), bytes());
    push();
  }
  void set(const T* p, std::size_t count)
  {
    if (count != rows_ * cols_)
      throw std::invalid_argument("count must equal rows*cols");
    std::memcpy(host_, p, bytes());
    push();
  }

  __host__ __device__ const T* data() const noexcept
  {
#ifdef __HIP_DEVICE_COMPILE__
    return dev_;
#else
    return host_;
#endif
  }
  T* hostData() noexcept { return host_; }
  const T* hostData() const noexcept { return host_; }
  T* deviceData() noexcept { return dev_; }
  const T* deviceData() const noexcept { return dev_; }
  T& operator()(std::size_t i, std::size_t j)
  {
    check(i, j);
    return host_[i * cols_ + j];
  }
  const T& operator()(std::size_t i, std::size_t j) const
  {
    check(i, j);
    return host_[i * cols_ + j];
  }
  __host__ __device__ std::size_t rows() const noexcept { return rows_; }
  __host__ __device__ std::size_t cols() const noexcept { return cols_; }
  void swap(Matrix& o) noexcept
  {
    std::swap(host_, o.host_);
    std::swap(dev_, o.dev_);
    std::swap(rows_, o.rows_);
    std::swap(cols_, o.cols_);
  }

private:
  T* host_ = nullptr;
  T* dev_ = nullptr;
  std::size_t rows_ = 0;
  std::size_t cols_ = 0;

  std::size_t bytes() const { return rows_ * cols_ * sizeof(T); }
  void check(std::size_t i, std::size_t j) const
  {
    if (i >= rows_ || j >= cols_)
      throw std::out_of_range("Matrix index out of range");
  }
  void alloc_both()
  {
    host_ = static_cast<T*>(std::malloc(bytes()));
    if (!host_)
      throw std::runtime_error("host malloc failed");
    void* p = nullptr;
    const int rc = zeus_alloc(&p, bytes());
    if (rc != ZEUS_OK) {
      std::free(host_);
      host_ = nullptr;
      throw HipError(hip_error_of(rc), "Matrix device allocation failed");
    }
    dev_ = static_cast<T*>(p);
  }
  void push()
  {
    if (zeus_memcpy_h2d(dev_, host_, bytes()) != ZEUS_OK)
      throw HipError(hipErrorInvalidValue, "Matrix upload failed");
  }
};

} // namespace zeus

// The reference exposes Matrix at global scope.
using zeus::Matrix;
