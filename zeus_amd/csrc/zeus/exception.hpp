// exception.hpp -- HipError (the role of zeus::CudaError,
// reference include/exception.hpp:9-23, src/exception.cpp:5-24).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace zeus {

class HipError : public std::runtime_error {
public:
  explicit HipError(hipError_t error, const std::string& msg = "")
    : std::runtime_error(format(error, msg)), code_(error)
  {}
  hipError_t code() const noexcept { return code_; }

private:
  hipError_t code_;
  static std::string format(hipError_t e, const std::string& msg)
  {
    std::string s = "HIP Error: ";
    s += hipGetErrorString(e);
    if (!msg.empty())
      s += " (" + msg + ")";
    return s;
  }
};

// Maps a C ABI return code onto the HIP error it stands for.
inline hipError_t
hip_error_of(int zeus_status)
{
  return zeus_status == 3 ? hipErrorOutOfMemory
                          : (zeus_status == 0 ? hipSuccess : hipErrorLaunchFailure);
}

} // namespace zeus
