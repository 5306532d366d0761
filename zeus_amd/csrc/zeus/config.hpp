// config.hpp -- common macros for the MI355X (gfx950) Zeus implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#define ZEUS_HD __host__ __device__
#define ZEUS_D __device__
#define ZEUS_FI __forceinline__

namespace zeus {

// Status codes of a start / of a call (reference README.md:27-35).
enum Status : int {
  kStatusUnset = -1,
  kStatusSurrendered = 0,
  kStatusConverged = 1,
  kStatusStopped = 2,
  kStatusMalloc = 3,
  kStatusKernel = 4,
  kStatusNonFinite = 5,
  kStatusZeroSteps = 6, // kept for API parity; unreachable (see line search)
};

// PSO coefficients (reference include/pso.cuh:302).
constexpr double kPsoW = 0.5;
constexpr double kPsoC1 = 1.2;
constexpr double kPsoC2 = 1.5;
// Line search (reference include/utils.cuh:389-395).
constexpr double kArmijoC1 = 0.1;
constexpr int kLineSearchTrials = 20;
// BFGS update is skipped when |s.y| is below this (include/utils.cuh:254).
constexpr double kCurvatureSkip = 1e-14;

constexpr int kWave = 64; // CDNA wavefront width

} // namespace zeus
