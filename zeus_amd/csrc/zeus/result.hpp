// result.hpp -- Result<D> and its companions (reference include/util.hpp:16-54).
#pragma once

#include <cstddef>

struct Metrics {
  const char* label = "";
  double ms_per_call = 0.0;
  double calls_per_thread_mean = 0.0;
  double fraction_of_kernel = 0.0;
  double block95 = 0.0;
  double serialized = 0.0;
  double fraction_of_thread = 0.0;
};

// Convergence counters of a call (reference util.hpp:151-200):
// claimed = status 1; actual = status 1 within 0.5 (Euclidean) of the known
// minimum implied by the function name; stopped = status 2; surrendered =
// every other status.
struct Convergence {
  int actual = 0;
  int claimed = 0;
  int surrendered = 0;
  int stopped = 0;
};

namespace zeus {

template <std::size_t DIM>
struct Result {
  int idx = -1;
  // -1 unset, 0 surrendered, 1 converged, 2 stopped early,
  // 3 allocation failure, 4 launch/runtime error, 5 non-finite
  int status = -1;
  double fval = 37.0;
  double gradientNorm = 0.0;
  double coordinates[DIM]{};
  int iter = 0;
  Convergence c;
  double ms_opt = 0.0; // BFGS phase time of the call (filled here)
  Metrics ad;
  Metrics bfgs;
};

} // namespace zeus
