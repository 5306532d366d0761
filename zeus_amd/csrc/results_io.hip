// results_io.hip -- host-side reporting of a call: timing metrics from
// per-start clock ticks, the distance to the known minimum, and the
// per-dimension results table.
//
// Restates reference report_metrics_and_cleanup (bfgs_sequential.cuh:
// 240-382), util::calculateEuclideanError (src/util.cpp:11-41) and
// util::appendResultsToTsv (include/util.hpp:78-149): same quantities, same
// columns, same number formats (the stream defaults there are printf's %g,
// and %e after std::scientific).
#include "zeus_c.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

extern "C" zeus_status_t
zeus_metrics_from_ticks(const uint64_t* ticks, const int32_t* calls, const uint64_t* total,
                        int64_t n, int32_t threads_per_block, double ms_kernel, double tick_hz,
                        zeus_metrics* out)
{
  if (!out || n < 0 || (n > 0 && (!ticks || !calls)) || !(tick_hz > 0.0))
    return ZEUS_E_ARG;
  *out = zeus_metrics{};
  if (n == 0)
    return ZEUS_OK;
  unsigned long long sum_ticks = 0;
  long long sum_calls = 0;
  for (int64_t i = 0; i < n; ++i) {
    sum_ticks += ticks[i];
    sum_calls += calls[i];
  }
  const double avg_ticks = double(sum_ticks) / double(n);
  const double avg_calls = sum_calls > 0 ? double(sum_calls) / double(n) : 0.0;
  const double avg_ticks_per_call = avg_calls > 0 ? avg_ticks / avg_calls : 0.0;
  const double avg_ms = avg_ticks / tick_hz * 1e3;
  out->ms_per_call = avg_ticks_per_call / tick_hz * 1e3;
  out->calls_per_thread_mean = avg_calls;
  out->fraction_of_kernel = (avg_ms > 0.0 && ms_kernel > 0.0) ? avg_ms / ms_kernel : 0.0;
  out->serialized = ms_kernel > 0.0 ? (double(sum_ticks) / tick_hz * 1e3) / ms_kernel : 0.0;
  if (total) {
    long double acc = 0.0L;
    long long cnt = 0;
    for (int64_t i = 0; i < n; ++i)
      if (total[i] > 0) {
        acc += (long double)(double(ticks[i]) / double(total[i]));
        ++cnt;
      }
    out->fraction_of_thread = cnt > 0 ? double(acc / (long double)cnt) : 0.0;
  }
  if (threads_per_block > 0 && ms_kernel > 0.0) {
    const int64_t nb = (n + threads_per_block - 1) / threads_per_block;
    std::vector<long double> bt(static_cast<size_t>(nb), 0.0L);
    for (int64_t i = 0; i < n; ++i)
      bt[size_t(i / threads_per_block)] += (long double)ticks[i];
    std::vector<double> fb(static_cast<size_t>(nb));
    for (int64_t b = 0; b < nb; ++b)
      fb[size_t(b)] = double(bt[size_t(b)] / tick_hz * 1e3L) / ms_kernel;
    std::sort(fb.begin(), fb.end());
    const size_t i95 =
      std::min<size_t>(fb.size() - 1, size_t(std::ceil(0.95 * double(fb.size()))) - 1);
    out->block95 = fb[i95];
  }
  return ZEUS_OK;
}

extern "C" double
zeus_euclidean_error(const char* fun_name, const double* x, int32_t dim)
{
  const std::string name = fun_name ? fun_name : "";
  double s = 0.0;
  if (name == "rosenbrock") {
    for (int i = 0; i < dim; ++i) {
      const double d = x[i] - 1.0;
      s += d * d;
    }
  } else if (name == "goldstein_price") {
    if (dim != 2)
      return std::nan("");
    const double dx = x[0] - 0.0, dy = x[1] - (-1.0);
    s = dx * dx + dy * dy;
  } else {
    for (int i = 0; i < dim; ++i)
      s += x[i] * x[i];
  }
  return std::sqrt(s);
}

extern "C" zeus_status_t
zeus_append_results_tsv(const char* path, int32_t dim, int64_t n, const char* fun_name,
                        double ms_init, double ms_pso, double ms_opt, double ms_rand,
                        int32_t max_iter, int32_t pso_iter, double error, int32_t run,
                        const zeus_result* best, const double* coords)
{
  if (dim < 0 || !best || (dim > 0 && !coords))
    return ZEUS_E_ARG;
  const std::string file =
    path ? std::string(path) : "zeus_" + std::to_string(dim) + "d_results.tsv";
  FILE* out = std::fopen(file.c_str(), "a");
  if (!out)
    return ZEUS_E_IO;
  std::fseek(out, 0, SEEK_END);
  const bool empty = std::ftell(out) == 0;
  bool ok = true;
  if (empty) {
    ok = std::fputs("fun\trun\tN\tkernel"
                    "\tad_ms_per_call\tad_calls_per_thread_mean\tad_fraction\tad_block95"
                    "\tad_serialized"
                    "\tbfgs_ms_per_call\tbfgs_calls_per_thread_mean\tbfgs_fraction"
                    "\tbfgs_block95\tbfgs_serialized"
                    "\tclaimed\tactual\tsurrender\tstopped\tidx\tstatus\tbfgs_iter\tpso_iter"
                    "\ttime\terror\tfval\tnorm",
                    out) >= 0;
    for (int i = 0; i < dim && ok; ++i)
      ok = std::fprintf(out, "\tcoord_%d", i) > 0;
    ok = ok && std::fputc('\n', out) != EOF;
  }
  // the reference passes these times as float and sums them in float
  const float fi = float(ms_init), fp = float(ms_pso), fo = float(ms_opt), fr = float(ms_rand);
  const float t = pso_iter > 0 ? (fi + fp + fo + fr) : (fo + fr);
  const zeus_metrics& a = best->ad;
  const zeus_metrics& b = best->bfgs;
  ok = ok &&
       std::fprintf(out,
                    "%s\t%d\t%lld\t%g\t%g\t%g\t%g\t%g\t%g\t%g\t%g\t%g\t%g\t%g\t%d\t%d\t%d\t%d\t%d"
                    "\t%d\t%d\t%d\t%g\t%e\t%e\t%e\t",
                    fun_name ? fun_name : "", run, (long long)n, double(fo), a.ms_per_call,
                    a.calls_per_thread_mean, a.fraction_of_thread, a.block95, a.serialized,
                    b.ms_per_call, b.calls_per_thread_mean, b.fraction_of_thread, b.block95,
                    b.serialized, best->claimed, best->actual, best->surrendered,
                    best->stopped, best->idx, best->status, max_iter, pso_iter, double(t),
                    error, best->fval, best->gradient_norm) > 0;
  for (int i = 0; i < dim && ok; ++i)
    ok = std::fprintf(out, i < dim - 1 ? "%e\t" : "%e", coords[i]) > 0;
  ok = ok && std::fputc('\n', out) != EOF;
  ok = (std::fclose(out) == 0) && ok;
  return ok ? ZEUS_OK : ZEUS_E_IO;
}
