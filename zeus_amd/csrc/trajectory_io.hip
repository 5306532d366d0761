// trajectory_io.hip -- host writer of recorded trajectories.
//
// Restates reference util::writeTrajectoryData and its TSV writer
// (src/utils.cu:126-286): the file format is chosen by extension; ".tsv"
// writes one row per (start, step), starts outer, with columns
//   traj  step  fval  grad  status  alpha  x0 .. x{dim-1}
// numbers in scientific notation with 17 digits after the point, NaN
// values of fval, grad and the coordinates as "NaN", status as an integer.
// NetCDF4 output is not built in (the reference's behaviour when built
// without it: an error); any other extension is an error.
#include "zeus_c.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

bool
ends_with(const char* s, const char* suffix)
{
  const size_t n = std::strlen(s), m = std::strlen(suffix);
  return n >= m && std::strcmp(s + n - m, suffix) == 0;
}

// "%.17e", or "NaN" where the reference writes it literally
int
put_num(char* p, size_t cap, double v, bool nan_literal)
{
  if (nan_literal && std::isnan(v))
    return std::snprintf(p, cap, "\tNaN");
  return std::snprintf(p, cap, "\t%.17e", v);
}

} // namespace

extern "C" zeus_status_t
zeus_write_trajectory(const char* filename, const double* coords, const double* fval,
                      const double* grad, const int8_t* status, const double* alpha, int32_t n,
                      int32_t steps, int32_t dim)
{
  if (!filename || n < 0 || steps < 0 || dim <= 0)
    return ZEUS_E_ARG;
  if (ends_with(filename, ".nc"))
    return ZEUS_E_UNSUPPORTED;
  if (!ends_with(filename, ".tsv"))
    return ZEUS_E_ARG;
  if (int64_t(n) * steps > 0 && (!coords || !fval || !grad || !status))
    return ZEUS_E_ARG;
  FILE* out = std::fopen(filename, "w");
  if (!out)
    return ZEUS_E_IO;
  std::vector<char> iobuf(1 << 20);
  std::setvbuf(out, iobuf.data(), _IOFBF, iobuf.size());
  bool ok = std::fputs("traj\tstep\tfval\tgrad\tstatus\talpha", out) >= 0;
  for (int d = 0; d < dim && ok; ++d)
    ok = std::fprintf(out, "\tx%d", d) > 0;
  ok = ok && std::fputc('\n', out) != EOF;
  std::vector<char> line(size_t(dim + 6) * 32 + 64);
  for (int i = 0; i < n && ok; ++i) {
    for (int it = 0; it < steps && ok; ++it) {
      const int64_t o = int64_t(it) * n + i;
      char* p = line.data();
      const size_t cap = line.size();
      size_t k = size_t(std::snprintf(p, cap, "%d\t%d", i, it));
      k += size_t(put_num(p + k, cap - k, fval[o], true));
      k += size_t(put_num(p + k, cap - k, grad[o], true));
      k += size_t(std::snprintf(p + k, cap - k, "\t%d", int(status[o])));
      k += size_t(put_num(p + k, cap - k, alpha ? alpha[o] : 0.0, false));
      for (int d = 0; d < dim; ++d)
        k += size_t(put_num(p + k, cap - k, coords[(int64_t(it) * dim + d) * n + i], true));
      p[k++] = '\n';
      ok = std::fwrite(p, 1, k, out) == k;
    }
  }
  ok = (std::fclose(out) == 0) && ok;
  return ok ? ZEUS_OK : ZEUS_E_IO;
}
