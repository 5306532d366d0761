/*
 * zeus_oracle.c -- CPU restatement of the reference multistart PSO -> BFGS
 * method.  TEST INFRASTRUCTURE ONLY (the checker); see zeus_oracle.h for what
 * pins each function.  Every function cites the reference file:line it
 * restates.  Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "zeus_oracle.h"

#include <math.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ======================================================================
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Replaces the cuRAND
 * states of src/utils.cu:17-50 / include/bfgs_common.cuh:15-81 with a
 * stateless counter-based stream keyed by (seed, stream); counter =
 * (draw/2, start).  Constants are the paper's published round multipliers
 * and Weyl key increments.
 * ==================================================================== */
void
orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* Draw `draw` of start `start`: words (2j, 2j+1), j = draw & 1, of block
   draw >> 1; top 53 bits -> (m + 1) * 2^-53 in (0, 1].  cuRAND's
   curand_uniform_double range (0,1] is kept (include/utils.cuh:353-358). */
double
orc_uniform(uint32_t seed, uint32_t stream, uint64_t start, uint64_t draw)
{
  uint64_t blk = draw >> 1;
  uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)start,
                     (uint32_t)(start >> 32)};
  uint32_t key[2] = {seed, stream};
  uint32_t o[4];
  orc_philox4x32_10(ctr, key, o);
  int j = (int)(draw & 1u) * 2;
  uint64_t m = ((((uint64_t)o[j]) << 32) | o[j + 1]) >> 11;
  return (double)(m + 1u) * 0x1p-53;
}

/* ---- selectable generators (include/traits.hpp:67-71) ----------------
   XORWOW: Marsaglia, "Xorshift RNGs", J. Stat. Softw. 8(14), 2003; one
   state per (start, coordinate), seeded from Philox counter blocks
   (2c, 2^32-1) and (2c+1, 2^32-1) of the start (never used by draws), giving
   that coordinate's draws in order (PSO init position and velocity, then r1
   and r2 of each iteration, or one BFGS random-start draw); a uniform takes
   two outputs, (m + 1) 2^-53 with m = top 32 bits of the first and top 21
   of the second.
   SOBOL: dimension 1 (direction numbers 2^(31-j)): point n is
   bitreverse32(n ^ (n >> 1)); draw k of start i is point i + k; (x + 1) 2^-32. */
void
orc_xorwow_seed(uint32_t seed, uint32_t stream, uint64_t start, uint32_t coord,
                uint32_t st[6])
{
  uint32_t key[2] = {seed, stream};
  uint32_t ca[4] = {2u * coord, 0xFFFFFFFFu, (uint32_t)start, (uint32_t)(start >> 32)};
  uint32_t cb[4] = {2u * coord + 1u, 0xFFFFFFFFu, (uint32_t)start, (uint32_t)(start >> 32)};
  uint32_t a[4], b[4];
  orc_philox4x32_10(ca, key, a);
  orc_philox4x32_10(cb, key, b);
  st[0] = a[0];
  st[1] = a[1];
  st[2] = a[2];
  st[3] = a[3];
  st[4] = b[0];
  st[5] = b[1];
  if ((st[0] | st[1] | st[2] | st[3] | st[4]) == 0u)
    st[4] = 0x2545F491u;
}

uint32_t
orc_xorwow_next(uint32_t st[6])
{
  uint32_t t = st[0] ^ (st[0] >> 2);
  st[0] = st[1];
  st[1] = st[2];
  st[2] = st[3];
  st[3] = st[4];
  st[4] = (st[4] ^ (st[4] << 4)) ^ (t ^ (t << 1));
  st[5] += 362437u;
  return st[4] + st[5];
}

static double
xorwow_uniform(uint32_t st[6])
{
  uint64_t hi = orc_xorwow_next(st);
  uint64_t lo = orc_xorwow_next(st);
  uint64_t m = (hi << 21) | (lo >> 11);
  return (double)(m + 1u) * 0x1p-53;
}

double
orc_sobol_uniform(uint64_t start, uint64_t draw)
{
  uint32_t n = (uint32_t)(start + draw);
  uint32_t v = n ^ (n >> 1), r = 0;
  for (int b = 0; b < 32; ++b)
    r |= ((v >> b) & 1u) << (31 - b);
  return ((double)r + 1.0) * 0x1p-32;
}

/* Draw `draw` of start `start` for the stateless generators. */
static double
index_uniform(const orc_problem* P, uint64_t start, uint64_t draw)
{
  return P->prng == ORC_PRNG_SOBOL ? orc_sobol_uniform(start, draw)
                                   : orc_uniform(P->seed, P->stream, start, draw);
}

/* Uniforms u0, u1 of coordinate d of start i at draw index k (= 2Dt + 2d):
   PHILOX / SOBOL by index, XORWOW from the (start, coordinate) state. */
static void
pso_pair(const orc_problem* P, uint64_t i, uint64_t k, uint32_t* xs, double* u0, double* u1)
{
  if (P->prng == ORC_PRNG_XORWOW) {
    *u0 = xorwow_uniform(xs);
    *u1 = xorwow_uniform(xs);
  } else {
    *u0 = index_uniform(P, i, k);
    *u1 = index_uniform(P, i, k + 1u);
  }
}

/* include/utils.cuh:353-358: lower + (upper + (-lower)) * u */
static double
to_range(double u, double lo, double hi)
{
  return lo + (hi + (-lo)) * u;
}

/* Coordinate d of a BFGS random start of start i (draw index d, or the
   first draw of the (start, coordinate) XORWOW stream). */
static double
start_uniform_in(const orc_problem* P, uint64_t i, int d)
{
  if (P->prng == ORC_PRNG_XORWOW) {
    uint32_t st[6];
    orc_xorwow_seed(P->seed, P->stream, i, (uint32_t)d, st);
    return to_range(xorwow_uniform(st), P->lower, P->upper);
  }
  return to_range(index_uniform(P, i, (uint64_t)d), P->lower, P->upper);
}

int
orc_stream_uniforms(int prng, uint32_t seed, uint32_t stream, uint64_t start, uint32_t coord,
                    int64_t n, double* out)
{
  if (prng < 0 || prng > 2 || n < 0 || (n > 0 && !out))
    return -1;
  orc_problem P;
  memset(&P, 0, sizeof(P));
  P.prng = prng;
  P.seed = seed;
  P.stream = stream;
  uint32_t st[6];
  orc_xorwow_seed(seed, stream, start, coord, st);
  for (int64_t k = 0; k < n; ++k)
    out[k] = prng == ORC_PRNG_XORWOW ? xorwow_uniform(st) : index_uniform(&P, start, (uint64_t)k);
  return 0;
}

/* ======================================================================
 * Dual numbers: include/duals.cuh:14-239, digamma src/duals.cu:8-39.
 * ==================================================================== */
double
orc_digamma(double x)
{
  if (x <= 0.0 && x == floor(x))
    return INFINITY;
  if (x < 0.5)
    return orc_digamma(1.0 - x) - ORC_PI / tan(ORC_PI * x);
  double acc = 0.0;
  while (x < 8.0) {
    acc -= 1.0 / x;
    x += 1.0;
  }
  const double inv = 1.0 / x;
  const double inv2 = inv * inv;
  const double s =
    inv2 *
    (-1.0 / 12.0 +
     inv2 * (1.0 / 120.0 +
             inv2 * (-1.0 / 252.0 +
                     inv2 * (1.0 / 240.0 +
                             inv2 * (-1.0 / 132.0 + inv2 * (691.0 / 32760.0))))));
  return acc + log(x) - 0.5 * inv + s;
}

static inline orc_dual
dn(double r, double d)
{
  orc_dual z = {r, d};
  return z;
}
static inline orc_dual
dn_add(orc_dual a, orc_dual b)
{
  return dn(a.re + b.re, a.du + b.du);
}
static inline orc_dual
dn_sub(orc_dual a, orc_dual b)
{
  return dn(a.re - b.re, a.du - b.du);
}
static inline orc_dual
dn_mul(orc_dual a, orc_dual b)
{
  return dn(a.re * b.re, a.du * b.re + a.re * b.du);
}
static inline orc_dual
dn_div(orc_dual a, orc_dual b)
{
  double denom = b.re * b.re;
  return dn(a.re / b.re, (a.du * b.re - a.re * b.du) / denom);
}
static inline orc_dual
dn_smul(double l, orc_dual r) /* operator*(double, DualNumber) */
{
  return dn(l * r.re, l * r.du);
}
static inline orc_dual
dn_cos(orc_dual x)
{
  return dn(cos(x.re), -x.du * sin(x.re));
}
static inline orc_dual
dn_sin(orc_dual x)
{
  return dn(sin(x.re), x.du * cos(x.re));
}
static inline orc_dual
dn_exp(orc_dual x)
{
  double ex = exp(x.re);
  return dn(ex, x.du * ex);
}
static inline orc_dual
dn_sqrt(orc_dual x)
{
  double sr = sqrt(x.re);
  return dn(sr, x.du / (2.0 * sr));
}
static inline orc_dual
dn_log(orc_dual x)
{
  return dn(log(x.re), x.du / x.re);
}
static inline orc_dual
dn_pow(orc_dual b, orc_dual e)
{
  const double br = b.re, er = e.re;
  const double pr = pow(br, er);
  const double pd = pr * (e.du * log(br) + er * b.du / br);
  return dn(pr, pd);
}
static inline orc_dual
dn_powd(orc_dual b, double e)
{
  const double br = b.re;
  const double pr = pow(br, e);
  if (br == 0.0)
    return dn(pr, e * pow(br, e - 1.0) * b.du);
  return dn(pr, e * pr / br * b.du);
}
static inline orc_dual
dn_lgamma(orc_dual x)
{
  return dn(lgamma(x.re), x.du * orc_digamma(x.re));
}
/* tanh is not in the reference's dual set; d tanh = (1 - tanh^2) dx. */
static inline orc_dual
dn_tanh(orc_dual x)
{
  double t = tanh(x.re);
  return dn(t, x.du * (1.0 - t * t));
}

orc_dual
orc_dual_apply(int op, orc_dual a, orc_dual b, double e)
{
  switch (op) {
    case ORC_D_ADD: return dn_add(a, b);
    case ORC_D_SUB: return dn_sub(a, b);
    case ORC_D_MUL: return dn_mul(a, b);
    case ORC_D_DIV: return dn_div(a, b);
    case ORC_D_NEG: return dn(-a.re, -a.du);
    case ORC_D_ABS:
      if (a.re == 0.0)
        return dn(0.0, NAN);
      return (a.re < 0.0) ? dn(-a.re, -a.du) : a;
    case ORC_D_SIN: return dn_sin(a);
    case ORC_D_COS: return dn_cos(a);
    case ORC_D_EXP: return dn_exp(a);
    case ORC_D_SQRT: return dn_sqrt(a);
    case ORC_D_ATAN2: {
      /* atan2(a = y, b = x) */
      double denom = b.re * b.re + a.re * a.re;
      return dn(atan2(a.re, b.re), (b.re * a.du - a.re * b.du) / denom);
    }
    case ORC_D_LOG: return dn_log(a);
    case ORC_D_POW: return dn_pow(a, b);
    case ORC_D_POWD: return dn_powd(a, e);
    case ORC_D_LGAMMA: return dn_lgamma(a);
    case ORC_D_TANH: return dn_tanh(a);
    default: return dn(NAN, NAN);
  }
}

/* ======================================================================
 * Objectives.  Each has a double path, a dual path that follows the
 * reference's own dual overload, and a hand-written reverse-mode adjoint
 * whose forward sweep repeats the double path operation for operation (so
 * its value equals orc_value bit for bit).
 * ==================================================================== */

/* --- Rosenbrock: examples/rosenbrock.hpp:19-30 (double), 6-17 (dual) --- */
static double
rosen_value(const double* x, int D)
{
  double sum = 0.0;
  for (int i = 0; i < D - 1; ++i) {
    double t1 = 1.0 - x[i];
    double t2 = x[i + 1] - x[i] * x[i];
    sum += t1 * t1 + 100.0 * t2 * t2;
  }
  return sum;
}
static orc_dual
rosen_dual(const orc_dual* x, int D)
{
  orc_dual sum = dn(0.0, 0.0);
  for (int i = 0; i < D - 1; ++i) {
    orc_dual t1 = dn_sub(dn(1.0, 0.0), x[i]);
    orc_dual t2 = dn_sub(x[i + 1], dn_mul(x[i], x[i]));
    sum = dn_add(dn_add(sum, dn_mul(t1, t1)),
                 dn_mul(dn_mul(dn(100.0, 0.0), t2), t2));
  }
  return sum;
}
static double
rosen_adjoint(const double* x, double* g, int D)
{
  for (int i = 0; i < D; ++i)
    g[i] = 0.0;
  double sum = 0.0;
  for (int i = 0; i < D - 1; ++i) {
    double t1 = 1.0 - x[i];
    double t2 = x[i + 1] - x[i] * x[i];
    sum += t1 * t1 + 100.0 * t2 * t2;
    double a = 200.0 * t2;
    g[i] = (g[i] - 2.0 * t1) - (2.0 * x[i]) * a;
    g[i + 1] = g[i + 1] + a;
  }
  return sum;
}

/* --- Rastrigin: examples/rastrigin.hpp:19-27 (double), 6-17 (dual) --- */
static double
rastrigin_value(const double* x, int D)
{
  double sum = 10.0 * D;
  for (int i = 0; i < D; ++i)
    sum += x[i] * x[i] - 10.0 * cos(2.0 * ORC_PI * x[i]);
  return sum;
}
static orc_dual
rastrigin_dual(const orc_dual* x, int D)
{
  orc_dual sum = dn(10.0 * D, 0.0);
  for (int i = 0; i < D; ++i)
    sum = dn_add(sum, dn_sub(dn_mul(x[i], x[i]),
                             dn_mul(dn(10.0, 0.0),
                                    dn_cos(dn_mul(x[i], dn(2.0 * ORC_PI, 0.0))))));
  return sum;
}
static double
rastrigin_adjoint(const double* x, double* g, int D)
{
  const double k = 10.0 * (2.0 * ORC_PI);
  double sum = 10.0 * D;
  for (int i = 0; i < D; ++i) {
    double w = 2.0 * ORC_PI * x[i];
    sum += x[i] * x[i] - 10.0 * cos(w);
    g[i] = 2.0 * x[i] + k * sin(w);
  }
  return sum;
}

/* --- Ackley: examples/ackley.hpp:25-38 (double), 10-23 (dual) --- */
static double
ackley_value(const double* x, int D)
{
  double sum_sq = 0.0, sum_cos = 0.0;
  for (int i = 0; i < D; ++i) {
    sum_sq += x[i] * x[i];
    sum_cos += cos(2.0 * ORC_PI * x[i]);
  }
  double term1 = -20.0 * exp(-0.2 * sqrt(sum_sq / D));
  double term2 = -exp(sum_cos / D);
  return term1 + term2 + 20.0 + exp(1.0);
}
static orc_dual
ackley_dual(const orc_dual* x, int D)
{
  orc_dual sum_sq = dn(0.0, 0.0), sum_cos = dn(0.0, 0.0);
  for (int i = 0; i < D; ++i) {
    sum_sq = dn_add(sum_sq, dn_powd(x[i], 2.0));
    sum_cos = dn_add(sum_cos, dn_cos(dn_smul(2.0 * ORC_PI, x[i])));
  }
  orc_dual Dd = dn((double)D, 0.0);
  orc_dual term1 =
    dn_mul(dn(-20.0, 0.0), dn_exp(dn_smul(-0.2, dn_sqrt(dn_div(sum_sq, Dd)))));
  orc_dual term2 = dn_sub(dn(0.0, 0.0), dn_exp(dn_div(sum_cos, Dd)));
  return dn_add(dn_add(dn_add(term1, term2), dn(20.0, 0.0)), dn_exp(dn(1.0, 0.0)));
}
static double
ackley_adjoint(const double* x, double* g, int D)
{
  double sum_sq = 0.0, sum_cos = 0.0;
  for (int i = 0; i < D; ++i) {
    sum_sq += x[i] * x[i];
    sum_cos += cos(2.0 * ORC_PI * x[i]);
  }
  double r = sqrt(sum_sq / D);
  double e1 = exp(-0.2 * r);
  double e2 = exp(sum_cos / D);
  double term1 = -20.0 * e1;
  double term2 = -e2;
  double f = term1 + term2 + 20.0 + exp(1.0);
  /* df/dx_i = 4 e1 x_i / (D r) + (2 pi / D) e2 sin(2 pi x_i) */
  double q = (4.0 * e1) / ((double)D * r);
  double c = (2.0 * ORC_PI / D) * e2;
  for (int i = 0; i < D; ++i)
    g[i] = q * x[i] + c * sin(2.0 * ORC_PI * x[i]);
  return f;
}

/* --- Gaussian: examples/gaussian.hpp:19-33, params = C row-major --- */
static double
gauss_value(const double* x, const double* C, int D)
{
  double q = 0.0;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      q += x[i] * C[i * D + j] * x[j];
  return 0.5 * q;
}
static orc_dual
gauss_dual(const orc_dual* x, const double* C, int D)
{
  orc_dual q = dn(0.0, 0.0);
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      q = dn_add(q, dn_mul(dn_mul(x[i], dn(C[i * D + j], 0.0)), x[j]));
  return dn_mul(dn(0.5, 0.0), q);
}
static double
gauss_adjoint(const double* x, double* g, const double* C, int D)
{
  double f = gauss_value(x, C, D);
  /* df/dx_k = 0.5 (sum_j C_kj x_j + sum_i C_ik x_i) */
  for (int k = 0; k < D; ++k) {
    double r = 0.0, c = 0.0;
    for (int j = 0; j < D; ++j) {
      r = fma(C[k * D + j], x[j], r);
      c = fma(C[j * D + k], x[j], c);
    }
    g[k] = 0.5 * (r + c);
  }
  return f;
}

/* --- Poisson NLL: simulation/fitting/LogLikelihood.h:29-58.
   params = [x (n), k (n), lgamma(k+1) (n)]; the oracle recomputes the
   lgamma term itself, as the reference does. --- */
static double
loglik_value(const double* th, const double* P, int64_t np)
{
  int64_t n = np / 3;
  const double *X = P, *K = P + n;
  const double a = th[0], b = th[1], c = th[2], d = th[3];
  double nll = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double x = X[i];
    double k = K[i];
    double mean = a * pow(1.0 - x, b) / pow(x, c + d * log(x));
    double log_p = k * log(mean) - mean - lgamma(k + 1.0);
    nll -= log_p;
  }
  return nll;
}
static orc_dual
loglik_dual(const orc_dual* th, const double* P, int64_t np)
{
  int64_t n = np / 3;
  const double *X = P, *K = P + n;
  orc_dual nll = dn(0.0, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    orc_dual x = dn(X[i], 0.0);
    orc_dual k = dn(K[i], 0.0);
    orc_dual num = dn_mul(th[0], dn_pow(dn_sub(dn(1.0, 0.0), x), th[1]));
    orc_dual ex = dn_add(th[2], dn_mul(th[3], dn_log(x)));
    orc_dual mean = dn_div(num, dn_pow(x, ex));
    orc_dual log_p = dn_sub(dn_sub(dn_mul(k, dn_log(mean)), mean),
                            dn_lgamma(dn_add(k, dn(1.0, 0.0))));
    nll = dn_sub(nll, log_p);
  }
  return nll;
}
static double
loglik_adjoint(const double* th, double* g, const double* P, int64_t np)
{
  int64_t n = np / 3;
  const double *X = P, *K = P + n;
  const double a = th[0], b = th[1], c = th[2], d = th[3];
  double nll = 0.0, ga = 0.0, gb = 0.0, gc = 0.0, gd = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double x = X[i];
    double k = K[i];
    double lx = log(x);
    double pw = pow(1.0 - x, b);
    double qw = pow(x, c + d * lx);
    double mean = a * pw / qw;
    double log_p = k * log(mean) - mean - lgamma(k + 1.0);
    nll -= log_p;
    /* d nll / d mean = 1 - k / mean */
    double w = 1.0 - k / mean;
    double wm = w * mean;
    ga += w * (pw / qw);
    gb += wm * log(1.0 - x);
    gc -= wm * lx;
    gd -= wm * (lx * lx);
  }
  g[0] = ga;
  g[1] = gb;
  g[2] = gc;
  g[3] = gd;
  return nll;
}

/* --- MLP 1-16-16-1, tanh hidden, linear output, MSE over the points.
   Parameter packing per neuron (weights then bias), layer by layer, in the
   order of examples/nn.hpp:52-68.  params = [x (P), y (P)]. --- */
#define MLP_H 16
static double
mlp_value(const double* th, const double* prm, int64_t np)
{
  int64_t P = np / 2;
  const double *xs = prm, *ys = prm + P;
  double loss = 0.0;
  for (int64_t p = 0; p < P; ++p) {
    int idx = 0;
    double h1[MLP_H], h2[MLP_H];
    for (int j = 0; j < MLP_H; ++j) {
      double sum = 0.0;
      sum += th[idx++] * xs[p];
      sum += th[idx++];
      h1[j] = tanh(sum);
    }
    for (int k = 0; k < MLP_H; ++k) {
      double sum = 0.0;
      for (int j = 0; j < MLP_H; ++j)
        sum += h1[j] * th[idx++];
      sum += th[idx++];
      h2[k] = tanh(sum);
    }
    double sum = 0.0;
    for (int k = 0; k < MLP_H; ++k)
      sum += h2[k] * th[idx++];
    sum += th[idx++];
    double dd = sum - ys[p];
    loss += dd * dd;
  }
  return loss / (double)P;
}
static orc_dual
mlp_dual(const orc_dual* th, const double* prm, int64_t np)
{
  int64_t P = np / 2;
  const double *xs = prm, *ys = prm + P;
  orc_dual loss = dn(0.0, 0.0);
  for (int64_t p = 0; p < P; ++p) {
    int idx = 0;
    orc_dual h1[MLP_H], h2[MLP_H];
    orc_dual xp = dn(xs[p], 0.0);
    for (int j = 0; j < MLP_H; ++j) {
      orc_dual sum = dn(0.0, 0.0);
      sum = dn_add(sum, dn_mul(th[idx++], xp));
      sum = dn_add(sum, th[idx++]);
      h1[j] = dn_tanh(sum);
    }
    for (int k = 0; k < MLP_H; ++k) {
      orc_dual sum = dn(0.0, 0.0);
      for (int j = 0; j < MLP_H; ++j)
        sum = dn_add(sum, dn_mul(h1[j], th[idx++]));
      sum = dn_add(sum, th[idx++]);
      h2[k] = dn_tanh(sum);
    }
    orc_dual sum = dn(0.0, 0.0);
    for (int k = 0; k < MLP_H; ++k)
      sum = dn_add(sum, dn_mul(h2[k], th[idx++]));
    sum = dn_add(sum, th[idx++]);
    orc_dual dd = dn_sub(sum, dn(ys[p], 0.0));
    loss = dn_add(loss, dn_mul(dd, dd));
  }
  return dn_div(loss, dn((double)P, 0.0));
}
static double
mlp_adjoint(const double* th, double* g, const double* prm, int64_t np)
{
  int64_t P = np / 2;
  const double *xs = prm, *ys = prm + P;
  const int W2 = 2 * MLP_H, W3 = W2 + MLP_H * (MLP_H + 1);
  const int NP = W3 + MLP_H + 1;
  for (int i = 0; i < NP; ++i)
    g[i] = 0.0;
  double loss = 0.0;
  const double scale = 2.0 / (double)P;
  for (int64_t p = 0; p < P; ++p) {
    int idx = 0;
    double h1[MLP_H], h2[MLP_H];
    for (int j = 0; j < MLP_H; ++j) {
      double sum = 0.0;
      sum += th[idx++] * xs[p];
      sum += th[idx++];
      h1[j] = tanh(sum);
    }
    for (int k = 0; k < MLP_H; ++k) {
      double sum = 0.0;
      for (int j = 0; j < MLP_H; ++j)
        sum += h1[j] * th[idx++];
      sum += th[idx++];
      h2[k] = tanh(sum);
    }
    double sum = 0.0;
    for (int k = 0; k < MLP_H; ++k)
      sum += h2[k] * th[idx++];
    sum += th[idx++];
    double dd = sum - ys[p];
    loss += dd * dd;
    /* reverse sweep */
    double dy = scale * dd;
    double d2[MLP_H];
    for (int k = 0; k < MLP_H; ++k) {
      g[W3 + k] += dy * h2[k];
      d2[k] = (dy * th[W3 + k]) * (1.0 - h2[k] * h2[k]);
    }
    g[W3 + MLP_H] += dy;
    double d1[MLP_H];
    for (int j = 0; j < MLP_H; ++j)
      d1[j] = 0.0;
    for (int k = 0; k < MLP_H; ++k) {
      int base = W2 + k * (MLP_H + 1);
      for (int j = 0; j < MLP_H; ++j) {
        g[base + j] += d2[k] * h1[j];
        d1[j] += d2[k] * th[base + j];
      }
      g[base + MLP_H] += d2[k];
    }
    for (int j = 0; j < MLP_H; ++j) {
      double dj = d1[j] * (1.0 - h1[j] * h1[j]);
      g[2 * j] += dj * xs[p];
      g[2 * j + 1] += dj;
    }
  }
  return loss / (double)P;
}

/* --- sphere: sum x_i^2 (test/bfgs_tests.cu GoodObjective, Quad) --- */
static double
sphere_value(const double* x, int D)
{
  double s = 0.0;
  for (int i = 0; i < D; ++i)
    s += x[i] * x[i];
  return s;
}
static orc_dual
sphere_dual(const orc_dual* x, int D)
{
  orc_dual s = dn(0.0, 0.0);
  for (int i = 0; i < D; ++i)
    s = dn_add(s, dn_mul(x[i], x[i]));
  return s;
}
static double
sphere_adjoint(const double* x, double* g, int D)
{
  double s = 0.0;
  for (int i = 0; i < D; ++i) {
    s += x[i] * x[i];
    g[i] = 2.0 * x[i];
  }
  return s;
}

double
orc_value(const orc_objective* f, const double* x)
{
  switch (f->kind) {
    case ORC_ROSENBROCK: return rosen_value(x, f->dim);
    case ORC_RASTRIGIN: return rastrigin_value(x, f->dim);
    case ORC_ACKLEY: return ackley_value(x, f->dim);
    case ORC_GAUSSIAN: return gauss_value(x, f->params, f->dim);
    case ORC_LOGLIK: return loglik_value(x, f->params, f->n_params);
    case ORC_MLP: return mlp_value(x, f->params, f->n_params);
    case ORC_SPHERE: return sphere_value(x, f->dim);
    default: return NAN;
  }
}

orc_dual
orc_value_dual(const orc_objective* f, const orc_dual* x)
{
  switch (f->kind) {
    case ORC_ROSENBROCK: return rosen_dual(x, f->dim);
    case ORC_RASTRIGIN: return rastrigin_dual(x, f->dim);
    case ORC_ACKLEY: return ackley_dual(x, f->dim);
    case ORC_GAUSSIAN: return gauss_dual(x, f->params, f->dim);
    case ORC_LOGLIK: return loglik_dual(x, f->params, f->n_params);
    case ORC_MLP: return mlp_dual(x, f->params, f->n_params);
    case ORC_SPHERE: return sphere_dual(x, f->dim);
    default: return dn(NAN, NAN);
  }
}

/* include/duals.cuh:242-266: one forward pass per dimension, seed e_i. */
void
orc_grad_dual(const orc_objective* f, const double* x, double* g)
{
  int D = f->dim;
  orc_dual* xd = (orc_dual*)malloc(sizeof(orc_dual) * (size_t)D);
  for (int i = 0; i < D; ++i)
    xd[i] = dn(x[i], 0.0);
  for (int i = 0; i < D; ++i) {
    xd[i].du = 1.0;
    orc_dual r = orc_value_dual(f, xd);
    g[i] = r.du;
    xd[i].du = 0.0;
  }
  free(xd);
}

double
orc_value_grad_adjoint(const orc_objective* f, const double* x, double* g)
{
  switch (f->kind) {
    case ORC_ROSENBROCK: return rosen_adjoint(x, g, f->dim);
    case ORC_RASTRIGIN: return rastrigin_adjoint(x, g, f->dim);
    case ORC_ACKLEY: return ackley_adjoint(x, g, f->dim);
    case ORC_GAUSSIAN: return gauss_adjoint(x, g, f->params, f->dim);
    case ORC_LOGLIK: return loglik_adjoint(x, g, f->params, f->n_params);
    case ORC_MLP: return mlp_adjoint(x, g, f->params, f->n_params);
    case ORC_SPHERE: return sphere_adjoint(x, g, f->dim);
    default: return NAN;
  }
}

/* ======================================================================
 * BFGS building blocks: include/utils.cuh.
 * ==================================================================== */

/* utils.cuh:111-119 (serial left-to-right sum, fused) */
double
orc_dot(const double* a, const double* b, int n)
{
  double s = 0.0;
  for (int i = 0; i < n; ++i)
    s = fma(a[i], b[i], s);
  return s;
}

/* utils.cuh:137-146 */
double
orc_norm(const double* g, int n)
{
  return sqrt(orc_dot(g, g, n));
}

/* utils.cuh:148-161: p_i = -sum_j H_ij g_j, j inner, row-major */
void
orc_search_direction(double* p, const double* H, const double* g, int n)
{
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j)
      s = fma(H[i * n + j], g[j], s);
    p[i] = -s;
  }
}

/* utils.cuh:245-287: H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T,
   evaluated literally in O(D^4) into a temporary, then copied back. */
void
orc_update_literal(double* H, const double* s, const double* y, double sTy,
                   int n)
{
  if (fabs(sTy) < 1e-14)
    return;
  double rho = 1.0 / sTy;
  double* Ht = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      double sum = 0.0;
      for (int k = 0; k < n; ++k) {
        double A_ik = ((i == k) ? 1.0 : 0.0) - rho * s[i] * y[k];
        double inner = 0.0;
        for (int m = 0; m < n; ++m) {
          double B_mj = ((m == j) ? 1.0 : 0.0) - rho * y[m] * s[j];
          inner = fma(H[k * n + m], B_mj, inner);
        }
        sum = fma(A_ik, inner, sum);
      }
      Ht[i * n + j] = sum + rho * s[i] * s[j];
    }
  }
  memcpy(H, Ht, sizeof(double) * (size_t)n * (size_t)n);
  free(Ht);
}

/* The same update in O(D^2) for symmetric H (H_0 = I and this update keep
   H exactly symmetric):  u = H y, beta = rho^2 y.u + rho,
   a = (beta/2) s - rho u,  H <- H + s a^T + a s^T.
   s_i a_j + a_i s_j is formed with plain products and one sum, which is
   symmetric under i <-> j bit for bit. */
void
orc_update_rank2(double* H, const double* s, const double* y, double sTy,
                 int n)
{
  if (fabs(sTy) < 1e-14)
    return;
  double rho = 1.0 / sTy;
  double* u = (double*)calloc((size_t)n * 2 + 1, sizeof(double));
  if (!u)
    return;
  double* a = u + n;
  for (int i = 0; i < n; ++i) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j)
      acc = fma(H[i * n + j], y[j], acc);
    u[i] = acc;
  }
  double yu = orc_dot(y, u, n);
  double beta = fma(rho * rho, yu, rho);
  double hb = 0.5 * beta;
  for (int i = 0; i < n; ++i)
    a[i] = fma(hb, s[i], -(rho * u[i]));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      H[i * n + j] = H[i * n + j] + (s[i] * a[j] + a[i] * s[j]);
  free(u);
}

/* utils.cuh:376-426: alpha = 1; ddir = g.p; up to 20 trials; accept iff
   f(x + alpha p) <= f0 + 0.1 alpha ddir and f(x + alpha p) < f0; else halve.
   Never returns 0 (the smallest value is 2^-20). */
double
orc_line_search(const orc_objective* f, double f0, const double* x,
                const double* p, const double* g, int n, int* n_evals,
                double* f_acc)
{
  double alpha = 1.0;
  double ddir = orc_dot(g, p, n);
  double* xt = (double*)malloc(sizeof(double) * (size_t)n);
  int ne = 0;
  *f_acc = NAN;
  for (int it = 0; it < 20; ++it) {
    for (int j = 0; j < n; ++j)
      xt[j] = fma(alpha, p[j], x[j]);
    double f1 = orc_value(f, xt);
    ++ne;
    if (f1 <= fma(0.1 * alpha, ddir, f0) && f1 < f0) {
      *f_acc = f1;
      break;
    }
    alpha *= 0.5;
  }
  free(xt);
  *n_evals = ne;
  return alpha;
}

int64_t
orc_argmin(const double* v, int64_t n)
{
  if (n <= 0)
    return -1;
  int64_t best = -1;
  for (int64_t i = 0; i < n; ++i) {
    if (isnan(v[i]))
      continue;
    if (best < 0 || v[i] < v[best])
      best = i;
  }
  return best < 0 ? 0 : best;
}

/* ======================================================================
 * PSO: include/pso.cuh:17-155 (kernels), 194-345 (launch).  Synchronous:
 * gBest is frozen during an iteration and replaced after it only on a
 * strict improvement, ties going to the lowest particle index (the
 * reference's racy CAS-then-copy is designed out).  No clamping, as in the
 * reference.  w = 0.5, c1 = 1.2, c2 = 1.5 (pso.cuh:302).
 * ==================================================================== */
void
orc_pso_step(int n, int dim, double w, double c1, double c2, const double* r,
             double* X, double* V, const double* pbx, const double* gbx)
{
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < dim; ++d) {
      size_t o = (size_t)i * dim + d;
      double r1 = r[2 * o], r2 = r[2 * o + 1];
      double xi = X[o], vi = V[o], pb = pbx[o], gb = gbx[d];
      double nv = w * vi + c1 * r1 * (pb - xi) + c2 * r2 * (gb - xi);
      V[o] = nv;
      X[o] = xi + nv;
    }
}

static void
pso_gbest(const double* pbv, const double* pbx, int64_t n, int D, double* gbv,
          double* gbx)
{
  int64_t m = orc_argmin(pbv, n);
  if (m >= 0 && pbv[m] < *gbv) {
    *gbv = pbv[m];
    memcpy(gbx, pbx + (size_t)m * D, sizeof(double) * (size_t)D);
  }
}

/* The swarm after PSO_ITER iterations: pBest (pbx, pbv), gBest (gbx, gbv)
   and, when xo / vo are not NULL, the final positions and velocities.
   Init follows pso.cuh:34-69 (per coordinate: position U(lo, hi], then
   velocity U(-0.1R, 0.1R], R = hi - lo), each iteration pso.cuh:97-155 (per
   coordinate r1 then r2, both U(0, 1]; the update is orc_pso_step). */
static int
pso_run(const orc_problem* P, int nt, double* pbx, double* pbv, double* gbx,
        double* gbv, double* xo, double* vo)
{
  const int64_t N = P->n_starts;
  const int D = P->dim;
  const double lo = P->lower, hi = P->upper;
  const double vr = (hi - lo) * 0.1;
  const double w = 0.5, c1 = 1.2, c2 = 1.5; /* pso.cuh:302 */
  orc_objective f = {P->objective, D, P->params, P->n_params};
  const size_t ND = (size_t)(N > 0 ? N : 1) * (size_t)D;
  double* X = (double*)malloc(sizeof(double) * ND);
  double* V = (double*)malloc(sizeof(double) * ND);
  double* R = (double*)malloc(sizeof(double) * 2 * ND);
  uint32_t* xs = (uint32_t*)malloc(sizeof(uint32_t) * 6 * ND);
  if (!X || !V || !R || !xs) {
    free(X);
    free(V);
    free(R);
    free(xs);
    return -2;
  }
  *gbv = INFINITY;
  for (int d = 0; d < D; ++d)
    gbx[d] = 0.0;

#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t i = 0; i < N; ++i) {
    for (int d = 0; d < D; ++d) {
      size_t o = (size_t)i * D + d;
      uint32_t* st = xs + 6 * o;
      orc_xorwow_seed(P->seed, P->stream, (uint64_t)i, (uint32_t)d, st);
      double u0, u1;
      pso_pair(P, (uint64_t)i, 2u * (uint64_t)d, st, &u0, &u1);
      double rx = to_range(u0, lo, hi);
      double rv = to_range(u1, -vr, vr);
      X[o] = rx;
      V[o] = rv;
      pbx[o] = rx;
    }
    pbv[i] = orc_value(&f, X + (size_t)i * D);
  }
  pso_gbest(pbv, pbx, N, D, gbv, gbx);

  for (int it = 1; it < P->pso_iter + 1; ++it) {
    const uint64_t base = 2u * (uint64_t)D * (uint64_t)it;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < N; ++i) {
      double* r = R + 2 * (size_t)i * D;
      for (int d = 0; d < D; ++d) {
        double u0, u1;
        pso_pair(P, (uint64_t)i, base + 2u * (uint64_t)d, xs + 6 * ((size_t)i * D + d), &u0,
                 &u1);
        r[2 * d] = to_range(u0, 0.0, 1.0);
        r[2 * d + 1] = to_range(u1, 0.0, 1.0);
      }
      double* Xi = X + (size_t)i * D;
      orc_pso_step(1, D, w, c1, c2, r, Xi, V + (size_t)i * D, pbx + (size_t)i * D, gbx);
      double fv = orc_value(&f, Xi);
      if (fv < pbv[i]) {
        pbv[i] = fv;
        memcpy(pbx + (size_t)i * D, Xi, sizeof(double) * D);
      }
    }
    pso_gbest(pbv, pbx, N, D, gbv, gbx);
  }
  if (xo)
    memcpy(xo, X, sizeof(double) * (size_t)N * D);
  if (vo)
    memcpy(vo, V, sizeof(double) * (size_t)N * D);
  free(R);
  free(xs);
  free(X);
  free(V);
  return 0;
}

int
orc_pso(const orc_problem* P, double* pbest_x, double* pbest_val,
        double* gbest_x, double* gbest_val)
{
  if (!P || P->n_starts < 0 || P->dim < 1)
    return -1;
  return pso_run(P, 0, pbest_x, pbest_val, gbest_x, gbest_val, NULL, NULL);
}

int
orc_pso_swarm(const orc_problem* P, double* x, double* v, double* pbest_x,
              double* pbest_val, double* gbest_x, double* gbest_val)
{
  if (!P || P->n_starts < 0 || P->dim < 1)
    return -1;
  return pso_run(P, 0, pbest_x, pbest_val, gbest_x, gbest_val, x, v);
}

/* ======================================================================
 * Per-start BFGS: include/bfgs_sequential.cuh:46-221 (step order),
 * termination rules include/bfgs_common.cuh:108-275.
 * ==================================================================== */
typedef struct {
  atomic_int stop_flag;
  atomic_int converged;
} orc_ctx;

typedef struct {
  int32_t status, iter;
  double fval, gnorm;
  int64_t evals, grads;
} start_out;

static void
grad_eval(const orc_objective* f, int mode, const double* x, double* g)
{
  if (mode == ORC_GRAD_DUAL)
    orc_grad_dual(f, x, g);
  else
    (void)orc_value_grad_adjoint(f, x, g);
}

/* Slot of start s at iteration it in a trajectory, or -1 if not kept. */
static int64_t
traj_slot(const orc_traj* T, int64_t s, int it)
{
  if (!T || !T->fval || s >= T->starts || it < 0 || it >= T->steps)
    return -1;
  return (int64_t)it * T->starts + s;
}

static void
run_start(const orc_problem* P, const orc_options* O, int64_t s,
          const double* starts, orc_ctx* ctx, start_out* out, double* xo)
{
  const int D = P->dim;
  orc_objective f = {P->objective, D, P->params, P->n_params};
  double* buf = (double*)malloc(sizeof(double) * ((size_t)7 * D + (size_t)D * D));
  double *x = buf, *xn = x + D, *g = xn + D, *gn = g + D, *p = gn + D;
  double *dx = p + D, *dg = dx + D, *H = dg + D;
  int64_t evals = 0, grads = 0;

  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      H[i * D + j] = (i == j) ? 1.0 : 0.0;
  for (int d = 0; d < D; ++d) {
    x[d] = starts ? starts[(size_t)s * D + d] : start_uniform_in(P, (uint64_t)s, d);
    g[d] = 0.0;
  }
  double f0 = orc_value(&f, x);
  ++evals;
  double bestVal = f0;
  grad_eval(&f, O->grad_mode, x, g);
  ++grads;

  int32_t status = -1;
  double fval = 333777.0, gnorm = 69.0;
  int iter;
  int done = 0;
  const orc_traj* T = O->traj;
  for (iter = 0; iter < P->max_iter; ++iter) {
    const int64_t to = traj_slot(T, s, iter);
    if (to >= 0) { /* bfgs_sequential.cuh:100-109 */
      for (int d = 0; d < D; ++d)
        T->coords[((int64_t)iter * D + d) * T->starts + s] = x[d];
      T->fval[to] = bestVal;
      T->grad[to] = orc_norm(g, D);
      T->status[to] = -1;
    }
    if (O->early_stop && atomic_load(&ctx->stop_flag) != 0) {
      status = 2;
      fval = orc_value(&f, x);
      gnorm = orc_norm(g, D);
      done = 1;
      break;
    }
    orc_search_direction(p, H, g, D);
    int ne = 0;
    double facc;
    double alpha = orc_line_search(&f, bestVal, x, p, g, D, &ne, &facc);
    evals += ne;
    if (to >= 0 && T->alpha) /* bfgs_parallel.cuh:182-184 */
      T->alpha[to] = alpha;
    for (int i = 0; i < D; ++i) {
      xn[i] = fma(alpha, p[i], x[i]);
      dx[i] = xn[i] - x[i];
    }
    double fnew = orc_value(&f, xn); /* reference recomputes f(x_new) */
    if (isnan(facc))
      ++evals; /* only the exhausted search needs a fresh evaluation */
    grad_eval(&f, O->grad_mode, xn, gn);
    ++grads;
    for (int i = 0; i < D; ++i)
      dg[i] = gn[i] - g[i];
    double sTy = orc_dot(dx, dg, D);
    if (O->update_mode == ORC_UPDATE_LITERAL)
      orc_update_literal(H, dx, dg, sTy, D);
    else
      orc_update_rank2(H, dx, dg, sTy, D);
    if (fnew < bestVal) {
      bestVal = fnew;
      memcpy(x, xn, sizeof(double) * D);
      memcpy(g, gn, sizeof(double) * D);
    }
    double gn2 = orc_norm(g, D);
    if (!isfinite(gn2) || !isfinite(fnew)) {
      status = 5;
      fval = fnew;
      gnorm = gn2;
      done = 1;
      break;
    }
    if (gn2 < P->tol) {
      int old = atomic_fetch_add(&ctx->converged, 1);
      status = 1;
      fval = orc_value(&f, x);
      gnorm = gn2;
      if (old + 1 == P->required)
        atomic_store(&ctx->stop_flag, 1);
      done = 1;
      break;
    }
  }
  if (!done && iter == P->max_iter) {
    status = 0;
    fval = orc_value(&f, x);
    gnorm = orc_norm(g, D);
  }
  { /* final status at the deciding iteration (bfgs_common.cuh:108-238) */
    const int64_t to = traj_slot(T, s, status == 0 ? P->max_iter - 1 : iter);
    if (to >= 0)
      T->status[to] = (int8_t)status;
  }
  out->status = status;
  out->iter = iter;
  out->fval = fval;
  out->gnorm = gnorm;
  out->evals = evals;
  out->grads = grads;
  if (xo)
    memcpy(xo, x, sizeof(double) * D);
  free(buf);
}

int
orc_run(const orc_problem* P, const orc_options* O, int32_t* status,
        int32_t* iter, double* fval, double* gnorm, double* x_out,
        int64_t* evals_out, int64_t* grads_out, double* best_x,
        orc_best* best)
{
  if (!P || !O || !best || P->dim < 1 || P->n_starts < 0)
    return -1;
  const int64_t N = P->n_starts;
  const int D = P->dim;
  best->idx = -1;
  best->status = -1;
  best->fval = 37.0;
  best->gnorm = 0.0;
  best->iter = 0;
  best->claimed = best->stopped = best->surrendered = 0;
  if (best_x)
    for (int d = 0; d < D; ++d)
      best_x[d] = 0.0;
  if (N == 0)
    return 0; /* zeus.cuh:58-60 */

  int nt = O->n_threads;
#ifdef _OPENMP
  if (nt <= 0)
    nt = omp_get_max_threads();
#else
  nt = 1;
#endif

  double* pso_x = NULL;
  const double* starts = P->x0;
  if (!starts && P->pso_iter >= 0) {
    pso_x = (double*)malloc(sizeof(double) * (size_t)N * D);
    double* pbv = (double*)malloc(sizeof(double) * (size_t)N);
    double* gbx = (double*)malloc(sizeof(double) * (size_t)D);
    double gbv;
    int rc = pso_run(P, nt, pso_x, pbv, gbx, &gbv, NULL, NULL);
    free(pbv);
    free(gbx);
    if (rc != 0) {
      free(pso_x);
      return rc;
    }
    starts = pso_x;
  }

  if (O->traj && O->traj->fval) { /* zeus.cuh:77-103 */
    const orc_traj* T = O->traj;
    const int64_t n = (int64_t)T->steps * T->starts;
    for (int64_t i = 0; i < n * D; ++i)
      T->coords[i] = NAN;
    for (int64_t i = 0; i < n; ++i) {
      T->fval[i] = NAN;
      T->grad[i] = NAN;
      T->status[i] = -1;
      if (T->alpha)
        T->alpha[i] = 0.0;
    }
  }
  start_out* outs = (start_out*)malloc(sizeof(start_out) * (size_t)N);
  double* xs = (double*)malloc(sizeof(double) * (size_t)N * D);
  orc_ctx ctx;
  atomic_init(&ctx.stop_flag, 0);
  atomic_init(&ctx.converged, 0);

#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t s = 0; s < N; ++s)
    run_start(P, O, s, starts, &ctx, &outs[s], xs + (size_t)s * D);

  double* fv = (double*)malloc(sizeof(double) * (size_t)N);
  for (int64_t s = 0; s < N; ++s) {
    fv[s] = outs[s].fval;
    if (status) status[s] = outs[s].status;
    if (iter) iter[s] = outs[s].iter;
    if (fval) fval[s] = outs[s].fval;
    if (gnorm) gnorm[s] = outs[s].gnorm;
    if (evals_out) evals_out[s] = outs[s].evals;
    if (grads_out) grads_out[s] = outs[s].grads;
    if (outs[s].status == 1)
      best->claimed++;
    else if (outs[s].status == 2)
      best->stopped++;
    else
      best->surrendered++;
  }
  if (x_out)
    memcpy(x_out, xs, sizeof(double) * (size_t)N * D);
  int64_t b = orc_argmin(fv, N);
  best->idx = (int32_t)b;
  best->status = outs[b].status;
  best->iter = outs[b].iter;
  best->fval = outs[b].fval;
  best->gnorm = outs[b].gnorm;
  if (best_x)
    memcpy(best_x, xs + (size_t)b * D, sizeof(double) * D);
  free(fv);
  free(xs);
  free(outs);
  free(pso_x);
  return 0;
}
