// frontend_test.hip -- exercises the C++ front end (zeus/zeus.hpp) the way
// reference users write objectives: templated functors over std::array<T, D>,
// a Matrix member uploaded once, an optional hand-written adjoint.
//
//   frontend_test --host-only   checks that need no GPU (exit status)
//   frontend_test               GPU runs; one JSON line per case on stdout,
//                               compared against the CPU oracle by pytest.
#include "zeus/zeus.hpp"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

// --- reference-style objectives ----------------------------------------------
template <std::size_t D>
struct UserRosenbrock { // shape of examples/rosenbrock.hpp
  static constexpr std::size_t arity = D;
  template <class T>
  __host__ __device__ T operator()(const std::array<T, D>& x) const
  {
    T sum = T(0.0);
    for (int i = 0; i < int(D) - 1; ++i) {
      const T t1 = T(1.0) - x[i];
      const T t2 = x[i + 1] - x[i] * x[i];
      if constexpr (std::is_same_v<T, double>)
        sum += t1 * t1 + 100.0 * t2 * t2;
      else
        sum = sum + t1 * t1 + T(100.0) * t2 * t2;
    }
    return sum;
  }
};

template <std::size_t N>
class UserGaussian { // shape of examples/gaussian.hpp (Matrix member)
  zeus::Matrix<double> C;

public:
  static constexpr std::size_t arity = N;
  explicit UserGaussian(const std::vector<double>& c) : C(N, N) { C.set(c.data(), N * N); }
  template <class T>
  __host__ __device__ T operator()(const std::array<T, N>& x) const
  {
    const double* Cp = C.data();
    T q = T(0);
    for (std::size_t i = 0; i < N; ++i)
      for (std::size_t j = 0; j < N; ++j)
        q += x[i] * T(Cp[i * N + j]) * x[j];
    return T(0.5) * q;
  }
};

struct UserSphereAdjoint { // with a hand-written adjoint
  static constexpr std::size_t arity = 3;
  template <class T>
  __host__ __device__ T operator()(const std::array<T, 3>& x) const
  {
    T s = T(0.0);
    for (int i = 0; i < 3; ++i)
      s += x[i] * x[i];
    return s;
  }
  __host__ __device__ double value_and_gradient(const double* x, double* g) const
  {
    double s = 0.0;
    for (int i = 0; i < 3; ++i) {
      s += x[i] * x[i];
      g[i] = 2.0 * x[i];
    }
    return s;
  }
};

// --- objectives the concept must reject (test/bad_objectives_concepts.cu) ----
struct NotTemplated {
  static constexpr std::size_t arity = 2;
  double operator()(const std::array<double, 2>& x) const { return x[0] * x[0] + x[1] * x[1]; }
};
struct WrongArgType {
  static constexpr std::size_t arity = 2;
  template <typename T>
  T operator()(T x, T y) const
  {
    return x * x + y * y;
  }
};
static_assert(!zeus::ZeusObjective<NotTemplated, 2>, "non-templated functor must be rejected");
static_assert(!zeus::ZeusObjective<WrongArgType, 2>, "wrong argument type must be rejected");
static_assert(zeus::ZeusObjective<UserRosenbrock<2>, 2>);
static_assert(zeus::FnTraits<UserGaussian<4>>::arity == 4);
static_assert(zeus::HasAdjoint<UserSphereAdjoint> && !zeus::HasAdjoint<UserRosenbrock<2>>);

// DeviceMatrix contract (test/device_matrix_tests.cu): owned device-heap
// storage that cannot be copied, assigned or moved.
static_assert(!std::is_copy_constructible_v<DeviceMatrix<int>>);
static_assert(!std::is_copy_assignable_v<DeviceMatrix<int>>);
static_assert(!std::is_move_constructible_v<DeviceMatrix<double>>);
static_assert(std::is_destructible_v<DeviceMatrix<int>>);

static int failures = 0;
#define CHECK(cond)                                                                    \
  do {                                                                                 \
    if (!(cond)) {                                                                     \
      std::fprintf(stderr, "CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
      ++failures;                                                                      \
    }                                                                                  \
  } while (0)

template <std::size_t D>
void
print_result(const char* name, const zeus::Result<D>& r)
{
  std::printf("{\"case\": \"%s\", \"idx\": %d, \"status\": %d, \"iter\": %d, \"fval\": %.17g, "
              "\"gnorm\": %.17g, \"claimed\": %d, \"actual\": %d, \"stopped\": %d, "
              "\"surrendered\": %d, \"x\": [",
              name, r.idx, r.status, r.iter, r.fval, r.gradientNorm, r.c.claimed, r.c.actual,
              r.c.stopped, r.c.surrendered);
  for (std::size_t d = 0; d < D; ++d)
    std::printf("%s%.17g", d ? ", " : "", r.coordinates[d]);
  std::printf("]}\n");
  std::fflush(stdout);
}

int
host_only()
{
  // Matrix: zero dimensions throw before touching the device (matrix_tests.cu)
  bool threw = false;
  try {
    zeus::Matrix<double> m(0, 3);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  CHECK(threw);
  // NonNull: construction from a null pointer throws on the host (non_null_tests.cu)
  threw = false;
  try {
    double* p = nullptr;
    util::NonNull<double*> nn(p);
    (void)nn;
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  CHECK(threw);
  double v = 1.5;
  util::NonNull<double*> ok(&v);
  CHECK(*ok == 1.5 && static_cast<double*>(ok) == &v);
  // HipBuffer of size 0 owns nothing (cuda_buffer_tests.cu)
  zeus::DoubleBuffer empty(0);
  CHECK(empty.data() == nullptr && empty.size() == 0);
  std::vector<double> out;
  CHECK(empty.copyToHost(out) == 0 && out.empty());
  double one;
  CHECK(empty.copyToHost(&one, 1) == 1);
  zeus::DoubleBuffer moved(std::move(empty));
  CHECK(moved.size() == 0);
  // Result defaults (util.hpp:36-54)
  zeus::Result<3> r;
  CHECK(r.idx == -1 && r.status == -1 && r.fval == 37.0 && r.iter == 0);
  // N == 0 returns the default Result without a device
  auto r0 = zeus::Zeus(UserRosenbrock<2>{}, -5.0, 5.0, 0, 10, 5, 1, "rosenbrock", 1e-8, 42, 0);
  CHECK(r0.status == -1 && r0.fval == 37.0);
  std::printf("host-only checks: %d failure(s)\n", failures);
  return failures == 0 ? 0 : 1;
}

// One thread builds a rows x cols DeviceMatrix on the device heap, fills
// element (i, j) with base + (i cols + j), copies it out row-major and
// reports its shape; the matrix is freed when it leaves scope.
template <class T>
__global__ void
devmat_fill_kernel(T* out, int rows, int cols, T base, int* shape)
{
  if (threadIdx.x != 0 || blockIdx.x != 0)
    return;
  DeviceMatrix<T> M(rows, cols);
  if (!M.valid()) {
    shape[0] = shape[1] = -1;
    return;
  }
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j)
      M(i, j) = static_cast<T>(base + static_cast<T>(i * cols + j));
  M.set(0, 0, M(0, 0)); // set() writes through the same storage
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j)
      out[i * cols + j] = M(i, j);
  shape[0] = M.rows();
  shape[1] = M.cols();
}

template <class T>
bool
devmat_fill_and_check(int rows, int cols, T base)
{
  const int n = rows * cols;
  void *d_out = nullptr, *d_shape = nullptr;
  if (zeus_alloc(&d_out, sizeof(T) * n) != ZEUS_OK || zeus_alloc(&d_shape, 2 * sizeof(int)) != ZEUS_OK)
    return false;
  devmat_fill_kernel<T><<<1, 1>>>(static_cast<T*>(d_out), rows, cols, base,
                                  static_cast<int*>(d_shape));
  bool ok = hipGetLastError() == hipSuccess && zeus_synchronize() == ZEUS_OK;
  std::vector<T> h(n);
  int shape[2] = {0, 0};
  ok = ok && zeus_memcpy_d2h(h.data(), d_out, sizeof(T) * n) == ZEUS_OK &&
       zeus_memcpy_d2h(shape, d_shape, sizeof(shape)) == ZEUS_OK;
  (void)zeus_free(d_out);
  (void)zeus_free(d_shape);
  ok = ok && shape[0] == rows && shape[1] == cols;
  for (int i = 0; ok && i < n; ++i)
    ok = h[i] == static_cast<T>(base + static_cast<T>(i));
  return ok;
}

int
gpu_cases()
{
  // DeviceMatrix: fill and read back (test/device_matrix_tests.cu), int and
  // double, small and rectangular, repeated construction and destruction
  CHECK(devmat_fill_and_check<int>(3, 4, 10));
  CHECK(devmat_fill_and_check<double>(5, 2, 1.5));
  for (int k = 0; k < 5; ++k)
    CHECK(devmat_fill_and_check<int32_t>(2, 3, k * 100));
  std::printf("{\"case\": \"device_matrix\", \"ok\": %d}\n", failures == 0);

  // buffers: deep copy, move, copyToHost (cuda_buffer_tests.cu)
  {
    std::array<double, 4> h{1, 2, 3, 4};
    zeus::DoubleBuffer a(h);
    zeus::DoubleBuffer b(a);
    CHECK(b.data() != a.data() && b.size() == 4 && !(a == b));
    CHECK(b.copyToHost() == std::vector<double>(h.begin(), h.end()));
    zeus::DoubleBuffer c(std::move(a));
    CHECK(a.data() == nullptr && a.size() == 0 && c.size() == 4);
    std::vector<double> out;
    CHECK(c.copyToHost(out) == 0 && out[3] == 4.0);
  }
  // Matrix: set, host access, deep copy (matrix_tests.cu)
  {
    zeus::Matrix<double> m(2, 2);
    m.set(std::array<double, 4>{1, 2, 3, 4});
    m.set(1, 0, 7.0);
    CHECK(m(1, 0) == 7.0 && m.data() == m.hostData());
    zeus::Matrix<double> m2(m);
    CHECK(m2.deviceData() != m.deviceData() && m2(1, 0) == 7.0);
    double dev[4];
    CHECK(zeus_memcpy_d2h(dev, m2.deviceData(), sizeof(dev)) == ZEUS_OK);
    CHECK(dev[0] == 1 && dev[1] == 2 && dev[2] == 7 && dev[3] == 4);
  }
  // C1 workload in the per-thread form, dual gradient (the reference's own)
  {
    auto r = zeus::Zeus(UserRosenbrock<2>{}, -5.0, 5.0, 1024, 10000, 10, 1025, "rosenbrock",
                        1e-8, 42, 0, /*parallel=*/false);
    print_result("rosenbrock2_thread_dual", r);
  }
  {
    auto r = zeus::Zeus(UserRosenbrock<2>{}, -5.0, 5.0, 256, 10000, 10, 257, "rosenbrock",
                        1e-8, 42, 0, /*parallel=*/true);
    print_result("rosenbrock2_group_dual", r);
  }
  // 20-D Gaussian with a Matrix member (the reference's README example shape)
  {
    constexpr std::size_t D = 20;
    std::vector<double> c(D * D, 0.5);
    for (std::size_t i = 0; i < D; ++i)
      c[i * D + i] = 1.0 + (D - 1) * 0.5;
    UserGaussian<D> g(c);
    auto r = zeus::Zeus(g, -5.0, 5.0, 64, 10000, 10, 65, "gaussian", 1e-8, 42, 0);
    print_result("gaussian20_group_dual", r);
  }
  // built-in objectives (zeus/builtins.hpp): the library's kernels, autotuned
  {
    constexpr std::size_t D = 150;
    auto C = std::make_unique<std::array<std::array<double, D>, D>>();
    for (std::size_t i = 0; i < D; ++i)
      for (std::size_t j = 0; j < D; ++j)
        (*C)[i][j] = i == j ? 1.0 + (D - 1) * 0.5 : 0.5;
    const zeus::Gaussian<D> g(*C);
    auto r = zeus::Zeus(g, -5.0, 5.0, 64, 10000, 10, 65, "gaussian", 1e-8, 42, 0);
    print_result("gaussian150_builtin", r);
    // the headline shape: 4096 starts, 100 required, Philox draws
    auto h = zeus::Zeus(g, -5.0, 5.0, 4096, 10000, 10, 100, "gaussian", 1e-8, 42, 0, true,
                        zeus::PRNGType::PHILOX);
    print_result("gaussian150_builtin_headline", h);
    int form = 0, threads = 0; // the shape the library ran the headline call with
    CHECK(zeus_tuned_config(ZEUS_OBJ_GAUSSIAN, int(D), ZEUS_GRAD_ADJOINT, &form, &threads) ==
          ZEUS_OK);
    std::printf("{\"case\": \"gaussian150_builtin_form\", \"form\": %d, \"threads\": %d}\n",
                form, threads);
  }
  {
    auto r = zeus::Zeus(zeus::obj::Rosenbrock<10>{}, -5.0, 5.0, 256, 10000, 10, 257,
                        "rosenbrock", 1e-8, 42, 0);
    print_result("rosenbrock10_builtin", r);
  }
  // adjoint picked up from value_and_gradient
  {
    auto r = zeus::Zeus(UserSphereAdjoint{}, -5.0, 5.0, 128, 1000, 5, 129, "sphere", 1e-10, 7, 0,
                        false);
    print_result("sphere3_thread_adjoint", r);
  }
  // early stop: required 10 of 512
  {
    auto r = zeus::Zeus(UserRosenbrock<2>{}, -5.0, 5.0, 512, 10000, 10, 10, "rosenbrock", 1e-8,
                        42, 0, false);
    print_result("rosenbrock2_thread_early", r);
  }
  // trajectory file (zeus.cuh:189-239), per-thread form, bounded to 50 steps
  if (const char* path = std::getenv("ZEUS_FRONTEND_TRAJ")) {
    auto r = zeus::Zeus(UserRosenbrock<2>{}, -5.0, 5.0, 32, 300, 3, 33, "rosenbrock", 1e-8, 42,
                        0, false, zeus::PRNGType::XORWOW, path, 0, zeus::Gradient::Dual, 50);
    print_result("rosenbrock2_thread_traj", r);
  }
  std::printf("{\"case\": \"checks\", \"failures\": %d}\n", failures);
  return failures == 0 ? 0 : 1;
}

int
main(int argc, char** argv)
{
  if (argc > 1 && std::strcmp(argv[1], "--host-only") == 0)
    return host_only();
  return gpu_cases();
}
