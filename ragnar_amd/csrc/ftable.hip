// F(x) = x * int_x^inf K_{5/3}(s) ds table, built with FP64 GPU integrals.
//
// REFERENCE mode reproduces sync::TabulateFfunc / Ffunc_integrand
// (synchrotron.cpp:28-65): the float Logspace node grid (snippets.cpp:39-62),
// y = 0 below the float node 1e-6 test and above x = 20, the small-x
// asymptote below 1e-5, and otherwise x * (a 100-point float trapezoid of
// K_{5/3} on Logspace(x, 20, 100)).  The 15.7k K_{5/3} values are evaluated in
// FP64 on the GPU; the float grids and the float running sum stay on the host
// because they are the reference's own float arithmetic.
//
// ACCURATE mode tabulates the true F on the same nodes.
//
// Both integrals use the representations
//   K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt
//   F(x)    = x int_0^inf exp(-x cosh t) cosh(5t/3) / cosh t dt
// with the trapezoid rule in t (step 1/64).  The integrands are even and
// analytic in the strip |Im t| < pi/2 and decay double-exponentially, so the
// rule converges geometrically in 1/h; at h = 1/64 the discretisation error
// is far below FP64 rounding.
#include <cmath>
#include <vector>

#include "rgn_core.hpp"

struct rgn_ftable {
  int n = 0;
  int mode = RGN_FTABLE_REFERENCE;
  float xmin = 0.f, xmax = 0.f;  // MinMax of the nodes (tabulation.cpp:84-102)
  std::vector<float> x, y;
};

namespace rgn {

// Shared with the emission plans (declared in emission.hip).
const std::vector<float>& ftable_y(const rgn_ftable* t) { return t->y; }
const std::vector<float>& ftable_x(const rgn_ftable* t) { return t->x; }
float ftable_xmin(const rgn_ftable* t) { return t->xmin; }
float ftable_xmax(const rgn_ftable* t) { return t->xmax; }

// snippets.cpp:39-62: float exponent, pow(10, float) evaluated in double.
void logspace_host(float start, float stop, size_t num, float* out) {
  for (size_t i = 0; i < num; ++i) {
    if (num == 1) {
      out[i] = start;
    } else {
      const float e = std::log10(start) +
                      static_cast<float>(i) * (std::log10(stop / start)) /
                        static_cast<float>(num - 1);
      out[i] = static_cast<float>(std::pow(10.0, static_cast<double>(e)));
    }
  }
}

namespace {

__device__ double cosh_integral(double x, double nu, bool divide_cosh) {
  const double h = 1.0 / 64.0;
  double sum = 0.5 * exp(-x);
  for (int k = 1; k < 8192; ++k) {
    const double t = k * h;
    const double ch = cosh(t);
    double term = exp(-x * ch) * cosh(nu * t);
    if (divide_cosh) term /= ch;
    sum += term;
    if (x * ch > 750.0 || term < 1e-300 || (k > 8 && term < 1e-18 * sum && x * ch > 40.0)) {
      break;
    }
  }
  return sum * h;
}

__global__ void __launch_bounds__(256) bessel_k_kernel(double nu, const double* __restrict__ x,
                                                       double* __restrict__ out, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = cosh_integral(x[i], nu, false);
}

// F(x), K_{5/3}(x) and dF/dx at n points.
__global__ void __launch_bounds__(256) ffunc_kernel(const double* __restrict__ x,
                                                    double* __restrict__ F,
                                                    double* __restrict__ k53,
                                                    double* __restrict__ dfdx, size_t n) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const double xi = x[i];
  const double f = xi * cosh_integral(xi, 5.0 / 3.0, true);
  const double k = cosh_integral(xi, 5.0 / 3.0, false);
  F[i] = f;
  k53[i] = k;
  dfdx[i] = f / xi - xi * k;
}

constexpr double kOneThird = 0.3333333333333333;   // synchrotron.cpp:21
constexpr double kFiveThird = 1.6666666666666667;  // synchrotron.cpp:22

void run_bessel(double nu, const std::vector<double>& xs, std::vector<double>& out) {
  out.assign(xs.size(), 0.0);
  if (xs.empty()) return;
  DeviceBuffer dx(xs.size() * sizeof(double)), dy(xs.size() * sizeof(double));
  RGN_HIP(hipMemcpy(dx.ptr, xs.data(), dx.bytes, hipMemcpyHostToDevice));
  const unsigned grid = static_cast<unsigned>(ceil_div(xs.size(), 256));
  bessel_k_kernel<<<grid, 256>>>(nu, dx.as<double>(), dy.as<double>(), xs.size());
  RGN_HIP(hipGetLastError());
  RGN_HIP(hipMemcpy(out.data(), dy.ptr, dy.bytes, hipMemcpyDeviceToHost));
}

void run_ffunc(const std::vector<double>& xs, std::vector<double>& F, std::vector<double>& k53,
               std::vector<double>& dfdx) {
  const size_t n = xs.size();
  F.assign(n, 0.0);
  k53.assign(n, 0.0);
  dfdx.assign(n, 0.0);
  if (!n) return;
  DeviceBuffer dx(n * sizeof(double)), dF(n * sizeof(double)), dk(n * sizeof(double)),
    dd(n * sizeof(double));
  RGN_HIP(hipMemcpy(dx.ptr, xs.data(), dx.bytes, hipMemcpyHostToDevice));
  ffunc_kernel<<<static_cast<unsigned>(ceil_div(n, 256)), 256>>>(
    dx.as<double>(), dF.as<double>(), dk.as<double>(), dd.as<double>(), n);
  RGN_HIP(hipGetLastError());
  RGN_HIP(hipMemcpy(F.data(), dF.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
  RGN_HIP(hipMemcpy(k53.data(), dk.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
  RGN_HIP(hipMemcpy(dfdx.data(), dd.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
}

void finish_minmax(rgn_ftable* t) {
  t->xmin = t->x[0];
  t->xmax = t->x[0];
  for (float v : t->x) {
    t->xmin = v < t->xmin ? v : t->xmin;
    t->xmax = v > t->xmax ? v : t->xmax;
  }
  // tabulation.cpp:104-117 (verify)
  RGN_REQUIRE(t->y.size() == t->x.size(), RGN_ERR_RANGE, "y.size != x.size in TabulatedFunction");
  RGN_REQUIRE(t->xmin < t->xmax, RGN_ERR_RANGE, "xmin >= xmax in TabulatedFunction");
  RGN_REQUIRE(t->xmin > 0.0f, RGN_ERR_RANGE, "xmin <= 0.0 in Logspace TabulatedFunction");
}

void build_reference(rgn_ftable* t, int n, float xmin, float xmax) {
  t->x.resize(n);
  t->y.assign(n, 0.0f);
  logspace_host(xmin, xmax, n, t->x.data());
  // Gather every inner-grid point of every node that takes the integral branch.
  std::vector<int> node_of;        // node index of each integral node
  std::vector<float> grid;         // 100 floats per integral node
  for (int i = 0; i < n; ++i) {
    const float x = t->x[i];
    if (x < 1e-6) continue;  // synchrotron.cpp:58 (float vs double literal)
    if (x < 1e-5) {          // synchrotron.cpp:29-31
      t->y[i] = static_cast<float>(4.0 * M_PI / (std::sqrt(3.0) * std::tgamma(kOneThird)) *
                                   std::pow(0.5 * x, 1.0 / 3.0));
      continue;
    }
    if (x > 20) continue;  // synchrotron.cpp:32-33
    node_of.push_back(i);
    const size_t base = grid.size();
    grid.resize(base + 100);
    logspace_host(x, 20.0f, 100, grid.data() + base);  // synchrotron.cpp:35
  }
  std::vector<double> gx(grid.begin(), grid.end()), k;
  run_bessel(kFiveThird, gx, k);  // synchrotron.cpp:37-39, FP64 on the GPU
  for (size_t j = 0; j < node_of.size(); ++j) {
    const float* xs = grid.data() + j * 100;
    float ys[100];
    for (int m = 0; m < 100; ++m) ys[m] = static_cast<float>(k[j * 100 + m]);
    float integrand = 0.0f;  // synchrotron.cpp:40-43: float sum of double terms
    for (int m = 0; m < 99; ++m) {
      integrand = static_cast<float>(static_cast<double>(integrand) +
                                     0.5 * (ys[m] + ys[m + 1]) * (xs[m + 1] - xs[m]));
    }
    const int i = node_of[j];
    t->y[i] = t->x[i] * integrand;  // synchrotron.cpp:44
  }
}

void build_accurate(rgn_ftable* t, int n, float xmin, float xmax) {
  t->x.resize(n);
  logspace_host(xmin, xmax, n, t->x.data());
  std::vector<double> xd(t->x.begin(), t->x.end()), F, k, d;
  run_ffunc(xd, F, k, d);
  t->y.resize(n);
  for (int i = 0; i < n; ++i) t->y[i] = static_cast<float>(F[i]);
}

}  // namespace
}  // namespace rgn

using namespace rgn;

extern "C" {

int rgn_ftable_build(int mode, int npoints, float xmin, float xmax, rgn_ftable** out) {
  return guarded([&] {
    RGN_REQUIRE(out != nullptr, RGN_ERR_INVALID, "null output handle");
    RGN_REQUIRE(npoints >= 2, RGN_ERR_INVALID, "F table needs at least 2 points");
    // snippets.cpp:41-47 (Logspace argument checks)
    RGN_REQUIRE(xmin > 0.0f && xmax > 0.0f, RGN_ERR_INVALID,
                "Logspace start and stop must be strictly positive");
    RGN_REQUIRE(xmin < xmax, RGN_ERR_INVALID, "Logspace start must be < stop");
    RGN_REQUIRE(mode == RGN_FTABLE_REFERENCE || mode == RGN_FTABLE_ACCURATE, RGN_ERR_INVALID,
                "unknown F table mode");
    auto* t = new rgn_ftable();
    try {
      t->n = npoints;
      t->mode = mode;
      if (mode == RGN_FTABLE_REFERENCE) {
        build_reference(t, npoints, xmin, xmax);
      } else {
        build_accurate(t, npoints, xmin, xmax);
      }
      finish_minmax(t);
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

int rgn_ftable_from_host(const float* h_x, const float* h_y, int n, rgn_ftable** out) {
  return guarded([&] {
    RGN_REQUIRE(out != nullptr && h_x != nullptr && h_y != nullptr, RGN_ERR_INVALID,
                "null argument");
    RGN_REQUIRE(n >= 2, RGN_ERR_RANGE, "F table needs at least 2 points");
    auto* t = new rgn_ftable();
    t->n = n;
    t->mode = -1;
    t->x.assign(h_x, h_x + n);
    t->y.assign(h_y, h_y + n);
    try {
      finish_minmax(t);
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

int rgn_ftable_nodes(const rgn_ftable* t, float* h_x, float* h_y) {
  return guarded([&] {
    RGN_REQUIRE(t != nullptr, RGN_ERR_INVALID, "null table");
    if (h_x) std::copy(t->x.begin(), t->x.end(), h_x);
    if (h_y) std::copy(t->y.begin(), t->y.end(), h_y);
  });
}

int rgn_ftable_size(const rgn_ftable* t) { return t ? t->n : 0; }

int rgn_ftable_aux(const rgn_ftable* t, double* h_k53, double* h_dfdx) {
  return guarded([&] {
    RGN_REQUIRE(t != nullptr, RGN_ERR_INVALID, "null table");
    std::vector<double> xd(t->x.begin(), t->x.end()), F, k, d;
    run_ffunc(xd, F, k, d);
    if (h_k53) std::copy(k.begin(), k.end(), h_k53);
    if (h_dfdx) std::copy(d.begin(), d.end(), h_dfdx);
  });
}

int rgn_ftable_destroy(rgn_ftable* t) {
  delete t;
  return RGN_OK;
}

// sync::Ffunc_integrand (synchrotron.cpp:28-46) for one x, K_{5/3} in FP64 on the GPU.
int rgn_ffunc_integrand(float x, float* out) {
  return guarded([&] {
    RGN_REQUIRE(out != nullptr, RGN_ERR_INVALID, "null output");
    if (x < 1e-5) {
      *out = static_cast<float>(4.0 * M_PI / (std::sqrt(3.0) * std::tgamma(kOneThird)) *
                                std::pow(0.5 * x, 1.0 / 3.0));
      return;
    }
    if (x > 20) {
      *out = 0.0f;
      return;
    }
    float xs[100];
    logspace_host(x, 20.0f, 100, xs);
    std::vector<double> gx(xs, xs + 100), k;
    run_bessel(kFiveThird, gx, k);
    float ys[100];
    for (int m = 0; m < 100; ++m) ys[m] = static_cast<float>(k[m]);
    float integrand = 0.0f;
    for (int m = 0; m < 99; ++m) {
      integrand = static_cast<float>(static_cast<double>(integrand) +
                                     0.5 * (ys[m] + ys[m + 1]) * (xs[m + 1] - xs[m]));
    }
    *out = x * integrand;
  });
}

int rgn_bessel_k(double nu, const double* d_x, double* d_out, size_t n, rgn_stream_t stream) {
  return guarded([&] {
    if (!n) return;
    bessel_k_kernel<<<static_cast<unsigned>(ceil_div(n, 256)), 256, 0, as_stream(stream)>>>(
      nu, d_x, d_out, n);
    RGN_HIP(hipGetLastError());
  });
}

}  // extern "C"
