// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference's radiation hot path (haykh/ragnar), used as
// the parity checker by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py.  Nothing in the product (ragnar_amd/) links,
// loads or calls this file.
//
// The reference itself cannot be compiled here: it needs Kokkos 4.5.01, which
// it fetches from GitHub at configure time (cmake/dependencies.cmake:25-35).
// This file therefore restates the reference arithmetic with the same
// FP32/FP64 promotion points, the same libm/libstdc++ calls and the same
// quirks.  It is compiled with g++ (libstdc++ provides std::cyl_bessel_k, the
// function the reference calls) and -ffp-contract=off so that no multiply-add
// is fused, as in the reference's x86-64 Release build.
//
// Three evaluation modes for the per-pair sums:
//   ORC_FAITHFUL (0): per-thread FP32 accumulators summed at the end, the way
//                     Kokkos::Experimental::ScatterView duplicates on OpenMP.
//                     This is the timed CPU baseline.
//   ORC_EXACT    (1): the reference's per-eval FP32 terms, summed in FP64.
//                     This is the parity oracle.
//   ORC_FP64     (2): the same discretised model (same float F table) with every
//                     operation in FP64: isolates the reference's own FP32
//                     rounding error.
//
// Parity pins for each function are listed in the test files that use them.

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

using real_t = float;

// synchrotron.cpp:21-22
constexpr double kOneThird  = 0.3333333333333333;
constexpr double kFiveThird = 1.6666666666666667;

// snippets.cpp:39-62.  The exponent is formed in float (log10f); the call
// math::pow(10, float) promotes both arguments to double (Kokkos
// promote_2_t<int,float> is double) and the result is rounded to float on
// store.
void logspace_f(real_t start, real_t stop, std::size_t num, real_t* out) {
  for (std::size_t i = 0; i < num; ++i) {
    if (num == 1) {
      out[i] = start;
    } else {
      const real_t e = std::log10(start) +
                       static_cast<real_t>(i) * (std::log10(stop / start)) /
                         static_cast<real_t>(num - 1);
      out[i] = static_cast<real_t>(std::pow(10.0, static_cast<double>(e)));
    }
  }
}

// snippets.cpp:21-37.  `i * (stop - start)` converts the size_t index to float.
void linspace_f(real_t start, real_t stop, std::size_t num, real_t* out) {
  for (std::size_t i = 0; i < num; ++i) {
    if (num == 1) {
      out[i] = start;
    } else {
      out[i] = start + static_cast<real_t>(i) * (stop - start) /
                         static_cast<real_t>(num - 1);
    }
  }
}

// synchrotron.cpp:28-46
real_t ffunc_integrand(real_t x) {
  if (x < 1e-5) {
    return static_cast<real_t>(4.0 * M_PI / (std::sqrt(3.0) * std::tgamma(kOneThird)) *
                               std::pow(0.5 * x, 1.0 / 3.0));
  } else if (x > 20) {
    return 0.0f;
  }
  real_t xs[100];
  logspace_f(x, 20.0f, 100, xs);
  real_t ys[100];
  for (int i = 0; i < 100; ++i) {
    ys[i] = static_cast<real_t>(std::cyl_bessel_k(kFiveThird, static_cast<double>(xs[i])));
  }
  real_t integrand = 0.0f;
  for (int i = 0; i < 99; ++i) {
    // 0.5 * (float) * (float) is a double; `float += double` rounds once.
    integrand = static_cast<real_t>(static_cast<double>(integrand) +
                                    0.5 * (ys[i] + ys[i + 1]) * (xs[i + 1] - xs[i]));
  }
  return x * integrand;
}

// tabulation.hpp:19-53 (LG = true branch)
inline real_t interp_log(real_t x0, const real_t* x, const real_t* y, std::size_t n,
                         real_t xmin, real_t xmax, real_t yfill) {
  if (x0 < xmin or x0 >= xmax) {
    return yfill;
  }
  const auto xi = static_cast<std::size_t>(static_cast<real_t>(n - 1) *
                                           std::fabs(std::log10(x0 / xmin)) /
                                           std::log10(xmax / xmin));
  if (xi >= n - 1) {
    return y[n - 1];
  }
  return (y[xi + 1] * std::log10(x0 / x[xi]) + y[xi] * std::log10(x[xi + 1] / x0)) /
         std::log10(x[xi + 1] / x[xi]);
}

// FP64 twin of interp_log used by ORC_FP64: same float nodes, double arithmetic.
inline double interp_log_d(double x0, const real_t* x, const real_t* y, std::size_t n,
                           double xmin, double xmax) {
  if (x0 < xmin or x0 >= xmax) {
    return 0.0;
  }
  auto xi = static_cast<std::size_t>(static_cast<double>(n - 1) *
                                     std::fabs(std::log10(x0 / xmin)) /
                                     std::log10(xmax / xmin));
  if (xi >= n - 1) {
    return y[n - 1];
  }
  // the double index can land one node off the float one near a node; the
  // interpolant is continuous so either segment gives the same value
  return (y[xi + 1] * std::log10(x0 / x[xi]) + y[xi] * std::log10(x[xi + 1] / x0)) /
         std::log10(static_cast<double>(x[xi + 1]) / x[xi]);
}

// synchrotron.hpp:181-232.  `1.0 + ux*ux + ...` squares in float, then sums in
// double; everything after gamma is double; chiR is rounded to float when
// stored through the reference parameter and is re-read as float for e_peak.
// The e' (eprime) block at :209-217 is dead code and is omitted.
inline void omega_sync_chiR(real_t& e_peak, real_t& chiR, real_t ux, real_t uy, real_t uz,
                            real_t ex, real_t ey, real_t ez, real_t bx, real_t by,
                            real_t bz, real_t B0, real_t g_syn, real_t e_at) {
  const double gamma  = std::sqrt(1.0 + ux * ux + uy * uy + uz * uz);
  const double beta_x = ux / gamma;
  const double beta_y = uy / gamma;
  const double beta_z = uz / gamma;

  const double beta_dot_e = beta_x * ex + beta_y * ey + beta_z * ez;

  const double bxb_x = beta_y * bz - beta_z * by;
  const double bxb_y = beta_z * bx - beta_x * bz;
  const double bxb_z = beta_x * by - beta_y * bx;

  const double v_x = ex + bxb_x;
  const double v_y = ey + bxb_y;
  const double v_z = ez + bxb_z;

  const double v_sqr = v_x * v_x + v_y * v_y + v_z * v_z;

  chiR   = static_cast<real_t>(std::sqrt(v_sqr - beta_dot_e * beta_dot_e) / B0);
  e_peak = static_cast<real_t>(e_at * gamma * gamma * chiR / (g_syn * g_syn));
}

// FP64 twin: no float rounding of the squares or of chiR / e_peak.
inline void omega_sync_chiR_d(double& e_peak, double& chiR, double ux, double uy,
                              double uz, double ex, double ey, double ez, double bx,
                              double by, double bz, double B0, double g_syn,
                              double e_at) {
  const double gamma  = std::sqrt(1.0 + ux * ux + uy * uy + uz * uz);
  const double beta_x = ux / gamma, beta_y = uy / gamma, beta_z = uz / gamma;
  const double beta_dot_e = beta_x * ex + beta_y * ey + beta_z * ez;
  const double v_x = ex + (beta_y * bz - beta_z * by);
  const double v_y = ey + (beta_z * bx - beta_x * bz);
  const double v_z = ez + (beta_x * by - beta_y * bx);
  chiR   = std::sqrt(v_x * v_x + v_y * v_y + v_z * v_z - beta_dot_e * beta_dot_e) / B0;
  e_peak = e_at * gamma * gamma * chiR / (g_syn * g_syn);
}

inline real_t comp(const real_t* p, std::size_t i) { return p ? p[i] : 0.0f; }

// ic::KNfunc (ic.hpp:20-25): float arguments; 2 * q * log(q) is float, the
// 0.5 literal makes the second part double, and the sum returns as float.
real_t ic_kn(real_t Gamma, real_t q) {
  const real_t Gq = Gamma * q;
  return 2 * q * std::log(q) + (1 - q) * ((1 + 2 * q) + 0.5 * (Gq) * (Gq) / (1 + Gq));
}

double ic_kn_d(double Gamma, double q) {
  const double Gq = Gamma * q;
  return 2 * q * std::log(q) + (1 - q) * ((1 + 2 * q) + 0.5 * Gq * Gq / (1 + Gq));
}

}  // namespace

extern "C" {

enum { ORC_FAITHFUL = 0, ORC_EXACT = 1, ORC_FP64 = 2 };

int orc_num_threads(void) { return omp_get_max_threads(); }

void orc_logspace(float start, float stop, std::size_t num, float* out) {
  logspace_f(start, stop, num, out);
}

void orc_linspace(float start, float stop, std::size_t num, float* out) {
  linspace_f(start, stop, num, out);
}

float orc_ffunc_integrand(float x) { return ffunc_integrand(x); }

// synchrotron.cpp:48-65.  The node test `xs[i] < 1e-6` compares a float with
// the double literal, so node 0 (9.99999997e-7f) gets y = 0.  The `ys[i] > 100`
// clause reads a zero-initialised vector and never fires.
void orc_tabulate_ffunc(std::size_t n, float xmin, float xmax, float* xs, float* ys) {
  logspace_f(xmin, xmax, n, xs);
  for (std::size_t i = 0; i < n; ++i) {
    ys[i] = 0.0f;
  }
  for (std::size_t i = 0; i < n; ++i) {
    if (xs[i] < 1e-6 or ys[i] > 100.0) {
      ys[i] = 0.0f;
    } else {
      ys[i] = ffunc_integrand(xs[i]);
    }
  }
}

// tabulation.cpp:84-102 (MinMax reduce) + tabulation.hpp:19-53.  Only the
// log-grid branch is restated: the reference instantiates its interpolant with
// LOGGRID = true only (global.h:32; its callers are synchrotron.hpp:82,163), so the linear branch
// of tabulation.hpp:43-52 is never evaluated.
float orc_interp(float x0, const float* x, const float* y, std::size_t n, float yfill) {
  float xmin = x[0], xmax = x[0];
  for (std::size_t i = 0; i < n; ++i) {
    xmin = x[i] < xmin ? x[i] : xmin;
    xmax = x[i] > xmax ? x[i] : xmax;
  }
  return interp_log(x0, x, y, n, xmin, xmax, yfill);
}

// synchrotron.hpp:181-232, exposed for closed-form tests.
void orc_omega_sync(const float* in9, float B0, float g_syn, float e_at, float* e_peak,
                    float* chiR) {
  omega_sync_chiR(*e_peak, *chiR, in9[0], in9[1], in9[2], in9[3], in9[4], in9[5],
                  in9[6], in9[7], in9[8], B0, g_syn, e_at);
}

// SynchrotronSpectrum<D> (synchrotron.cpp:107-145) with Kernel<D>
// (synchrotron.hpp:145-172).  U/E/B are SoA component pointers (a null E or W
// component reads as zero / one).  `w` is a per-particle weight that the
// reference does not have; w == nullptr reproduces it exactly.  The F table is
// passed in (fx, fy, fn) so tests can use the reference default table or a
// planted one.  Out: out[nbins] (float), and for ORC_EXACT / ORC_FP64 also the
// FP64 sum in out_d (may be null).
int orc_sync_spectrum(int mode, const float* ux, const float* uy, const float* uz,
                      const float* ex, const float* ey, const float* ez, const float* bx,
                      const float* by, const float* bz, const float* w, std::size_t np,
                      const float* bins, std::size_t nbins, const float* fx,
                      const float* fy, std::size_t fn, float B0, float g_syn,
                      float e_at, float* out, double* out_d) {
  real_t fxmin = fx[0], fxmax = fx[0];
  for (std::size_t i = 0; i < fn; ++i) {
    fxmin = fx[i] < fxmin ? fx[i] : fxmin;
    fxmax = fx[i] > fxmax ? fx[i] : fxmax;
  }
  if (mode == ORC_FAITHFUL) {
    const int nt = omp_get_max_threads();
    std::vector<real_t> dup(static_cast<std::size_t>(nt) * nbins, 0.0f);
#pragma omp parallel
    {
      real_t* acc = dup.data() + static_cast<std::size_t>(omp_get_thread_num()) * nbins;
#pragma omp for schedule(static)
      for (std::size_t p = 0; p < np; ++p) {
        for (std::size_t e = 0; e < nbins; ++e) {
          const real_t e_syn = bins[e];
          real_t e_peak, chiR;
          omega_sync_chiR(e_peak, chiR, ux[p], uy[p], uz[p], comp(ex, p), comp(ey, p),
                          comp(ez, p), bx[p], by[p], bz[p], B0, g_syn, e_at);
          if (e_peak > 0.0) {
            const real_t F = interp_log(e_syn / e_peak, fx, fy, fn, fxmin, fxmax, 0.0f);
            real_t term    = e_syn * chiR * F;
            if (w) term = term * w[p];
            acc[e] += term;
          }
        }
      }
    }
    for (std::size_t e = 0; e < nbins; ++e) {
      real_t s = 0.0f;
      for (int t = 0; t < nt; ++t) s += dup[static_cast<std::size_t>(t) * nbins + e];
      out[e] = s;
    }
    return 0;
  }
  std::vector<double> sum(nbins, 0.0);
#pragma omp parallel
  {
    std::vector<double> acc(nbins, 0.0);
#pragma omp for schedule(static)
    for (std::size_t p = 0; p < np; ++p) {
      if (mode == ORC_EXACT) {
        real_t e_peak, chiR;
        omega_sync_chiR(e_peak, chiR, ux[p], uy[p], uz[p], comp(ex, p), comp(ey, p),
                        comp(ez, p), bx[p], by[p], bz[p], B0, g_syn, e_at);
        if (!(e_peak > 0.0)) continue;
        for (std::size_t e = 0; e < nbins; ++e) {
          const real_t e_syn = bins[e];
          const real_t F = interp_log(e_syn / e_peak, fx, fy, fn, fxmin, fxmax, 0.0f);
          real_t term    = e_syn * chiR * F;
          if (w) term = term * w[p];
          acc[e] += static_cast<double>(term);
        }
      } else {
        double e_peak, chiR;
        omega_sync_chiR_d(e_peak, chiR, ux[p], uy[p], uz[p], comp(ex, p), comp(ey, p),
                          comp(ez, p), bx[p], by[p], bz[p], B0, g_syn, e_at);
        if (!(e_peak > 0.0)) continue;
        for (std::size_t e = 0; e < nbins; ++e) {
          const double e_syn = bins[e];
          const double F     = interp_log_d(e_syn / e_peak, fx, fy, fn, fxmin, fxmax);
          acc[e] += e_syn * chiR * F * (w ? static_cast<double>(w[p]) : 1.0);
        }
      }
    }
#pragma omp critical
    for (std::size_t e = 0; e < nbins; ++e) sum[e] += acc[e];
  }
  for (std::size_t e = 0; e < nbins; ++e) {
    out[e] = static_cast<real_t>(sum[e]);
    if (out_d) out_d[e] = sum[e];
  }
  return 0;
}

// Gradient of L = sum_e gbar[e] S[e] for the ORC_FP64 spectrum model (same
// float table and nodes, double arithmetic), with respect to each particle's
// weight, momentum, magnetic and electric field.  No reference counterpart: the
// reference is forward-only.  Exact derivative of that model: the interpolant
// is piecewise linear in ln x, so dF/dln x is the slope of the segment that
// interp_log_d evaluates (zero outside the table and past its last node).
// Per particle, with x_e = e_e / eps, eps = e_at gamma^2 chi / g_syn^2:
//   A = sum_e gbar_e e_e F(x_e),  C = sum_e gbar_e e_e dF/dln x(x_e),
//   dL/dw = chi A,  dL/dchi = w (A - C),  dL/dgamma = -2 w chi C / gamma,
// chained through chi = sqrt(q) / B0, q = |E + beta x B|^2 - (beta.E)^2,
// beta = u / gamma, gamma = sqrt(1 + u^2):
//   dq/dE = 2 V - 2 (beta.E) beta,  dq/dB = 2 V x beta,
//   dq/dbeta = 2 B x V - 2 (beta.E) E,  dbeta/du = (I - beta beta^T) / gamma,
//   dgamma/du = beta.
// g* outputs are (3, n) component-major; a null ex/ey/ez or w reads as zero /
// one.  dnode[p] = the smallest distance, in units of the segment, of any of
// the particle's x_e from a table node (where dF/dln x jumps).
int orc_sync_adjoint_fp64(const float* ux, const float* uy, const float* uz, const float* ex,
                          const float* ey, const float* ez, const float* bx, const float* by,
                          const float* bz, const float* w, std::size_t np, const float* bins,
                          std::size_t nbins, const float* fx, const float* fy, std::size_t fn,
                          double B0, double g_syn, double e_at, const double* gbar, double* gw,
                          double* gu, double* gb, double* ge, double* dnode) {
  real_t fxmin = fx[0], fxmax = fx[0];
  for (std::size_t i = 0; i < fn; ++i) {
    fxmin = fx[i] < fxmin ? fx[i] : fxmin;
    fxmax = fx[i] > fxmax ? fx[i] : fxmax;
  }
  const double lxmin = std::log10(static_cast<double>(fxmin));
  const double lspan = std::log10(static_cast<double>(fxmax) / fxmin);
#pragma omp parallel for schedule(static)
  for (std::size_t p = 0; p < np; ++p) {
    const double u[3] = {ux[p], uy[p], uz[p]};
    const double E[3] = {comp(ex, p), comp(ey, p), comp(ez, p)};
    const double B[3] = {bx[p], by[p], bz[p]};
    const double wp = w ? static_cast<double>(w[p]) : 1.0;
    const double gamma = std::sqrt(1.0 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double be[3] = {u[0] / gamma, u[1] / gamma, u[2] / gamma};
    const double bdote = be[0] * E[0] + be[1] * E[1] + be[2] * E[2];
    const double V[3] = {E[0] + (be[1] * B[2] - be[2] * B[1]), E[1] + (be[2] * B[0] - be[0] * B[2]),
                         E[2] + (be[0] * B[1] - be[1] * B[0])};
    const double q = V[0] * V[0] + V[1] * V[1] + V[2] * V[2] - bdote * bdote;
    const double chi = std::sqrt(q) / B0;
    const double eps = e_at * gamma * gamma * chi / (g_syn * g_syn);
    double A = 0.0, Cs = 0.0, dmin = 1.0;
    if (eps > 0.0) {
      for (std::size_t e = 0; e < nbins; ++e) {
        const double x0 = bins[e] / eps;
        if (x0 < fxmin or x0 >= fxmax) continue;
        auto i = static_cast<std::size_t>(static_cast<double>(fn - 1) *
                                          std::fabs(std::log10(x0) - lxmin) / lspan);
        if (i >= fn - 1) {  // the constant y[n-1] (interp_log_d): no slope
          A += gbar[e] * bins[e] * fy[fn - 1];
          continue;
        }
        // the double index can fall one node off near a node; take the
        // segment that holds x0
        if (i + 1 < fn - 1 && x0 >= fx[i + 1]) ++i;
        if (i > 0 && x0 < fx[i]) --i;
        const double l = std::log(static_cast<double>(fx[i + 1]) / fx[i]);
        const double f = std::log(x0 / fx[i]) / l;
        const double slope = (static_cast<double>(fy[i + 1]) - fy[i]) / l;
        const double F = fy[i] + f * (static_cast<double>(fy[i + 1]) - fy[i]);
        A += gbar[e] * bins[e] * F;
        Cs += gbar[e] * bins[e] * slope;
        dmin = std::fmin(dmin, std::fmin(f, 1.0 - f));
      }
    }
    dnode[p] = dmin;
    double o_w = 0.0, o_u[3] = {0, 0, 0}, o_b[3] = {0, 0, 0}, o_e[3] = {0, 0, 0};
    if (eps > 0.0) {
      o_w = chi * A;
      const double gchi = wp * (A - Cs);
      const double ggam = -2.0 * wp * chi * Cs / gamma;
      const double dchi_dq = 1.0 / (2.0 * B0 * B0 * chi);
      const double k = gchi * dchi_dq;
      const double dqdb[3] = {2 * (B[1] * V[2] - B[2] * V[1]) - 2 * bdote * E[0],
                              2 * (B[2] * V[0] - B[0] * V[2]) - 2 * bdote * E[1],
                              2 * (B[0] * V[1] - B[1] * V[0]) - 2 * bdote * E[2]};
      const double bq = be[0] * dqdb[0] + be[1] * dqdb[1] + be[2] * dqdb[2];
      for (int c = 0; c < 3; ++c) {
        o_u[c] = k * (dqdb[c] - be[c] * bq) / gamma + ggam * be[c];
        o_e[c] = k * (2 * V[c] - 2 * bdote * be[c]);
      }
      o_b[0] = k * 2 * (V[1] * be[2] - V[2] * be[1]);
      o_b[1] = k * 2 * (V[2] * be[0] - V[0] * be[2]);
      o_b[2] = k * 2 * (V[0] * be[1] - V[1] * be[0]);
    }
    gw[p] = o_w;
    for (int c = 0; c < 3; ++c) {
      gu[c * np + p] = o_u[c];
      gb[c * np + p] = o_b[c];
      ge[c * np + p] = o_e[c];
    }
  }
  return 0;
}

// ICSpectrum (ic.cpp:15-46) with ic::Kernel::operator() (ic.hpp:62-85): every
// (gamma bin, soft-photon bin, output bin) triple adds
// f_g (f_s / e_s) e_ic^2 KN / g (g^2 for linear gamma bins) below the kinematic
// limit e_ic <= g Gamma / (1 + Gamma), Gamma = 4 g e_s.  The unit checks of the
// kernel's constructor (ic.hpp:52-59) are the caller's.
int orc_ic_spectrum(int mode, const float* g, const float* fg, std::size_t ng, int g_log,
                    const float* es, const float* fes, std::size_t nes, const float* eic,
                    std::size_t neic, float* out, double* out_d) {
#pragma omp parallel for schedule(static)
  for (std::size_t e = 0; e < neic; ++e) {
    real_t sf = 0.0f;
    double sd = 0.0;
    for (std::size_t gi = 0; gi < ng; ++gi) {
      for (std::size_t si = 0; si < nes; ++si) {
        if (mode == ORC_FP64) {
          const double gp = g[gi], sp = es[si], ep = eic[e];
          const double Gamma = 4 * gp * sp;
          if (ep > gp * Gamma / (1 + Gamma)) continue;
          const double q = (ep / gp) / (Gamma * (1 - (ep / gp)));
          sd += static_cast<double>(fg[gi]) * (static_cast<double>(fes[si]) / sp) * ep * ep *
                ic_kn_d(Gamma, q) / (g_log ? gp : gp * gp);
          continue;
        }
        const real_t g_prtls = g[gi];
        const real_t f_prtls = fg[gi];
        const real_t e_soft_photons = es[si];
        const real_t f_soft_photons = fes[si];
        const real_t e_ic = eic[e];
        const real_t Gamma = 4 * g_prtls * e_soft_photons;
        if (e_ic > g_prtls * Gamma / (1 + Gamma)) continue;
        const real_t q = (e_ic / g_prtls) / (Gamma * (1 - (e_ic / g_prtls)));
        const real_t KNval = ic_kn(Gamma, q);
        const real_t term = g_log ? f_prtls * (f_soft_photons / e_soft_photons) * e_ic * e_ic * KNval / g_prtls
                                  : f_prtls * (f_soft_photons / e_soft_photons) * e_ic * e_ic * KNval /
                                      (g_prtls * g_prtls);
        sf += term;
        sd += static_cast<double>(term);
      }
    }
    out[e] = mode == ORC_FAITHFUL ? sf : static_cast<real_t>(sd);
    if (out_d) out_d[e] = mode == ORC_FAITHFUL ? static_cast<double>(sf) : sd;
  }
  return 0;
}

// SynchrotronSpectrumFromDist (synchrotron.cpp:69-105) with KernelFromDist
// (synchrotron.hpp:72-95).  The peak energy is float arithmetic here; there is
// no bins-unit check on this path.
int orc_sync_fromdist(int mode, const float* gbeta, const float* f, std::size_t ng,
                      int dist_log, const float* bins, std::size_t nbins, const float* fx,
                      const float* fy, std::size_t fn, float g_syn, float e_at,
                      float* out, double* out_d) {
  real_t fxmin = fx[0], fxmax = fx[0];
  for (std::size_t i = 0; i < fn; ++i) {
    fxmin = fx[i] < fxmin ? fx[i] : fxmin;
    fxmax = fx[i] > fxmax ? fx[i] : fxmax;
  }
  for (std::size_t e = 0; e < nbins; ++e) {
    real_t sf = 0.0f;
    double sd = 0.0;
    for (std::size_t g = 0; g < ng; ++g) {
      if (mode == ORC_FP64) {
        const double gb = gbeta[g];
        const double ep = static_cast<double>(e_at) * gb * gb /
                          (static_cast<double>(g_syn) * g_syn);
        if (ep > 0.0) {
          const double F = interp_log_d(bins[e] / ep, fx, fy, fn, fxmin, fxmax);
          sd += static_cast<double>(f[g]) * bins[e] * (dist_log ? gb : 1.0) * F;
        }
        continue;
      }
      const real_t e_syn  = bins[e];
      const real_t gb     = gbeta[g];
      const real_t e_peak = e_at * gb * gb / (g_syn * g_syn);
      if (e_peak > 0.0) {
        const real_t F = interp_log(e_syn / e_peak, fx, fy, fn, fxmin, fxmax, 0.0f);
        const real_t term = dist_log ? f[g] * e_syn * gb * F : f[g] * e_syn * F;
        sf += term;
        sd += static_cast<double>(term);
      }
    }
    if (mode == ORC_FAITHFUL) {
      out[e] = sf;
    } else {
      out[e] = static_cast<real_t>(sd);
    }
    if (out_d) out_d[e] = (mode == ORC_FAITHFUL) ? static_cast<double>(sf) : sd;
  }
  return 0;
}

// Particles<D>::energyDistribution (particles.cpp:189-260).  `energy` is a
// double (the ternary mixes sqrt(float) with sqrt(double)); the index uses
// double log10 of energy/emin over float log10f(emax/emin); out-of-range
// energies clamp to the first / last bin; log bins add (float)(1/E), other bins
// add 1.  The index is always logarithmic, also for linear bins (reference
// behaviour).  counts (uint64, may be null) is the number of particles per bin;
// hist gets the FP64 sum of the float terms (ORC_EXACT) or per-thread float
// sums (ORC_FAITHFUL), rounded to float.
int orc_energy_hist(int mode, const float* ux, const float* uy, const float* uz,
                    std::size_t np, const float* bins, std::size_t nb, int log_bins,
                    int fourvel, float* hist, double* hist_d, std::uint64_t* counts) {
  real_t emin = bins[0], emax = bins[0];
  for (std::size_t i = 0; i < nb; ++i) {
    emin = bins[i] < emin ? bins[i] : emin;
    emax = bins[i] > emax ? bins[i] : emax;
  }
  const int nt = omp_get_max_threads();
  std::vector<real_t> dupf(static_cast<std::size_t>(nt) * nb, 0.0f);
  std::vector<double> dupd(static_cast<std::size_t>(nt) * nb, 0.0);
  std::vector<std::uint64_t> dupc(static_cast<std::size_t>(nt) * nb, 0);
#pragma omp parallel
  {
    const std::size_t t = static_cast<std::size_t>(omp_get_thread_num());
    real_t* af = dupf.data() + t * nb;
    double* ad = dupd.data() + t * nb;
    std::uint64_t* ac = dupc.data() + t * nb;
#pragma omp for schedule(static)
    for (std::size_t p = 0; p < np; ++p) {
      const real_t usqr = ux[p] * ux[p] + uy[p] * uy[p] + uz[p] * uz[p];
      const double energy = fourvel ? static_cast<double>(std::sqrt(usqr))
                                    : std::sqrt(1.0 + usqr);
      std::size_t idx;
      if (energy < emin) {
        idx = 0;
      } else if (energy >= emax) {
        idx = nb - 1;
      } else {
        const auto ei = static_cast<std::size_t>(
          static_cast<real_t>(nb - 1) * std::fabs(std::log10(energy / emin)) /
          std::log10(emax / emin));
        idx = ei > nb - 1 ? nb - 1 : ei;
      }
      const real_t term = log_bins ? static_cast<real_t>(1.0 / energy) : 1.0f;
      af[idx] += term;
      ad[idx] += static_cast<double>(term);
      ac[idx] += 1;
    }
  }
  for (std::size_t i = 0; i < nb; ++i) {
    real_t sf = 0.0f;
    double sd = 0.0;
    std::uint64_t c = 0;
    for (int t = 0; t < nt; ++t) {
      sf += dupf[static_cast<std::size_t>(t) * nb + i];
      sd += dupd[static_cast<std::size_t>(t) * nb + i];
      c += dupc[static_cast<std::size_t>(t) * nb + i];
    }
    hist[i] = (mode == ORC_FAITHFUL) ? sf : static_cast<real_t>(sd);
    if (hist_d) hist_d[i] = sd;
    if (counts) counts[i] = c;
  }
  return 0;
}

// Generators, distributions.cpp:37-136.  kind: 0 = Plaw(p, emin, emax),
// 1 = BrokenPlaw(e_break, p1, p2, emin, emax), 2 = Delta(energy0, denergy).
// par holds the constructor arguments in that order.  Returns 0, or 1 when the
// reference constructor would throw.
int orc_generator(int kind, const float* par, const float* e, std::size_t n, float* f) {
  if (kind == 0) {
    const real_t p = par[0], emin = par[1], emax = par[2];
    real_t norm;
    if (emin < 0.0) return 1;
    if (emin > emax and emax > 0.0) return 1;
    if (emin == 0.0 and p <= -1.0) return 1;
    if (emax == 0.0 and p >= -1.0) return 1;
    if (emax == 0.0) {
      norm = -std::pow(emin, p + 1) / (p + 1);
    } else if (p != -1.0) {
      norm = (std::pow(emax, p + 1) - std::pow(emin, p + 1)) / (p + 1);
    } else {
      norm = std::log(emax / emin);
    }
    for (std::size_t i = 0; i < n; ++i) {
      f[i] = (e[i] < emin or (emax > 0.0 and e[i] >= emax)) ? 0.0f
                                                              : std::pow(e[i], p) / norm;
    }
    return 0;
  }
  if (kind == 1) {
    const real_t e_break = par[0], p1 = par[1], p2 = par[2], emin = par[3], emax = par[4];
    if (e_break <= 0.0) return 1;
    if (emin < 0.0) return 1;
    if (emin > emax and emax > 0.0) return 1;
    real_t pre, post;
    if (emin == 0.0 and p1 <= -1.0) return 1;
    if (e_break <= emin) {
      pre = 0.0f;
    } else {
      pre = (1 - std::pow(emin / e_break, p1 + 1)) / (p1 + 1);
    }
    if (emax == 0.0 and p2 >= -1.0) return 1;
    if (emax == 0.0) {
      post = 1 / (-p2 - 1);
    } else if (p2 == -1.0) {
      post = std::log(emax / e_break);
    } else {
      post = (std::pow(emax / e_break, p2 + 1) - 1) / (p2 + 1);
    }
    const real_t norm = (pre + post) * e_break;
    for (std::size_t i = 0; i < n; ++i) {
      if (e[i] < emin or (emax > 0.0 and e[i] >= emax)) {
        f[i] = 0.0f;
      } else if (e[i] < e_break) {
        f[i] = std::pow(e[i] / e_break, p1) / norm;
      } else {
        f[i] = std::pow(e[i] / e_break, p2) / norm;
      }
    }
    return 0;
  }
  if (kind == 2) {
    const real_t energy0 = par[0], denergy = par[1];
    const real_t norm = static_cast<real_t>(1.0 / denergy);
    for (std::size_t i = 0; i < n; ++i) {
      f[i] = (std::abs(e[i] - energy0) < denergy * 0.5) ? static_cast<real_t>(1.0 * norm)
                                                          : 0.0f;
    }
    return 0;
  }
  return 1;
}

// ---------------------------------------------------------------------------
// FP64 physics, independent of the reference's discretisation.
//
// K_nu(x)  = int_0^inf exp(-x cosh t) cosh(nu t) dt
// F(x)     = x int_x^inf K_{5/3}(s) ds = x int_0^inf exp(-x cosh t) cosh(5t/3)/cosh t dt
// The integrands are even, analytic in the strip |Im t| < pi/2 and decay
// double-exponentially, so the trapezoid rule converges geometrically in the
// step; h = 1/64 on [0, tmax] is far beyond FP64 accuracy for x >= 1e-7.
// ---------------------------------------------------------------------------
static double trap_cosh_integral(double x, double nu, bool divide_cosh) {
  const double h = 1.0 / 64.0;
  double sum     = 0.5 * std::exp(-x) * (divide_cosh ? 1.0 : 1.0);
  for (int k = 1;; ++k) {
    const double t  = k * h;
    const double ch = std::cosh(t);
    double term     = std::exp(-x * ch) * std::cosh(nu * t);
    if (divide_cosh) term /= ch;
    sum += term;
    if (x * ch > 750.0 || (term < 1e-300) || (k > 8 && term < 1e-18 * sum && x * ch > 40.0)) {
      break;
    }
  }
  return sum * h;
}

double orc_k53_fp64(double x) { return trap_cosh_integral(x, kFiveThird, false); }

double orc_bessel_k_libstdcxx(double nu, double x) { return std::cyl_bessel_k(nu, x); }

double orc_F_fp64(double x) { return x * trap_cosh_integral(x, kFiveThird, true); }

// dF/dx = F/x - x K_{5/3}(x)
double orc_dFdx_fp64(double x) { return orc_F_fp64(x) / x - x * orc_k53_fp64(x); }

// Gradient oracle for the per-particle emission (FP64 central differences are
// done by the tests; this gives the analytic FP64 per-particle quantities):
// e_peak, chiR in FP64 for one particle.
void orc_omega_sync_fp64(const double* in9, double B0, double g_syn, double e_at,
                         double* e_peak, double* chiR) {
  omega_sync_chiR_d(*e_peak, *chiR, in9[0], in9[1], in9[2], in9[3], in9[4], in9[5],
                    in9[6], in9[7], in9[8], B0, g_syn, e_at);
}

}  // extern "C"
