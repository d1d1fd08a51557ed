// Device-side pieces shared by the emission kernels (per-particle spectrum,
// adjoint, batched cells and the distribution forms).
//
// Interpolation coordinate.  The reference interpolates the F table linearly
// in log10 x between log-uniform nodes (tabulation.hpp:19-53).  With
//   t(x) = (log2 x - log2 x_min) * K,   K = (n - 1) / log2(x_max / x_min),
// node i sits at t = i.  For a photon bin e and a particle p, x = e_e / eps_p,
// so t = A_e - B_p with A_e = (log2 e_e - log2 x_min) K and B_p = log2(eps_p) K.
// (The float nodes sit within a few 1e-6 of integer t; each segment of the
// device table interpolates between the two nodes' exact coordinates.)
// The kernels use the reversed coordinate u = Z - t = U_e + B_p, where Z is the
// first node of the table's all-zero tail (183 for the reference table).  In
// u the F table is a padded array of segments: segment j covers u in [j, j+1)
// with F = c_j + u * s_j.  Entry 0 is (0, 0), so every u < 1 (x beyond the zero
// tail) evaluates to 0 without a range test, and a negative u saturates to
// entry 0 in the float->uint conversion.  Entries past the low end of the table
// (x < x_min) are (0, 0) as well.
//
// Precision.  Rounding errors of u scale with |u|; the table's relative slope
// |s/F| is large only in the exponential tail, which sits at small u, so the
// per-eval error stays ~1e-7 relative.  U_e is stored as an exact-product pair
// (c_e, S_e): c_e * S_e equals U_e to ~1e-9 (a 48-bit product inside one FMA),
// and B_p as an integer-valued float V_hi plus a remainder V_lo in [-0.5, 0.5],
// so u = fma(c_e, S_e, V_hi) + V_lo is rounded only at the magnitude of u.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rgn {

struct ParticlesDev {
  const float* u[3];
  const float* e[3];
  const float* b[3];
  const float* w;
};

struct PlanDev {
  const float2* table;  // (c_j, s_j), TP entries
  const float2* bincs;  // (c_e, S_e), nbins_pad entries
  const float* bine;    // e_e, 0 for dead bins
  int TP;
  int nbins;
  int nbins_pad;
  int Z;
  float B0, g_syn, e_at;
  double K;         // table units per unit of log2 x
  double win_lo;    // a particle contributes iff win_lo < B_p < win_hi
  double win_hi;
  float v_safe;     // V_hi for a particle that contributes nowhere
};

// Hardware float -> uint32 conversion: negative inputs and NaN give 0, inputs
// above 2^32 give 0xffffffff.  Written as an ALU instruction because the C++
// conversion is undefined for negative values and the compiler would add two
// clamps.
__device__ __forceinline__ unsigned cvt_sat_u32(float u) {
  unsigned r;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(u));
  return r;
}

// synchrotron.hpp:181-232 (Kernel<D>::OmegaSync_ChiR), evaluated with the
// reference's promotions: squares of the float momenta are float products,
// everything from gamma on is double, chiR is rounded to float and re-read as
// float in e_peak.  The library is compiled with -ffp-contract=off, so these
// expressions round exactly as written.
__device__ __forceinline__ void omega_sync(float ux, float uy, float uz, float ex, float ey,
                                           float ez, float bx, float by, float bz, float B0,
                                           float g_syn, float e_at, float& e_peak,
                                           float& chiR) {
  const double gamma = sqrt(1.0 + ux * ux + uy * uy + uz * uz);
  const double beta_x = ux / gamma;
  const double beta_y = uy / gamma;
  const double beta_z = uz / gamma;
  const double beta_dot_e = beta_x * ex + beta_y * ey + beta_z * ez;
  const double vx = ex + (beta_y * bz - beta_z * by);
  const double vy = ey + (beta_z * bx - beta_x * bz);
  const double vz = ez + (beta_x * by - beta_y * bx);
  const double v_sqr = vx * vx + vy * vy + vz * vz;
  chiR = static_cast<float>(sqrt(v_sqr - beta_dot_e * beta_dot_e) / B0);
  e_peak = static_cast<float>(e_at * gamma * gamma * chiR / (g_syn * g_syn));
}

// Per-source interpolation parameters: weight W (multiplies F) and B split as
// V_hi + V_lo.  A source that contributes to no bin gets W = 0 and a V that
// keeps u inside the padded table.
struct Source {
  float W, vhi, vlo;
};

__device__ __forceinline__ Source make_source(float W, float eps, const PlanDev& P) {
  Source s{0.0f, P.v_safe, 0.0f};
  if (eps > 0.0f) {  // synchrotron.hpp:162 (`e_peak > 0.0`); NaN fails too
    const double B = log2(static_cast<double>(eps)) * P.K;
    if (B > P.win_lo && B < P.win_hi) {
      const double hi = rint(B);
      s.W = W;
      s.vhi = static_cast<float>(hi);
      s.vlo = static_cast<float>(B - hi);
    }
  }
  return s;
}

__device__ __forceinline__ float load_or_zero(const float* p, size_t i) {
  return p ? p[i] : 0.0f;
}

// The raw inputs of one particle (loaded ahead of their use).
struct RawParticle {
  float ux, uy, uz, ex, ey, ez, bx, by, bz, w;
};

__device__ __forceinline__ RawParticle load_particle(const ParticlesDev& D, size_t p) {
  RawParticle r;
  r.ux = D.u[0][p];
  r.uy = D.u[1][p];
  r.uz = D.u[2][p];
  r.ex = load_or_zero(D.e[0], p);
  r.ey = load_or_zero(D.e[1], p);
  r.ez = load_or_zero(D.e[2], p);
  r.bx = D.b[0][p];
  r.by = D.b[1][p];
  r.bz = D.b[2][p];
  r.w = D.w ? D.w[p] : 1.0f;
  return r;
}

__device__ __forceinline__ Source particle_source(const RawParticle& r, const PlanDev& P) {
  float e_peak, chiR;
  omega_sync(r.ux, r.uy, r.uz, r.ex, r.ey, r.ez, r.bx, r.by, r.bz, P.B0, P.g_syn, P.e_at, e_peak,
             chiR);
  return make_source(chiR * r.w, e_peak, P);
}

// Source of particle p of a particle store.
__device__ __forceinline__ Source particle_source(const ParticlesDev& D, size_t p,
                                                  const PlanDev& P) {
  return particle_source(load_particle(D, p), P);
}

// DPP row/bank controls (GFX9 encoding).
constexpr int kDppQuadPerm1032 = 0xb1;
constexpr int kDppQuadPerm2301 = 0x4e;
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(
    float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xf, false));
}

// Sum over the 64 lanes of a wave; the total lands in lane 63.
__device__ __forceinline__ float wave_sum_to_lane63(float v) {
  v += dpp_f<kDppQuadPerm1032>(v);
  v += dpp_f<kDppQuadPerm2301>(v);
  v += dpp_f<kDppRowHalfMirror>(v);
  v += dpp_f<kDppRowMirror>(v);
  v += dpp_f<kDppRowBcast15, 0xa>(v);
  v += dpp_f<kDppRowBcast31, 0xc>(v);
  return v;
}

}  // namespace rgn
