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

// Constants of the emission formula with the correctly rounded reciprocals of
// its fixed divisors (see omega_terms).
struct EmitConst {
  float B0, g_syn, e_at;
  double rB0;  // RN(1 / B0)
  double Dg;   // the reference's e_peak divisor: float(g_syn * g_syn) as double
  double rDg;  // RN(1 / Dg)
  int fast;    // B0 and Dg are normal and their reciprocals finite
};

struct PlanDev {
  const float2* table;  // (c_j, s_j), TP entries
  const float2* bincs;  // (c_e, S_e), nbins_pad entries
  const float* bine;    // e_e, 0 for dead bins
  int TP;
  int nbins;
  int nbins_pad;
  int Z;
  EmitConst em;
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

// The raw inputs of one particle (loaded ahead of their use).
struct RawParticle {
  float ux, uy, uz, ex, ey, ez, bx, by, bz, w;
};

// Correctly rounded a / b from y = RN(1 / b).  Markstein's theorem (radix 2,
// round to nearest, no underflow or overflow): if y is within half an ulp of
// 1/b and q0 within one ulp of a/b, the residual a - b q0 is exact in one FMA
// and q0 + (a - b q0) y rounds to RN(a/b).  q0 = RN(a y) is within one ulp.
// Three dependent FMA-class operations instead of a full division.
__device__ __forceinline__ double div_rn(double a, double b, double y) {
  const double q0 = a * y;
  return fma(fma(-q0, b, a), y, q0);
}

// sqrt(x) correctly rounded, for finite x with 2^-767 < x: the reciprocal
// square root refinement the compiler expands sqrt into (hardware estimate, one
// Goldschmidt step, two Newton corrections of the root), without the input
// scaling it adds for tiny x and the special-value selects.  Equal to sqrt()
// on this range (checked on the CPU with the estimate perturbed by 2^-22).
__device__ __forceinline__ double sqrt_rn_normal(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

// Intermediate terms of OmegaSync_ChiR for one particle (the adjoint reuses them).
struct OmegaTerms {
  double gamma, ig;        // gamma and RN(1 / gamma)
  double btx, bty, btz;    // beta = u / gamma
  double bde;              // beta . E
  double vx, vy, vz;       // V = E + beta x B
  double rootQ;            // sqrt(|V|^2 - (beta . E)^2)
  float e_peak, chiR;
};

// synchrotron.hpp:181-232 (Kernel<D>::OmegaSync_ChiR), evaluated with the
// reference's promotions: squares of the float momenta are float products,
// everything from gamma on is double, chiR is rounded to float and re-read as
// float in e_peak.  The library is compiled with -ffp-contract=off, so these
// expressions round exactly as written.  The five double divisions (three by
// gamma, one by B0, one by g_syn^2) are correctly rounded through div_rn, which
// returns the same doubles as the division operator; a particle with a
// non-finite gamma, or a plan whose divisors are not normal, takes the
// division operator itself.
// HAS_E = false: the particle store has no electric field.  Then beta.E is
// exactly 0 and V = beta x B, the reference's values up to the sign of zero
// terms that only enter squared.
template <bool HAS_E = true>
__device__ __forceinline__ OmegaTerms omega_terms(const RawParticle& r, const EmitConst& C) {
  OmegaTerms o;
  const float ux = r.ux, uy = r.uy, uz = r.uz, ex = r.ex, ey = r.ey, ez = r.ez;
  const float bx = r.bx, by = r.by, bz = r.bz;
  const double g2 = 1.0 + ux * ux + uy * uy + uz * uz;  // >= 1, or inf / NaN
  const bool fast = C.fast && g2 < 1e300;               // false for inf and NaN
  const double gamma = fast ? sqrt_rn_normal(g2) : sqrt(g2);
  o.gamma = gamma;
  if (fast) {
    const double ig = 1.0 / gamma;
    o.ig = ig;
    o.btx = div_rn(ux, gamma, ig);
    o.bty = div_rn(uy, gamma, ig);
    o.btz = div_rn(uz, gamma, ig);
  } else {
    o.ig = 1.0 / gamma;
    o.btx = ux / gamma;
    o.bty = uy / gamma;
    o.btz = uz / gamma;
  }
  if (HAS_E) {
    o.bde = o.btx * ex + o.bty * ey + o.btz * ez;
    o.vx = ex + (o.bty * bz - o.btz * by);
    o.vy = ey + (o.btz * bx - o.btx * bz);
    o.vz = ez + (o.btx * by - o.bty * bx);
  } else {
    o.bde = 0.0;
    o.vx = o.bty * bz - o.btz * by;
    o.vy = o.btz * bx - o.btx * bz;
    o.vz = o.btx * by - o.bty * bx;
  }
  const double v_sqr = o.vx * o.vx + o.vy * o.vy + o.vz * o.vz;
  const double q = HAS_E ? v_sqr - o.bde * o.bde : v_sqr;
  if (fast) {
    // from float inputs q is 0, negative (rounding: NaN, as sqrt), or above
    // 1e-90 and finite
    o.rootQ = q > 0.0 ? sqrt_rn_normal(q) : (q == 0.0 ? 0.0 : __builtin_nan(""));
  } else {
    o.rootQ = sqrt(q);
  }
  if (fast) {
    o.chiR = static_cast<float>(div_rn(o.rootQ, static_cast<double>(C.B0), C.rB0));
    o.e_peak = static_cast<float>(div_rn(C.e_at * gamma * gamma * o.chiR, C.Dg, C.rDg));
  } else {
    o.chiR = static_cast<float>(o.rootQ / C.B0);
    o.e_peak = static_cast<float>(C.e_at * gamma * gamma * o.chiR / (C.g_syn * C.g_syn));
  }
  return o;
}

// 1 / x to about one ulp: hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  return fma(y, fma(-x, y, 1.0), y);
}

// sqrt(x) to about one ulp for finite x > 0 (hardware reciprocal square root
// estimate and two Goldschmidt steps), without special cases: x = 0 and
// x = inf give NaN (so do negative and NaN x).  omega_terms_approx uses it:
// there gamma^2 >= 1, and a zero rootQ (chiR = 0, e_peak = 0) and a NaN one
// both leave e_peak outside (0, inf), which drops the particle, so the NaN is
// never used.
__device__ __forceinline__ double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  r = fma(-g, h, 0.5);
  return fma(g, r, g);
}

// The OmegaSync terms to about one ulp per operation, for the adjoint's
// gradients (which need no correct rounding): reciprocal products instead of
// divisions, refined square roots.  A particle's e_peak may differ from the
// reference's in the last float bit; the forward kernels use omega_terms.
template <bool HAS_E = true>
__device__ __forceinline__ OmegaTerms omega_terms_approx(const RawParticle& r, const EmitConst& C) {
  OmegaTerms o;
  const float ux = r.ux, uy = r.uy, uz = r.uz, ex = r.ex, ey = r.ey, ez = r.ez;
  const float bx = r.bx, by = r.by, bz = r.bz;
  const double gamma = sqrt_nr(1.0 + ux * ux + uy * uy + uz * uz);
  const double ig = rcp_nr(gamma);
  o.gamma = gamma;
  o.ig = ig;
  o.btx = ux * ig;
  o.bty = uy * ig;
  o.btz = uz * ig;
  if (HAS_E) {
    o.bde = o.btx * ex + o.bty * ey + o.btz * ez;
    o.vx = ex + (o.bty * bz - o.btz * by);
    o.vy = ey + (o.btz * bx - o.btx * bz);
    o.vz = ez + (o.btx * by - o.bty * bx);
  } else {
    o.bde = 0.0;
    o.vx = o.bty * bz - o.btz * by;
    o.vy = o.btz * bx - o.btx * bz;
    o.vz = o.btx * by - o.bty * bx;
  }
  const double v_sqr = o.vx * o.vx + o.vy * o.vy + o.vz * o.vz;
  o.rootQ = sqrt_nr(HAS_E ? v_sqr - o.bde * o.bde : v_sqr);
  o.chiR = static_cast<float>(o.rootQ * C.rB0);
  o.e_peak = static_cast<float>(C.e_at * gamma * gamma * o.chiR * C.rDg);
  return o;
}

// log2 of a positive finite float, accurate to ~1e-15 absolute:
// x = m 2^E with m in [sqrt(1/2), sqrt(2)), log2 m = (2 / ln 2) atanh(z),
// z = (m - 1) / (m + 1), |z| <= 0.1716, odd series through z^15 (truncation
// below 2e-14).  z is formed with a refined reciprocal (relative error ~1e-16).
__device__ __forceinline__ double log2_pos(float xf) {
  int E;
  double m = frexp(static_cast<double>(xf), &E);  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    E -= 1;
  }
  const double num = m - 1.0, den = m + 1.0;  // both exact
  double y = __builtin_amdgcn_rcp(den);
  y = fma(y, fma(-den, y, 1.0), y);
  y = fma(y, fma(-den, y, 1.0), y);
  double z = num * y;
  z = fma(fma(-z, den, num), y, z);
  const double s = z * z;
  constexpr double c1 = 2.0 / 0.69314718055994530942;
  double p = c1 / 15.0;
  p = fma(p, s, c1 / 13.0);
  p = fma(p, s, c1 / 11.0);
  p = fma(p, s, c1 / 9.0);
  p = fma(p, s, c1 / 7.0);
  p = fma(p, s, c1 / 5.0);
  p = fma(p, s, c1 / 3.0);
  p = fma(p, s, c1);
  return static_cast<double>(E) + z * p;
}

// Per-source interpolation parameters: weight W (multiplies F) and B split as
// V_hi + V_lo.  A source that contributes to no bin gets W = 0 and a V that
// keeps u inside the padded table.
struct Source {
  float W, vhi, vlo;
};

__device__ __forceinline__ Source make_source(float W, float eps, const PlanDev& P) {
  Source s{0.0f, P.v_safe, 0.0f};
  if (eps > 0.0f && eps < INFINITY) {  // synchrotron.hpp:162 (`e_peak > 0.0`); NaN fails too
    const double B = log2_pos(eps) * P.K;
    if (B > P.win_lo && B < P.win_hi) {
      const double hi = rint(B);
      s.W = W;
      s.vhi = static_cast<float>(hi);
      s.vlo = static_cast<float>(B - hi);
    }
  }
  return s;
}

// One entry of the distribution kernel matrix K[g, e] (KernelFromDist,
// synchrotron.hpp:72-95): the source of gamma-beta bin g, eps = e_at gb^2 /
// g_syn^2 and W = gb for log-spaced distribution bins (synchrotron.hpp:78-79),
// evaluated at photon bin e as the emission kernels evaluate a source.
__device__ __forceinline__ Source dist_source(float gb, int dist_log, const PlanDev& P) {
  const float eps = P.em.e_at * gb * gb / (P.em.g_syn * P.em.g_syn);
  return make_source(dist_log ? gb : 1.0f, eps, P);
}

__device__ __forceinline__ float dist_kernel_value(const Source& s, int e, const PlanDev& P) {
  if (s.W == 0.0f) return 0.0f;
  const float2 cs_e = P.bincs[e];
  const float u = fmaf(cs_e.x, cs_e.y, s.vhi) + s.vlo;
  const float2 cs = P.table[cvt_sat_u32(u)];
  return P.bine[e] * (s.W * fmaf(u, cs.y, cs.x));
}

__device__ __forceinline__ float load_or_zero(const float* p, size_t i) {
  return p ? p[i] : 0.0f;
}

// Element i through a 32-bit byte offset (callers guarantee i * sizeof(T) <
// 2^32): the compiler then addresses with the uniform 64-bit base in scalar
// registers plus a 32-bit per-lane offset, instead of 64-bit per-lane
// address arithmetic for every access.
// NT = true marks a non-temporal access (scalar types only): data streamed
// through once, which the caches then hold only briefly, leaving them to the
// tables that other lanes re-read.
template <bool NT = false, class T>
__device__ __forceinline__ T ld32(const T* base, unsigned i) {
  const T* a = reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * static_cast<unsigned>(sizeof(T)));
  if constexpr (NT) {
    return __builtin_nontemporal_load(a);
  } else {
    return *a;
  }
}
template <bool NT = false, class T>
__device__ __forceinline__ void st32(T* base, unsigned i, T v) {
  T* a = reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * static_cast<unsigned>(sizeof(T)));
  if constexpr (NT) {
    __builtin_nontemporal_store(v, a);
  } else {
    *a = v;
  }
}

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

// load_particle for stores of fewer than 2^29 particles (32-bit byte offsets).
template <bool NT = false>
__device__ __forceinline__ RawParticle load_particle32(const ParticlesDev& D, unsigned p) {
  RawParticle r;
  r.ux = ld32<NT>(D.u[0], p);
  r.uy = ld32<NT>(D.u[1], p);
  r.uz = ld32<NT>(D.u[2], p);
  r.ex = D.e[0] ? ld32<NT>(D.e[0], p) : 0.0f;
  r.ey = D.e[1] ? ld32<NT>(D.e[1], p) : 0.0f;
  r.ez = D.e[2] ? ld32<NT>(D.e[2], p) : 0.0f;
  r.bx = ld32<NT>(D.b[0], p);
  r.by = ld32<NT>(D.b[1], p);
  r.bz = ld32<NT>(D.b[2], p);
  r.w = D.w ? ld32<NT>(D.w, p) : 1.0f;
  return r;
}

__device__ __forceinline__ Source particle_source(const RawParticle& r, const PlanDev& P) {
  const OmegaTerms o = omega_terms(r, P.em);
  return make_source(o.chiR * r.w, o.e_peak, P);
}

// Source of particle p of a particle store.
__device__ __forceinline__ Source particle_source(const ParticlesDev& D, size_t p,
                                                  const PlanDev& P) {
  return particle_source(load_particle(D, p), P);
}

// Gradients of L = sum_e gbar_e S_e with respect to one particle's w, u, B, E,
// given A = sum_e g'_e F_e and Cl = sum_e g'_e dF_e/dln x (g'_e = gbar_e e_e)
// at that particle.  With chi = rootQ / B0, W = chi w and x_e = e_e / eps,
// eps = e_at gamma^2 chi / g_syn^2:
//   dL/dw = chi A,  dL/dchi = w (A - Cl),  dL/dgamma = -2 w chi Cl / gamma,
//   dchi/dV = V / (B0 rootQ), dchi/d(beta.E) = -(beta.E) / (B0 rootQ),
//   dbeta_i/du_j = delta_ij / gamma - u_i u_j / gamma^3, dgamma/du_j = u_j / gamma.
// Gradients need no correct rounding: the divisions are reciprocal products.
struct ParticleGrads {
  double gw, gu[3], gb[3], ge[3];
};


// E_TERMS = false: E is zero and dL/dE is not wanted (ge is left zero), so
// the terms in E drop out.
template <bool E_TERMS = true>
__device__ __forceinline__ ParticleGrads emission_grads(const OmegaTerms& o, const RawParticle& r,
                                                        double A, double Cl, const EmitConst& C) {
  ParticleGrads g;
  const double ig = o.ig;
  const double chi = o.rootQ * C.rB0;
  const double w = r.w;
  g.gw = chi * A;
  const double gchi = w * (A - Cl);
  const double ggamma = -2.0 * w * chi * Cl * ig;
  const double coef = gchi * C.rB0 * rcp_nr(o.rootQ);
  const double ex = r.ex, ey = r.ey, ez = r.ez, bx = r.bx, by = r.by, bz = r.bz;
  // V / (B0 rootQ) scaled by dL/dchi: the common factor of the terms below
  // (fused multiply-adds throughout: the gradients need no correct rounding)
  const double cx = coef * o.vx, cy = coef * o.vy, cz = coef * o.vz;
  // dchi/dB = (V x beta) / (B0 rootQ)
  g.gb[0] = fma(cy, o.btz, -cz * o.bty);
  g.gb[1] = fma(cz, o.btx, -cx * o.btz);
  g.gb[2] = fma(cx, o.bty, -cy * o.btx);
  double gbtx, gbty, gbtz;
  if constexpr (E_TERMS) {
    const double cb = coef * o.bde;
    // dchi/dE = (V - (beta.E) beta) / (B0 rootQ)
    g.ge[0] = fma(-cb, o.btx, cx);
    g.ge[1] = fma(-cb, o.bty, cy);
    g.ge[2] = fma(-cb, o.btz, cz);
    // dchi/dbeta = (B x V - (beta.E) E) / (B0 rootQ)
    gbtx = fma(by, cz, fma(-bz, cy, -cb * ex));
    gbty = fma(bz, cx, fma(-bx, cz, -cb * ey));
    gbtz = fma(bx, cy, fma(-by, cx, -cb * ez));
  } else {
    g.ge[0] = g.ge[1] = g.ge[2] = 0.0;
    gbtx = fma(by, cz, -bz * cy);
    gbty = fma(bz, cx, -bx * cz);
    gbtz = fma(bx, cy, -by * cx);
  }
  const double ux = r.ux, uy = r.uy, uz = r.uz;
  const double udg = fma(ux, gbtx, fma(uy, gbty, uz * gbtz));
  const double t = fma(-udg * ig, ig, ggamma);
  g.gu[0] = ig * fma(ux, t, gbtx);
  g.gu[1] = ig * fma(uy, t, gbty);
  g.gu[2] = ig * fma(uz, t, gbtz);
  return g;
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
