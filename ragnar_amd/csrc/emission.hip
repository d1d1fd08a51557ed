// Synchrotron emission on gfx950: per-particle spectrum (forward), its adjoint,
// and the spectrum of a tabulated distribution.
//
// Work layout (forward).  A block of 256 threads (4 waves) walks tiles of 256
// sources.  Each thread turns one source into (W, V_hi, V_lo) in FP64 and parks
// it in LDS; then each wave takes 64 of the tile's sources and every lane
// evaluates them against its own photon bins (bin e = lane + 64k, k < NB), so
// the per-source operands are wave-uniform broadcasts and each bin accumulates
// in a register: no atomics.  A lane's 64-term float partial is folded into a
// double per tile; the four waves and then the blocks are combined in a fixed
// order, so results are bitwise reproducible.  The F table lives in LDS as
// (c_j, s_j) segments; one eval is 6 VALU ops + one ds_read_b64.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "emission.hpp"
#include "rgn_core.hpp"

namespace rgn {

const std::vector<float>& ftable_y(const rgn_ftable* t);
const std::vector<float>& ftable_x(const rgn_ftable* t);
float ftable_xmin(const rgn_ftable* t);
float ftable_xmax(const rgn_ftable* t);

namespace {

constexpr int kThreads = 256;
constexpr int kMaxTableEntries = 12288;  // 96 KiB of (c, s) pairs in LDS

// ---------------------------------------------------------------------------
// Exact-product representation of the per-bin coordinate U_e: search c near 1
// such that S = float(U / c) makes c * S closest to U.  One wave per bin.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) product_pair_kernel(const double* __restrict__ U,
                                                          float2* __restrict__ out, int n) {
  const int e = blockIdx.x;
  if (e >= n) return;
  const double target = U[e];
  double best_r = INFINITY;
  int best_k = 0;
  float best_c = 1.0f, best_s = static_cast<float>(target);
  // k in [-2047, 2047]: c = 1 + k 2^-23 above 1, 1 + k 2^-24 below 1
  for (int k = static_cast<int>(threadIdx.x) - 2047; k <= 2047; k += 64) {
    const float c = (k >= 0) ? 1.0f + static_cast<float>(k) * 0x1p-23f
                             : 1.0f + static_cast<float>(k) * 0x1p-24f;
    const float s = static_cast<float>(target / static_cast<double>(c));
    const double r = fabs(target - static_cast<double>(c) * static_cast<double>(s));
    if (r < best_r || (r == best_r && abs(k) < abs(best_k))) {
      best_r = r;
      best_k = k;
      best_c = c;
      best_s = s;
    }
  }
  __shared__ double rs[64];
  __shared__ int ks[64];
  rs[threadIdx.x] = best_r;
  ks[threadIdx.x] = best_k;
  __syncthreads();
  if (threadIdx.x == 0) {
    int arg = 0;
    for (int i = 1; i < 64; ++i) {
      if (rs[i] < rs[arg] || (rs[i] == rs[arg] && abs(ks[i]) < abs(ks[arg]))) arg = i;
    }
    ks[0] = arg;
  }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) == ks[0]) out[e] = make_float2(best_c, best_s);
}

// ---------------------------------------------------------------------------
// Shared inner loop: 64 sources of `src` against the lane's NB bins.
// ---------------------------------------------------------------------------
template <int NB>
__device__ __forceinline__ void eval_sources(const float4* __restrict__ src, int count,
                                             const float (&c)[NB], const float (&S)[NB],
                                             const float2* __restrict__ ltab, float (&acc)[NB]) {
#pragma unroll 4
  for (int j = 0; j < count; ++j) {
    const float4 q = src[j];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const float u = fmaf(c[k], S[k], q.y) + q.z;
      const float2 cs = ltab[cvt_sat_u32(u)];
      acc[k] = fmaf(q.x, fmaf(u, cs.y, cs.x), acc[k]);
    }
  }
}

struct ParticleSourceFn {
  ParticlesDev D;
  using Raw = RawParticle;
  __device__ Raw load(size_t p) const { return load_particle(D, p); }
  __device__ Source make(const Raw& r, const PlanDev& P) const { return particle_source(r, P); }
};

// synchrotron.hpp:72-95: eps = e_at * gb * gb / (g_syn * g_syn) in float;
// the weight is f (times gb for log-spaced distribution bins).
struct DistSourceFn {
  const float* gbeta;
  const float* f;
  int dist_log;
  struct Raw {
    float gb, f;
  };
  __device__ Raw load(size_t g) const { return Raw{gbeta[g], f[g]}; }
  __device__ Source make(const Raw& r, const PlanDev& P) const {
    const float eps = P.em.e_at * r.gb * r.gb / (P.em.g_syn * P.em.g_syn);
    const float W = dist_log ? r.f * r.gb : r.f;
    return make_source(W, eps, P);
  }
};

// With out non-null (a one-block grid) the block writes the spectrum itself,
// out[e] = e_e * sum, exactly what the combine pass would write for one block.
template <int NB, class SourceFn>
__global__ void __launch_bounds__(kThreads) forward_kernel(SourceFn fn, size_t n, PlanDev P,
                                                           double* __restrict__ partials,
                                                           float* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  float4* tile = reinterpret_cast<float4*>(smem);
  float2* ltab = reinterpret_cast<float2*>(smem + kThreads * sizeof(float4));
  for (int j = threadIdx.x; j < P.TP; j += kThreads) ltab[j] = P.table[j];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  float c[NB], S[NB];
  double dacc[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const float2 cs = P.bincs[lane + 64 * k];
    c[k] = cs.x;
    S[k] = cs.y;
    dacc[k] = 0.0;
  }

  const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
  size_t base = static_cast<size_t>(blockIdx.x) * kThreads;
  typename SourceFn::Raw raw{};
  if (base + threadIdx.x < n) raw = fn.load(base + threadIdx.x);
  for (; base < n; base += stride) {
    const size_t p = base + threadIdx.x;
    Source s{0.0f, P.v_safe, 0.0f};
    if (p < n) s = fn.make(raw, P);
    // the next tile's inputs load while this tile is evaluated
    if (p + stride < n) raw = fn.load(p + stride);
    __syncthreads();  // the previous tile has been consumed
    tile[threadIdx.x] = make_float4(s.W, s.vhi, s.vlo, 0.0f);
    __syncthreads();
    float acc[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k] = 0.0f;
    eval_sources<NB>(tile + wave * 64, 64, c, S, ltab, acc);
#pragma unroll
    for (int k = 0; k < NB; ++k) dacc[k] += static_cast<double>(acc[k]);
  }

  // Fixed-order combine of the four waves, then one column of `partials` (or,
  // for a one-block grid, the output).
  __syncthreads();
  double* red = reinterpret_cast<double*>(smem);
#pragma unroll
  for (int k = 0; k < NB; ++k) red[(wave * NB + k) * 64 + lane] = dacc[k];
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * NB; e += kThreads) {
    const int k = e >> 6, l = e & 63;
    const double v = ((red[(0 * NB + k) * 64 + l] + red[(1 * NB + k) * 64 + l]) +
                      red[(2 * NB + k) * 64 + l]) +
                     red[(3 * NB + k) * 64 + l];
    if (out) {
      if (e < P.nbins) {
        const float ee = P.bine[e];
        out[e] = (ee == 0.0f) ? 0.0f : static_cast<float>(static_cast<double>(ee) * v);
      }
    } else {
      partials[static_cast<size_t>(e) * gridDim.x + blockIdx.x] = v;
    }
  }
}

// out[e] = e_e * sum_b partials[e][b] (fixed order), 0 for dead bins.
__global__ void __launch_bounds__(kThreads) combine_kernel(const double* __restrict__ partials,
                                                           int nblocks, const float* __restrict__ bine,
                                                           int nbins, float* __restrict__ out) {
  const int e = blockIdx.x;
  if (e >= nbins) return;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kThreads) s += partials[static_cast<size_t>(e) * nblocks + b];
  __shared__ double red[kThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float ee = bine[e];
    out[e] = (ee == 0.0f) ? 0.0f : static_cast<float>(static_cast<double>(ee) * red[0]);
  }
}

// ---------------------------------------------------------------------------
// Adjoint.  Same tile walk; for each source the lane sums g_e F and g_e s over
// its bins, the wave reduces the two sums with DPP, and the owning thread
// turns (A, C) into gradients in FP64 (recomputing gamma and the pitch terms).
// ---------------------------------------------------------------------------
struct GradOut {
  float* gw;
  float* gu[3];
  float* gb[3];
  float* ge[3];
};

// Sum over groups of LPP lanes (16 or 32); the group total lands in the
// group's last lane (for 16, in every lane of the group).
template <int LPP>
__device__ __forceinline__ float group_sum(float v) {
  v += dpp_f<kDppQuadPerm1032>(v);
  v += dpp_f<kDppQuadPerm2301>(v);
  v += dpp_f<kDppRowHalfMirror>(v);
  v += dpp_f<kDppRowMirror>(v);
  if constexpr (LPP == 32) v += dpp_f<kDppRowBcast15, 0xa>(v);
  return v;
}

// LPP lanes share one source; a wave evaluates 64 / LPP sources at a time, so
// the per-source reduction (2 x log2(LPP) DPP adds) is amortised over the
// group's LPP x NBL bins.  Lane i of a group owns bins e = i + LPP k, k < NBL.
template <int LPP, int NBL>
__global__ void __launch_bounds__(kThreads) adjoint_kernel(ParticlesDev D, size_t n, PlanDev P,
                                                           const float* __restrict__ gbar,
                                                           GradOut G) {
  extern __shared__ __align__(16) unsigned char smem[];
  float4* tile = reinterpret_cast<float4*>(smem);
  float2* ac = reinterpret_cast<float2*>(smem + kThreads * sizeof(float4));
  float2* ltab = reinterpret_cast<float2*>(smem + kThreads * (sizeof(float4) + sizeof(float2)));
  for (int j = threadIdx.x; j < P.TP; j += kThreads) ltab[j] = P.table[j];

  constexpr int PPW = 64 / LPP;  // sources per wave step
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int grp = lane / LPP;
  const int li = lane % LPP;
  float c[NBL], S[NBL], g[NBL];
#pragma unroll
  for (int k = 0; k < NBL; ++k) {
    const int e = li + LPP * k;
    const float2 cs = P.bincs[e];
    c[k] = cs.x;
    S[k] = cs.y;
    g[k] = (e < P.nbins) ? gbar[e] * P.bine[e] : 0.0f;
  }
  const double dlnx = -P.K / 0.69314718055994530942;  // dF/dln x = s * du/dln x

  const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
  size_t base = static_cast<size_t>(blockIdx.x) * kThreads;
  RawParticle raw{};
  if (base + threadIdx.x < n) raw = load_particle(D, base + threadIdx.x);
  for (; base < n; base += stride) {
    const size_t p = base + threadIdx.x;
    const RawParticle cur = raw;  // kept for the gradient epilogue
    Source s{0.0f, P.v_safe, 0.0f};
    if (p < n) s = particle_source(cur, P);
    if (p + stride < n) raw = load_particle(D, p + stride);  // next tile, in flight
    __syncthreads();
    tile[threadIdx.x] = make_float4(s.W, s.vhi, s.vlo, 0.0f);
    __syncthreads();
    for (int step = 0; step < 64 / PPW; ++step) {
      const int j = wave * 64 + step * PPW + grp;
      const float4 q = tile[j];
      float a = 0.0f, cc = 0.0f;
#pragma unroll
      for (int k = 0; k < NBL; ++k) {
        const float u = fmaf(c[k], S[k], q.y) + q.z;
        const float2 cs = ltab[cvt_sat_u32(u)];
        a = fmaf(g[k], fmaf(u, cs.y, cs.x), a);
        cc = fmaf(g[k], cs.y, cc);
      }
      a = group_sum<LPP>(a);
      cc = group_sum<LPP>(cc);
      if (li == LPP - 1) ac[j] = make_float2(a, cc);
    }
    __syncthreads();
    if (p >= n) continue;
    const float2 mine = ac[threadIdx.x];
    ParticleGrads g{};
    if (s.W != 0.0f) {
      const OmegaTerms o = omega_terms(cur, P.em);
      g = emission_grads(o, cur, mine.x, dlnx * mine.y, P.em);
    }
    const double gw = g.gw, gux = g.gu[0], guy = g.gu[1], guz = g.gu[2];
    const double gbx = g.gb[0], gby = g.gb[1], gbz = g.gb[2];
    const double gex = g.ge[0], gey = g.ge[1], gez = g.ge[2];
    if (G.gw) G.gw[p] = static_cast<float>(gw);
    if (G.gu[0]) G.gu[0][p] = static_cast<float>(gux);
    if (G.gu[1]) G.gu[1][p] = static_cast<float>(guy);
    if (G.gu[2]) G.gu[2][p] = static_cast<float>(guz);
    if (G.gb[0]) G.gb[0][p] = static_cast<float>(gbx);
    if (G.gb[1]) G.gb[1][p] = static_cast<float>(gby);
    if (G.gb[2]) G.gb[2][p] = static_cast<float>(gbz);
    if (G.ge[0]) G.ge[0][p] = static_cast<float>(gex);
    if (G.ge[1]) G.ge[1][p] = static_cast<float>(gey);
    if (G.ge[2]) G.ge[2][p] = static_cast<float>(gez);
  }
}

// Per-particle (e_peak, chiR, W, V_hi, V_lo) for diagnostics; HAS_E as the
// factorised kernels instantiate it (no electric field: the shorter form).
template <bool HAS_E>
__global__ void __launch_bounds__(kThreads) particle_terms_kernel(ParticlesDev D, size_t n,
                                                                  PlanDev P, float* __restrict__ out) {
  const size_t p = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  if (p >= n) return;
  const RawParticle r = load_particle(D, p);
  const OmegaTerms o = omega_terms<HAS_E>(r, P.em);
  const Source s = particle_source(r, P);
  out[5 * p + 0] = o.e_peak;
  out[5 * p + 1] = o.chiR;
  out[5 * p + 2] = s.W;
  out[5 * p + 3] = s.vhi;
  out[5 * p + 4] = s.vlo;
}

}  // namespace

// Component pointers may be null only when there are no particles.
ParticlesDev make_particles_checked(const float* const d_u[3], const float* const d_e[3],
                                    const float* const d_b[3], const float* d_w, size_t n) {
  ParticlesDev D{};
  for (int k = 0; k < 3; ++k) {
    RGN_REQUIRE(n == 0 || (d_u && d_u[k]), RGN_ERR_INVALID, "null momentum component");
    RGN_REQUIRE(n == 0 || (d_b && d_b[k]), RGN_ERR_INVALID, "null magnetic-field component");
    D.u[k] = d_u ? d_u[k] : nullptr;
    D.b[k] = d_b ? d_b[k] : nullptr;
    D.e[k] = d_e ? d_e[k] : nullptr;
  }
  D.w = d_w;
  return D;
}

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------
void build_sync_plan(SyncPlan& plan, const rgn_ftable* table, const float* h_bins, int nbins,
                     float B0, float g_syn, float e_at) {
  RGN_REQUIRE(table != nullptr, RGN_ERR_INVALID, "null F table");
  RGN_REQUIRE(nbins > 0, RGN_ERR_INVALID, "no photon bins");
  RGN_REQUIRE(h_bins != nullptr, RGN_ERR_INVALID, "null photon bins");
  const std::vector<float>& y = ftable_y(table);
  const int T = static_cast<int>(y.size());
  const double xmin = ftable_xmin(table), xmax = ftable_xmax(table);
  const double K = (T - 1) / std::log2(xmax / xmin);

  plan.nbins = nbins;
  plan.bins.assign(h_bins, h_bins + nbins);
  plan.B0 = B0;
  plan.g_syn = g_syn;
  plan.e_at = e_at;
  plan.K = K;
  plan.log2_xmin = std::log2(xmin);

  // Zero tail: Z - 1 and Z are zero nodes (or past the end of the table).
  int last = -1;
  for (int i = 0; i < T; ++i)
    if (y[i] != 0.0f) last = i;
  plan.Z = std::min(last + 2, T);

  std::vector<double> A(nbins, 0.0);
  std::vector<char> live(nbins, 0);
  double amin = std::numeric_limits<double>::infinity(), amax = -amin;
  for (int e = 0; e < nbins; ++e) {
    const double ee = h_bins[e];
    if (ee > 0.0 && std::isfinite(ee)) {
      live[e] = 1;
      A[e] = (std::log2(ee) - plan.log2_xmin) * K;
      amin = std::min(amin, A[e]);
      amax = std::max(amax, A[e]);
    }
  }
  if (!(amin <= amax)) {  // no live bin: every output is 0
    amin = amax = 0.0;
  }
  plan.Amin = amin;
  plan.Amax = amax;
  const double span = amax - amin;
  RGN_REQUIRE(plan.Z + span + 8 < kMaxTableEntries, RGN_ERR_RANGE,
              "photon bins span too many decades for the LDS table");
  plan.TP = plan.Z + static_cast<int>(std::ceil(span)) + 4;

  // Padded segment table in u = Z - t.  Segment j interpolates between nodes
  // a = Z - j - 1 and b = Z - j at their true coordinates: the float Logspace
  // nodes sit up to a few 1e-6 table units off the uniform grid, and the
  // reference interpolates between the float nodes (tabulation.hpp:39-41).
  const std::vector<float>& xs = ftable_x(table);
  auto node_u = [&](int i) { return plan.Z - (std::log2(static_cast<double>(xs[i])) - plan.log2_xmin) * K; };
  std::vector<float2> tab(plan.TP);
  for (int j = 0; j < plan.TP; ++j) {
    const int a = plan.Z - j - 1, b = plan.Z - j;
    if (a < 0 || b > T - 1) {
      tab[j] = make_float2(0.0f, 0.0f);
    } else {
      const double ya = y[a], yb = y[b];
      const double ua = node_u(a), ub = node_u(b);
      const double s = (ya - yb) / (ua - ub);
      tab[j] = make_float2(static_cast<float>(yb - ub * s), static_cast<float>(s));
    }
  }
  plan.table.reset(plan.TP * sizeof(float2));
  RGN_HIP(hipMemcpy(plan.table.ptr, tab.data(), plan.table.bytes, hipMemcpyHostToDevice));

  // Per-bin coordinates; dead bins and padding lanes reuse the lowest live A.
  plan.NB = static_cast<int>(ceil_div(nbins, 64));
  plan.nbins_pad = 64 * plan.NB;
  // the per-bin arrays also cover the adjoint's lane layouts (64, 128, 256 or
  // 512 bins, see rgn_sync_adjoint)
  const int adj_cover = nbins <= 64 ? 64 : nbins <= 128 ? 128 : nbins <= 256 ? 256 : 512;
  const int alloc = std::max(plan.nbins_pad, adj_cover);
  std::vector<double> U(alloc, plan.Z - amin);
  std::vector<float> bine(alloc, 0.0f);
  for (int e = 0; e < nbins; ++e) {
    if (live[e]) {
      U[e] = plan.Z - A[e];
      bine[e] = h_bins[e];
    }
  }
  DeviceBuffer dU(alloc * sizeof(double));
  RGN_HIP(hipMemcpy(dU.ptr, U.data(), dU.bytes, hipMemcpyHostToDevice));
  plan.bincs.reset(alloc * sizeof(float2));
  product_pair_kernel<<<alloc, 64>>>(dU.as<double>(), plan.bincs.as<float2>(), alloc);
  RGN_HIP(hipGetLastError());
  plan.bine.reset(alloc * sizeof(float));
  RGN_HIP(hipMemcpy(plan.bine.ptr, bine.data(), plan.bine.bytes, hipMemcpyHostToDevice));
  RGN_HIP(hipDeviceSynchronize());  // dU goes out of scope

  plan.win_lo = amin - plan.Z;
  plan.win_hi = amax;
  plan.v_safe = static_cast<float>(std::rint(amax));

  plan.T = T;
  plan.A = A;
  plan.live = live;
  build_binned_tables(plan, xs, y);
}

PlanDev device_view(const SyncPlan& plan) {
  PlanDev P{};
  P.table = plan.table.as<float2>();
  P.bincs = plan.bincs.as<float2>();
  P.bine = plan.bine.as<float>();
  P.TP = plan.TP;
  P.nbins = plan.nbins;
  P.nbins_pad = plan.nbins_pad;
  P.Z = plan.Z;
  P.em = emit_const(plan);
  P.K = plan.K;
  P.win_lo = plan.win_lo;
  P.win_hi = plan.win_hi;
  P.v_safe = plan.v_safe;
  return P;
}

size_t forward_smem(const SyncPlan& plan) {
  const size_t main = kThreads * sizeof(float4) + plan.TP * sizeof(float2);
  const size_t red = 4 * 64 * plan.NB * sizeof(double);
  return std::max(main, red);
}

int forward_grid(size_t n, int blocks_per_cu) {
  const size_t tiles = ceil_div(n, kThreads);
  const size_t cap = static_cast<size_t>(num_cus()) * blocks_per_cu;
  return static_cast<int>(std::max<size_t>(1, std::min(tiles, cap)));
}

template <class SourceFn>
void launch_forward(const SyncPlan& plan, SourceFn fn, size_t n, int grid, double* partials,
                    hipStream_t st, float* out = nullptr) {
  const PlanDev P = device_view(plan);
  const size_t smem = forward_smem(plan);
  switch (plan.NB) {
#define RGN_FWD_CASE(NB_)                                                          \
  case NB_:                                                                        \
    forward_kernel<NB_, SourceFn><<<grid, kThreads, smem, st>>>(fn, n, P, partials, out); \
    break;
    RGN_FWD_CASE(1)
    RGN_FWD_CASE(2)
    RGN_FWD_CASE(3)
    RGN_FWD_CASE(4)
    RGN_FWD_CASE(5)
    RGN_FWD_CASE(6)
    RGN_FWD_CASE(7)
    RGN_FWD_CASE(8)
#undef RGN_FWD_CASE
    default:
      throw Error(RGN_ERR_UNSUPPORTED, "at most 512 photon bins per plan");
  }
  RGN_HIP(hipGetLastError());
  stage_mark(st, "forward_kernel");
}

void launch_combine(const SyncPlan& plan, const double* partials, int grid, float* out,
                    hipStream_t st) {
  combine_kernel<<<plan.nbins, kThreads, 0, st>>>(partials, grid, plan.bine.as<float>(),
                                                  plan.nbins, out);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "combine_kernel");
}

// Automatic choice between the direct and the factorised kernels from a cost
// model measured on MI355X (2^16 .. 2^24 particles, 32 .. 512 bins; constants
// below).  RAGNAR_SYNC_METHOD (direct | binned) overrides the automatic choice.
int resolve_method(const SyncPlan& plan, size_t n, int method, bool adjoint) {
  // the factorised kernels address particles with 32-bit byte offsets
  const bool binned_fits = plan.binned_ok && n < kBinnedMaxParticles;
  if (method == RGN_METHOD_AUTO) {
    if (const char* m = std::getenv("RAGNAR_SYNC_METHOD")) {
      if (std::string(m) == "direct") return RGN_METHOD_DIRECT;
      if (std::string(m) == "binned" && binned_fits) return RGN_METHOD_BINNED;
    }
    if (!binned_fits) return RGN_METHOD_DIRECT;
    // Seconds per call, fitted on MI355X to 2^16 .. 2^24 particles at 32 .. 512
    // bins.  The pair-by-pair kernels cost per particle by NB, the number of
    // 64-bin blocks (seconds per particle below, by NB); the factorised ones
    // have a floor of fixed passes, then a cost per particle that hardly
    // depends on the bin count.
    static const double c_fwd[9] = {0, 20e-12, 26e-12, 31e-12, 36e-12, 50e-12, 65e-12, 80e-12, 94e-12};
    static const double c_adj[9] = {0, 27e-12, 34e-12, 43e-12, 52e-12, 63e-12, 74e-12, 84e-12, 95e-12};
    const double nd = static_cast<double>(n);
    const int nbk = std::min(std::max(plan.NB, 1), 8);
    const double t_direct = adjoint ? 25e-6 + nd * c_adj[nbk] : 18e-6 + nd * c_fwd[nbk];
    const double t_binned = adjoint ? std::max(33e-6, 28e-6 + nd * 13.5e-12)
                                    : std::max(62e-6, 35e-6 + nd * 10.1e-12);
    return t_binned < t_direct ? RGN_METHOD_BINNED : RGN_METHOD_DIRECT;
  }
  RGN_REQUIRE(method == RGN_METHOD_DIRECT || method == RGN_METHOD_BINNED, RGN_ERR_INVALID,
              "unknown method");
  RGN_REQUIRE(method != RGN_METHOD_BINNED || binned_fits, RGN_ERR_UNSUPPORTED,
              "the binned method needs fewer than 2^29 particles, at most 1024 distinct bin "
              "offsets and a bin range of at most 3072 table units");
  return method;
}

void sync_forward(const SyncPlan& plan, const ParticlesDev& D, size_t n, float* out,
                  hipStream_t st, int blocks_per_cu) {
  RGN_REQUIRE(out != nullptr, RGN_ERR_INVALID, "null output");
  const int grid = forward_grid(n, blocks_per_cu);
  double* partials =
    static_cast<double*>(scratch(sizeof(double) * plan.nbins_pad * static_cast<size_t>(grid), 0));
  if (grid == 1) {  // one block: it writes the spectrum, no combine launch
    launch_forward(plan, ParticleSourceFn{D}, n, grid, partials, st, out);
    return;
  }
  launch_forward(plan, ParticleSourceFn{D}, n, grid, partials, st);
  launch_combine(plan, partials, grid, out, st);
}

}  // namespace rgn

using namespace rgn;

extern "C" {

int rgn_sync_plan_create(const rgn_ftable* table, const float* h_bins, int nbins, float B0,
                         float g_syn, float e_at, rgn_sync_plan** out) {
  return guarded([&] {
    RGN_REQUIRE(out != nullptr, RGN_ERR_INVALID, "null output handle");
    RGN_REQUIRE(nbins <= 512, RGN_ERR_UNSUPPORTED, "at most 512 photon bins per plan");
    auto* plan = new rgn_sync_plan();
    try {
      build_sync_plan(plan->p, table, h_bins, nbins, B0, g_syn, e_at);
    } catch (...) {
      delete plan;
      throw;
    }
    *out = plan;
  });
}

int rgn_sync_plan_destroy(rgn_sync_plan* plan) {
  if (plan) forget_plan_graphs(&plan->p);
  delete plan;
  return RGN_OK;
}

int rgn_sync_spectrum_method(const rgn_sync_plan* plan, const float* const d_u[3],
                             const float* const d_e[3], const float* const d_b[3],
                             const float* d_w, size_t n, int method, float* d_out,
                             rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    RGN_REQUIRE(d_out != nullptr, RGN_ERR_INVALID, "null output");
    const ParticlesDev D = make_particles_checked(d_u, d_e, d_b, d_w, n);
    ScratchUse order(as_stream(stream));
    if (resolve_method(plan->p, n, method, false) == RGN_METHOD_BINNED) {
      sync_forward_binned(plan->p, D, n, d_out, as_stream(stream));
    } else {
      // 16 blocks per CU measured fastest on MI355X (2^24 x 256: 0.68 ms vs 0.74 at 4)
      sync_forward(plan->p, D, n, d_out, as_stream(stream), env_int("RAGNAR_FWD_BPC", 16));
    }
  });
}

int rgn_sync_spectrum(const rgn_sync_plan* plan, const float* const d_u[3],
                      const float* const d_e[3], const float* const d_b[3], const float* d_w,
                      size_t n, float* d_out, rgn_stream_t stream) {
  return rgn_sync_spectrum_method(plan, d_u, d_e, d_b, d_w, n, RGN_METHOD_AUTO, d_out, stream);
}

int rgn_sync_adjoint(const rgn_sync_plan* plan, const float* const d_u[3],
                     const float* const d_e[3], const float* const d_b[3], const float* d_w,
                     size_t n, const float* d_gbar, float* d_gw, float* const d_gu[3],
                     float* const d_gb[3], float* const d_ge[3], rgn_stream_t stream) {
  return rgn_sync_adjoint_method(plan, d_u, d_e, d_b, d_w, n, d_gbar, RGN_METHOD_AUTO, d_gw,
                                 d_gu, d_gb, d_ge, stream);
}

int rgn_sync_adjoint_method(const rgn_sync_plan* plan, const float* const d_u[3],
                            const float* const d_e[3], const float* const d_b[3],
                            const float* d_w, size_t n, const float* d_gbar, int method,
                            float* d_gw, float* const d_gu[3], float* const d_gb[3],
                            float* const d_ge[3], rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    RGN_REQUIRE(d_gbar != nullptr, RGN_ERR_INVALID, "null upstream gradient");
    const ParticlesDev D = make_particles_checked(d_u, d_e, d_b, d_w, n);
    ScratchUse order(as_stream(stream));
    if (resolve_method(plan->p, n, method, true) == RGN_METHOD_BINNED) {
      sync_adjoint_binned(plan->p, D, n, d_gbar, d_gw, d_gu, d_gb, d_ge, as_stream(stream));
      return;
    }
    GradOut G{};
    G.gw = d_gw;
    for (int k = 0; k < 3; ++k) {
      G.gu[k] = d_gu ? d_gu[k] : nullptr;
      G.gb[k] = d_gb ? d_gb[k] : nullptr;
      G.ge[k] = d_ge ? d_ge[k] : nullptr;
    }
    if (n == 0) return;
    const SyncPlan& p = plan->p;
    const PlanDev P = device_view(p);
    const int grid = forward_grid(n, env_int("RAGNAR_ADJ_BPC", 16));
    const size_t smem = kThreads * (sizeof(float4) + sizeof(float2)) + p.TP * sizeof(float2);
    hipStream_t st = as_stream(stream);
    // (lanes per source, bins per lane) covering nbins; the plan pads its
    // per-bin arrays to cover every configuration (adjoint_coverage).
    if (p.nbins <= 64) {
      adjoint_kernel<16, 4><<<grid, kThreads, smem, st>>>(D, n, P, d_gbar, G);
    } else if (p.nbins <= 128) {
      adjoint_kernel<16, 8><<<grid, kThreads, smem, st>>>(D, n, P, d_gbar, G);
    } else if (p.nbins <= 256) {
      adjoint_kernel<16, 16><<<grid, kThreads, smem, st>>>(D, n, P, d_gbar, G);
    } else {
      adjoint_kernel<32, 16><<<grid, kThreads, smem, st>>>(D, n, P, d_gbar, G);
    }
    RGN_HIP(hipGetLastError());
    stage_mark(st, "adjoint_kernel");
  });
}

namespace {
// Rethrows the status of a nested C-ABI call.
void rethrow(int status) {
  if (status != RGN_OK) throw Error(status, rgn_last_error());
}
}  // namespace

int rgn_sync_spectrum_stages(const rgn_sync_plan* plan, const float* const d_u[3],
                             const float* const d_e[3], const float* const d_b[3],
                             const float* d_w, size_t n, int method, float* d_out,
                             rgn_stream_t stream, float* stage_ms, const char** names,
                             int max_stages, int* nstages) {
  return guarded([&] {
    RGN_REQUIRE(stage_ms != nullptr && nstages != nullptr, RGN_ERR_INVALID, "null timing output");
    StageScope scope(as_stream(stream));
    rethrow(rgn_sync_spectrum_method(plan, d_u, d_e, d_b, d_w, n, method, d_out, stream));
    *nstages = scope.finish(stage_ms, names, max_stages);
  });
}

int rgn_sync_adjoint_stages(const rgn_sync_plan* plan, const float* const d_u[3],
                            const float* const d_e[3], const float* const d_b[3],
                            const float* d_w, size_t n, const float* d_gbar, int method,
                            float* d_gw, float* const d_gu[3], float* const d_gb[3],
                            float* const d_ge[3], rgn_stream_t stream, float* stage_ms,
                            const char** names, int max_stages, int* nstages) {
  return guarded([&] {
    RGN_REQUIRE(stage_ms != nullptr && nstages != nullptr, RGN_ERR_INVALID, "null timing output");
    StageScope scope(as_stream(stream));
    rethrow(rgn_sync_adjoint_method(plan, d_u, d_e, d_b, d_w, n, d_gbar, method, d_gw, d_gu, d_gb,
                                    d_ge, stream));
    *nstages = scope.finish(stage_ms, names, max_stages);
  });
}

int rgn_sync_fromdist(const rgn_sync_plan* plan, const float* d_gbeta, const float* d_f, int ng,
                      int dist_log, float* d_out, rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    RGN_REQUIRE(d_gbeta != nullptr && d_f != nullptr && d_out != nullptr, RGN_ERR_INVALID,
                "null argument");
    RGN_REQUIRE(ng >= 0, RGN_ERR_INVALID, "negative distribution size");
    ScratchUse order(as_stream(stream));
    const SyncPlan& p = plan->p;
    const int grid = forward_grid(static_cast<size_t>(ng), 4);
    double* partials =
      static_cast<double*>(scratch(sizeof(double) * p.nbins_pad * static_cast<size_t>(grid), 1));
    hipStream_t st = as_stream(stream);
    if (grid == 1) {  // one block: it writes the spectrum, no combine launch
      launch_forward(p, DistSourceFn{d_gbeta, d_f, dist_log}, static_cast<size_t>(ng), grid, partials, st,
                     d_out);
      return;
    }
    launch_forward(p, DistSourceFn{d_gbeta, d_f, dist_log}, static_cast<size_t>(ng), grid,
                   partials, st);
    launch_combine(p, partials, grid, d_out, st);
  });
}

// Diagnostics: out[5 * n] = (e_peak, chiR, W, V_hi, V_lo) per particle.
int rgn_sync_particle_terms(const rgn_sync_plan* plan, const float* const d_u[3],
                            const float* const d_e[3], const float* const d_b[3],
                            const float* d_w, size_t n, float* d_out, rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    const ParticlesDev D = make_particles_checked(d_u, d_e, d_b, d_w, n);
    if (!n) return;
    const unsigned grid = static_cast<unsigned>(ceil_div(n, kThreads));
    if (D.e[0] || D.e[1] || D.e[2]) {
      particle_terms_kernel<true><<<grid, kThreads, 0, as_stream(stream)>>>(D, n, device_view(plan->p), d_out);
    } else {
      particle_terms_kernel<false><<<grid, kThreads, 0, as_stream(stream)>>>(D, n, device_view(plan->p), d_out);
    }
    RGN_HIP(hipGetLastError());
  });
}

// Plan scalars for tests: Z, TP, NB, K, Amin, Amax.
int rgn_sync_plan_info(const rgn_sync_plan* plan, int* ints6, double* dbl3) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    if (ints6) {
      ints6[0] = plan->p.Z;
      ints6[1] = plan->p.TP;
      ints6[2] = plan->p.NB;
      ints6[3] = plan->p.M;
      ints6[4] = plan->p.NK;
      ints6[5] = plan->p.binned_ok ? 1 : 0;
    }
    if (dbl3) {
      dbl3[0] = plan->p.K;
      dbl3[1] = plan->p.Amin;
      dbl3[2] = plan->p.Amax;
    }
  });
}

}  // extern "C"
