// Factorised ("binned") form of the per-particle spectrum and its adjoint.
//
// The interpolant F is piecewise linear in t = A_e - B (table units), so the
// N x B sum factors exactly.  Write B = k + phi (k integer, phi in [0, 1)) and
// A_e = a_e + alpha_e.  Then t lies on segment m = a_e - k when phi <= alpha_e
// and on segment m - 1 otherwise, and on segment s
//     F = C_s(e, k) - sigma_s * phi,   C_s = y_s + (A_e - k - tau_s) sigma_s.
// The sorted distinct alpha_e cut [0, 1) into M + 1 sub-cells; a source in
// sub-cell j lies left of alpha_e exactly when j <= r_e (the rank of alpha_e).
// So every bin needs, per k, only the prefix sums over sub-cells of
//     m0 = sum W   and   m1 = sum W phi,
// and the spectrum is
//     S_e = e_e sum_k [ C_m P0 - sigma_m P1 + C_{m-1} (T0 - P0) - sigma_{m-1} (T1 - P1) ]
// with P = prefix at r_e and T = totals.  This is the same piecewise-linear
// function the direct kernel evaluates pair by pair, summed in FP64.
//
// Forward (a scatter): the sources are partitioned by k with a counting sort
// (count, scan, scatter) and each bucket is reduced into its M + 1 sub-cells in
// LDS; no global atomics (device-scope FP64 atomics run at ~14 G/s on MI355X).
// Adjoint (a gather): dL/dS is folded into a table U[k][j] = (sum_e g'_e C,
// sum_e g'_e sigma) and every particle reads one entry.
#include <algorithm>
#include <cmath>
#include <vector>

#include "emission.hpp"
#include "rgn_core.hpp"

namespace rgn {

void build_binned_tables(SyncPlan& plan, const std::vector<float>& xs, const std::vector<float>& ys) {
  const int T = plan.T;
  std::vector<double> alphas;
  for (int e = 0; e < plan.nbins; ++e) {
    if (plan.live[e]) alphas.push_back(plan.A[e] - std::floor(plan.A[e]));
  }
  std::sort(alphas.begin(), alphas.end());
  alphas.erase(std::unique(alphas.begin(), alphas.end()), alphas.end());
  plan.M = static_cast<int>(alphas.size());
  std::vector<BinInfo> bi(plan.nbins);
  for (int e = 0; e < plan.nbins; ++e) {
    if (!plan.live[e]) {
      bi[e] = BinInfo{0.0, 0, -1};
      continue;
    }
    const double a = std::floor(plan.A[e]);
    const double al = plan.A[e] - a;
    const int r = static_cast<int>(std::lower_bound(alphas.begin(), alphas.end(), al) - alphas.begin());
    bi[e] = BinInfo{plan.A[e], static_cast<int>(a), r};
  }
  std::vector<SegInfo> seg(std::max(T - 1, 1));
  auto tau = [&](int i) { return (std::log2(static_cast<double>(xs[i])) - plan.log2_xmin) * plan.K; };
  for (int s = 0; s + 1 < T; ++s) {
    const double t0 = tau(s), t1 = tau(s + 1);
    seg[s] = SegInfo{t0, ys[s], (static_cast<double>(ys[s + 1]) - ys[s]) / (t1 - t0)};
  }
  if (plan.M > 0) {
    // sources can contribute only for B in (A_min - tau_{T-1}, A_max)
    plan.kmin = static_cast<int>(std::floor(plan.Amin - (T - 1))) - 1;
    const int kmax = static_cast<int>(std::floor(plan.Amax)) + 1;
    plan.NK = kmax - plan.kmin + 1;
  } else {
    plan.kmin = 0;
    plan.NK = 0;
  }
  plan.alpha.reset(std::max<size_t>(1, alphas.size()) * sizeof(double));
  if (!alphas.empty()) {
    RGN_HIP(hipMemcpy(plan.alpha.ptr, alphas.data(), alphas.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  plan.binfo.reset(bi.size() * sizeof(BinInfo));
  RGN_HIP(hipMemcpy(plan.binfo.ptr, bi.data(), bi.size() * sizeof(BinInfo), hipMemcpyHostToDevice));
  plan.seg.reset(seg.size() * sizeof(SegInfo));
  RGN_HIP(hipMemcpy(plan.seg.ptr, seg.data(), seg.size() * sizeof(SegInfo), hipMemcpyHostToDevice));
}

namespace {

constexpr int kThreads = 256;
constexpr int kItem = 4096;  // sources per sub-histogram work item

struct BinnedDev {
  const double* alpha;
  const BinInfo* binfo;
  const SegInfo* seg;
  const float* bine;
  int M, T, kmin, NK, nbins;
  float B0, g_syn, e_at;
  double K;
};

BinnedDev binned_view(const SyncPlan& p) {
  return BinnedDev{p.alpha.as<double>(), p.binfo.as<BinInfo>(), p.seg.as<SegInfo>(), p.bine.as<float>(),
                   p.M, p.T, p.kmin, p.NK, p.nbins, p.B0, p.g_syn, p.e_at, p.K};
}

// W and B of a particle; false when it contributes nowhere.
__device__ __forceinline__ bool particle_wb(const RawParticle& r, const BinnedDev& P, float& W, double& B) {
  float e_peak, chiR;
  omega_sync(r.ux, r.uy, r.uz, r.ex, r.ey, r.ez, r.bx, r.by, r.bz, P.B0, P.g_syn, P.e_at, e_peak, chiR);
  W = chiR * r.w;
  if (!(e_peak > 0.0f)) return false;
  B = log2(static_cast<double>(e_peak)) * P.K;
  return B >= static_cast<double>(P.kmin) && B < static_cast<double>(P.kmin + P.NK);
}

// Number of alphas strictly below phi (binary search in LDS).
__device__ __forceinline__ int subcell(const double* __restrict__ alpha, int M, double phi) {
  int lo = 0, hi = M;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (alpha[mid] < phi) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Pass 1: per particle (W, phi, k) records and per-block bucket counts.
__global__ void __launch_bounds__(kThreads) binned_count(ParticlesDev D, size_t n, size_t Q, BinnedDev P,
                                                         float* __restrict__ recW, float* __restrict__ recPhi,
                                                         int* __restrict__ recK, unsigned* __restrict__ counts) {
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned* cnt = reinterpret_cast<unsigned*>(smem);
  for (int i = threadIdx.x; i < P.NK; i += kThreads) cnt[i] = 0u;
  __syncthreads();
  const size_t p0 = static_cast<size_t>(blockIdx.x) * Q;
  const size_t p1 = min(n, p0 + Q);
  size_t p = p0 + threadIdx.x;
  RawParticle raw{};
  if (p < p1) raw = load_particle(D, p);
  for (; p < p1; p += kThreads) {
    const RawParticle cur = raw;
    if (p + kThreads < p1) raw = load_particle(D, p + kThreads);
    float W;
    double B;
    int k = -1;
    float phi = 0.0f;
    if (particle_wb(cur, P, W, B)) {
      const double fl = floor(B);
      k = static_cast<int>(fl) - P.kmin;
      phi = fminf(static_cast<float>(B - fl), 0x1.fffffep-1f);
      atomicAdd(cnt + k, 1u);
    }
    recW[p] = W;
    recPhi[p] = phi;
    recK[p] = k;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P.NK; i += kThreads) counts[static_cast<size_t>(i) * gridDim.x + blockIdx.x] = cnt[i];
}

// Pass 2a: exclusive scan of each bucket's per-block counts (one block per bucket).
__global__ void __launch_bounds__(kThreads) binned_scan_rows(unsigned* __restrict__ counts, int G,
                                                             unsigned* __restrict__ totals) {
  const int k = blockIdx.x;
  unsigned* row = counts + static_cast<size_t>(k) * G;
  __shared__ unsigned part[kThreads];
  // each thread owns a contiguous run of the row
  const int per = (G + kThreads - 1) / kThreads;
  const int b0 = threadIdx.x * per, b1 = min(G, b0 + per);
  unsigned s = 0;
  for (int b = b0; b < b1; ++b) s += row[b];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned acc = 0;
    for (int i = 0; i < kThreads; ++i) {
      const unsigned v = part[i];
      part[i] = acc;
      acc += v;
    }
    totals[k] = acc;
  }
  __syncthreads();
  unsigned acc = part[threadIdx.x];
  for (int b = b0; b < b1; ++b) {
    const unsigned v = row[b];
    row[b] = acc;
    acc += v;
  }
}

// Pass 2b: bucket starts and sub-histogram work items (one block).
__global__ void __launch_bounds__(kThreads) binned_scan_totals(const unsigned* __restrict__ totals, int NK,
                                                               unsigned* __restrict__ bstart,
                                                               unsigned* __restrict__ istart) {
  if (threadIdx.x == 0) {
    unsigned s = 0, it = 0;
    for (int k = 0; k < NK; ++k) {
      bstart[k] = s;
      istart[k] = it;
      s += totals[k];
      it += (totals[k] + kItem - 1) / kItem;
    }
    bstart[NK] = s;
    istart[NK] = it;
  }
}

// Pass 3: scatter (W, phi) into bucket order.
__global__ void __launch_bounds__(kThreads) binned_scatter(size_t n, size_t Q, int NK,
                                                           const float* __restrict__ recW,
                                                           const float* __restrict__ recPhi,
                                                           const int* __restrict__ recK,
                                                           const unsigned* __restrict__ counts,
                                                           const unsigned* __restrict__ bstart,
                                                           float* __restrict__ sortW, float* __restrict__ sortPhi) {
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned* cur = reinterpret_cast<unsigned*>(smem);
  for (int i = threadIdx.x; i < NK; i += kThreads) {
    cur[i] = bstart[i] + counts[static_cast<size_t>(i) * gridDim.x + blockIdx.x];
  }
  __syncthreads();
  const size_t p0 = static_cast<size_t>(blockIdx.x) * Q;
  const size_t p1 = min(n, p0 + Q);
  for (size_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
    const int k = recK[p];
    if (k < 0) continue;
    const unsigned pos = atomicAdd(cur + k, 1u);
    sortW[pos] = recW[p];
    sortPhi[pos] = recPhi[p];
  }
}

// Pass 4: sub-histogram of one work item (a chunk of one bucket).
__global__ void __launch_bounds__(kThreads) binned_subhist(BinnedDev P, const unsigned* __restrict__ bstart,
                                                           const unsigned* __restrict__ istart,
                                                           const float* __restrict__ sortW,
                                                           const float* __restrict__ sortPhi,
                                                           double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int M1 = P.M + 1;
  double* m0 = reinterpret_cast<double*>(smem);
  double* m1 = m0 + M1;
  double* alpha = m1 + M1;
  const unsigned it = blockIdx.x;
  if (it >= istart[P.NK]) return;  // uniform per block
  for (int i = threadIdx.x; i < M1; i += kThreads) {
    m0[i] = 0.0;
    m1[i] = 0.0;
  }
  for (int i = threadIdx.x; i < P.M; i += kThreads) alpha[i] = P.alpha[i];
  int lo = 0, hi = P.NK - 1;  // bucket of this item: last k with istart[k] <= it
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (istart[mid] <= it) lo = mid; else hi = mid - 1;
  }
  const int k = lo;
  const unsigned s0 = bstart[k] + (it - istart[k]) * kItem;
  const unsigned s1 = min(bstart[k + 1], s0 + kItem);
  __syncthreads();
  for (unsigned s = s0 + threadIdx.x; s < s1; s += kThreads) {
    const double w = sortW[s];
    const double phi = sortPhi[s];
    const int j = subcell(alpha, P.M, phi);
    atomicAdd(m0 + j, w);
    atomicAdd(m1 + j, w * phi);
  }
  __syncthreads();
  double* out = part + static_cast<size_t>(it) * 2 * M1;
  for (int i = threadIdx.x; i < M1; i += kThreads) {
    out[i] = m0[i];
    out[M1 + i] = m1[i];
  }
}

// Pass 5: per bucket, sum its items in order and take prefix sums over j.
__global__ void __launch_bounds__(kThreads) binned_prefix(int M, const unsigned* __restrict__ istart,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ P0, double* __restrict__ P1) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int M1 = M + 1;
  double* s0 = reinterpret_cast<double*>(smem);
  double* s1 = s0 + M1;
  const int k = blockIdx.x;
  for (int j = threadIdx.x; j < M1; j += kThreads) {
    double a = 0.0, b = 0.0;
    for (unsigned it = istart[k]; it < istart[k + 1]; ++it) {
      a += part[static_cast<size_t>(it) * 2 * M1 + j];
      b += part[static_cast<size_t>(it) * 2 * M1 + M1 + j];
    }
    s0[j] = a;
    s1[j] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int j = 0; j < M1; ++j) {
      a += s0[j];
      b += s1[j];
      s0[j] = a;
      s1[j] = b;
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < M1; j += kThreads) {
    P0[static_cast<size_t>(k) * M1 + j] = s0[j];
    P1[static_cast<size_t>(k) * M1 + j] = s1[j];
  }
}

__device__ __forceinline__ void seg_line(const BinnedDev& P, const BinInfo& b, int k_abs, int s, double& C,
                                         double& sig) {
  if (s < 0 || s > P.T - 2) {
    C = 0.0;
    sig = 0.0;
    return;
  }
  const SegInfo g = P.seg[s];
  sig = g.sigma;
  C = g.y + (b.A - k_abs - g.tau) * g.sigma;
}

// Pass 6: S_e = e_e sum_k (...), one block per bin, fixed-order reduction.
__global__ void __launch_bounds__(kThreads) binned_evaluate(BinnedDev P, const double* __restrict__ P0,
                                                            const double* __restrict__ P1,
                                                            float* __restrict__ out) {
  const int e = blockIdx.x;
  const BinInfo b = P.binfo[e];
  __shared__ double red[kThreads];
  const int M1 = P.M + 1;
  double acc = 0.0;
  if (b.r >= 0) {
    const int klo = max(P.kmin, b.a - (P.T - 1)), khi = min(P.kmin + P.NK - 1, b.a);
    for (int k = klo + static_cast<int>(threadIdx.x); k <= khi; k += kThreads) {
      const int kk = k - P.kmin;
      const int m = b.a - k;
      double Cp, sp, Cm, sm;
      seg_line(P, b, k, m, Cp, sp);
      seg_line(P, b, k, m - 1, Cm, sm);
      const double* r0 = P0 + static_cast<size_t>(kk) * M1;
      const double* r1 = P1 + static_cast<size_t>(kk) * M1;
      const double p0 = r0[b.r], p1 = r1[b.r], t0 = r0[P.M], t1 = r1[P.M];
      acc += (Cp * p0 - sp * p1) + (Cm * (t0 - p0) - sm * (t1 - p1));
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float ee = P.bine[e];
    out[e] = (b.r < 0 || ee == 0.0f) ? 0.0f : static_cast<float>(static_cast<double>(ee) * red[0]);
  }
}

// ---------------------------------------------------------------------------
// Adjoint: U[k][j] = (sum_e g'_e C_s, sum_e g'_e sigma_s), s = m or m - 1.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) binned_adj_table(BinnedDev P, const float* __restrict__ gbar,
                                                             double2* __restrict__ U) {
  extern __shared__ __align__(16) unsigned char smem[];
  BinInfo* bi = reinterpret_cast<BinInfo*>(smem);
  double* gp = reinterpret_cast<double*>(bi + P.nbins);
  for (int e = threadIdx.x; e < P.nbins; e += kThreads) {
    bi[e] = P.binfo[e];
    gp[e] = static_cast<double>(gbar[e]) * P.bine[e];
  }
  __syncthreads();
  const int M1 = P.M + 1;
  const int kk = blockIdx.x;
  const int k = P.kmin + kk;
  for (int j = threadIdx.x; j < M1; j += kThreads) {
    double u0 = 0.0, u1 = 0.0;
    for (int e = 0; e < P.nbins; ++e) {
      const BinInfo b = bi[e];
      if (b.r < 0 || gp[e] == 0.0) continue;
      const int s = (j <= b.r) ? b.a - k : b.a - k - 1;
      double C, sig;
      seg_line(P, b, k, s, C, sig);
      u0 += gp[e] * C;
      u1 += gp[e] * sig;
    }
    U[static_cast<size_t>(kk) * M1 + j] = make_double2(u0, u1);
  }
}

struct GradPtrs {
  float* gw;
  float* gu[3];
  float* gb[3];
  float* ge[3];
};

__global__ void __launch_bounds__(kThreads) binned_adj_gather(ParticlesDev D, size_t n, BinnedDev P,
                                                              const double2* __restrict__ U, GradPtrs G) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* alpha = reinterpret_cast<double*>(smem);
  for (int i = threadIdx.x; i < P.M; i += kThreads) alpha[i] = P.alpha[i];
  __syncthreads();
  const double dlnx = P.K / 0.69314718055994530942;  // dt / dln x
  const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
  size_t p = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
  RawParticle raw{};
  if (p < n) raw = load_particle(D, p);
  for (; p < n; p += stride) {
    const RawParticle cur = raw;
    if (p + stride < n) raw = load_particle(D, p + stride);
    float W;
    double B;
    double gw = 0, gux = 0, guy = 0, guz = 0, gbx = 0, gby = 0, gbz = 0, gex = 0, gey = 0, gez = 0;
    if (particle_wb(cur, P, W, B)) {
      const double fl = floor(B);
      const double phi = B - fl;
      const int kk = static_cast<int>(fl) - P.kmin;
      const int j = subcell(alpha, P.M, phi);
      const double2 u = U[static_cast<size_t>(kk) * (P.M + 1) + j];
      const double A = u.x - u.y * phi;  // sum_e g'_e F
      const double C = dlnx * u.y;        // sum_e g'_e dF/dln x
      const double ux = cur.ux, uy = cur.uy, uz = cur.uz, ex = cur.ex, ey = cur.ey, ez = cur.ez;
      const double bx = cur.bx, by = cur.by, bz = cur.bz, w = cur.w;
      const double gamma = sqrt(1.0 + ux * ux + uy * uy + uz * uz);
      const double btx = ux / gamma, bty = uy / gamma, btz = uz / gamma;
      const double bde = btx * ex + bty * ey + btz * ez;
      const double vx = ex + (bty * bz - btz * by);
      const double vy = ey + (btz * bx - btx * bz);
      const double vz = ez + (btx * by - bty * bx);
      const double rootQ = sqrt(vx * vx + vy * vy + vz * vz - bde * bde);
      const double chi = rootQ / P.B0;
      gw = chi * A;
      const double gchi = w * (A - C);
      const double ggamma = -2.0 * w * chi * C / gamma;
      const double coef = gchi / (static_cast<double>(P.B0) * rootQ);
      gbx = coef * (vy * btz - vz * bty);
      gby = coef * (vz * btx - vx * btz);
      gbz = coef * (vx * bty - vy * btx);
      gex = coef * (vx - bde * btx);
      gey = coef * (vy - bde * bty);
      gez = coef * (vz - bde * btz);
      const double gbtx = coef * ((by * vz - bz * vy) - bde * ex);
      const double gbty = coef * ((bz * vx - bx * vz) - bde * ey);
      const double gbtz = coef * ((bx * vy - by * vx) - bde * ez);
      const double udg = ux * gbtx + uy * gbty + uz * gbtz;
      const double g3 = gamma * gamma * gamma;
      gux = gbtx / gamma - ux * udg / g3 + ggamma * ux / gamma;
      guy = gbty / gamma - uy * udg / g3 + ggamma * uy / gamma;
      guz = gbtz / gamma - uz * udg / g3 + ggamma * uz / gamma;
    }
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

// Scratch slots 6..: records, counts, sorted sources, partials, tables.
struct BinnedScratch {
  float* recW;
  float* recPhi;
  int* recK;
  unsigned* counts;
  unsigned* totals;
  unsigned* bstart;
  unsigned* istart;
  float* sortW;
  float* sortPhi;
  double* part;
  double* P0;
  double* P1;
};

}  // namespace

void sync_forward_binned(const SyncPlan& plan, const ParticlesDev& D, size_t n, float* out, hipStream_t st) {
  const BinnedDev P = binned_view(plan);
  if (plan.NK == 0 || n == 0) {
    RGN_HIP(hipMemsetAsync(out, 0, sizeof(float) * plan.nbins, st));
    return;
  }
  const int G = static_cast<int>(std::max<size_t>(1, std::min<size_t>(ceil_div(n, kThreads),
                                                                         static_cast<size_t>(num_cus()) * 8)));
  const size_t Q = ceil_div(n, G);
  const size_t M1 = plan.M + 1;
  const size_t max_items = n / kItem + plan.NK + 1;
  BinnedScratch S{};
  S.recW = static_cast<float*>(scratch(n * sizeof(float), 6));
  S.recPhi = static_cast<float*>(scratch(n * sizeof(float), 7));
  S.recK = static_cast<int*>(scratch(n * sizeof(int), 8));
  S.counts = static_cast<unsigned*>(scratch(sizeof(unsigned) * plan.NK * static_cast<size_t>(G), 9));
  unsigned* small = static_cast<unsigned*>(scratch(sizeof(unsigned) * (3 * (plan.NK + 1) + 64), 10));
  S.totals = small;
  S.bstart = small + (plan.NK + 1);
  S.istart = small + 2 * (plan.NK + 1);
  S.sortW = static_cast<float*>(scratch(n * sizeof(float), 11));
  S.sortPhi = static_cast<float*>(scratch(n * sizeof(float), 12));
  S.part = static_cast<double*>(scratch(sizeof(double) * 2 * M1 * max_items, 13));
  double* pref = static_cast<double*>(scratch(sizeof(double) * 2 * M1 * plan.NK, 14));
  S.P0 = pref;
  S.P1 = pref + M1 * plan.NK;

  const size_t cnt_smem = sizeof(unsigned) * plan.NK;
  RGN_REQUIRE(cnt_smem <= 64 * 1024, RGN_ERR_RANGE, "photon bins span too many decades for the binned form");
  binned_count<<<G, kThreads, cnt_smem, st>>>(D, n, Q, P, S.recW, S.recPhi, S.recK, S.counts);
  RGN_HIP(hipGetLastError());
  binned_scan_rows<<<plan.NK, kThreads, 0, st>>>(S.counts, G, S.totals);
  RGN_HIP(hipGetLastError());
  binned_scan_totals<<<1, kThreads, 0, st>>>(S.totals, plan.NK, S.bstart, S.istart);
  RGN_HIP(hipGetLastError());
  binned_scatter<<<G, kThreads, cnt_smem, st>>>(n, Q, plan.NK, S.recW, S.recPhi, S.recK, S.counts, S.bstart,
                                               S.sortW, S.sortPhi);
  RGN_HIP(hipGetLastError());
  const size_t sub_smem = sizeof(double) * (2 * M1 + plan.M);
  binned_subhist<<<static_cast<unsigned>(max_items), kThreads, sub_smem, st>>>(P, S.bstart, S.istart, S.sortW,
                                                                             S.sortPhi, S.part);
  RGN_HIP(hipGetLastError());
  binned_prefix<<<plan.NK, kThreads, sizeof(double) * 2 * M1, st>>>(plan.M, S.istart, S.part, S.P0, S.P1);
  RGN_HIP(hipGetLastError());
  binned_evaluate<<<plan.nbins, kThreads, 0, st>>>(P, S.P0, S.P1, out);
  RGN_HIP(hipGetLastError());
}

void sync_adjoint_binned(const SyncPlan& plan, const ParticlesDev& D, size_t n, const float* gbar, float* gw,
                         float* const gu[3], float* const gb[3], float* const ge[3], hipStream_t st) {
  const BinnedDev P = binned_view(plan);
  GradPtrs G{};
  G.gw = gw;
  for (int k = 0; k < 3; ++k) {
    G.gu[k] = gu ? gu[k] : nullptr;
    G.gb[k] = gb ? gb[k] : nullptr;
    G.ge[k] = ge ? ge[k] : nullptr;
  }
  if (n == 0) return;
  const size_t M1 = plan.M + 1;
  double2* U = static_cast<double2*>(scratch(sizeof(double2) * M1 * std::max(plan.NK, 1), 15));
  if (plan.NK > 0) {
    const size_t tab_smem = plan.nbins * (sizeof(BinInfo) + sizeof(double));
    binned_adj_table<<<plan.NK, kThreads, tab_smem, st>>>(P, gbar, U);
    RGN_HIP(hipGetLastError());
  }
  BinnedDev Pg = P;
  if (plan.NK == 0) Pg.NK = 0;  // every particle falls outside: gradients are zero
  const int grid = static_cast<int>(std::max<size_t>(1, std::min<size_t>(ceil_div(n, kThreads),
                                                                            static_cast<size_t>(num_cus()) * 8)));
  binned_adj_gather<<<grid, kThreads, sizeof(double) * std::max(plan.M, 1), st>>>(D, n, Pg, U, G);
  RGN_HIP(hipGetLastError());
}

}  // namespace rgn
