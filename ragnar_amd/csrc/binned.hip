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
// Forward (a scatter), no global atomics (device-scope FP64 atomics run at
// ~14 G/s on MI355X):
//   bucket    each block of 8192 particles computes (k, W, phi), counting-sorts
//             its sources by k in LDS and writes the sorted (W, phi) chunk
//             coalesced, plus one run descriptor per non-empty bucket;
//   scan      per bucket, its non-empty runs compacted in block order with
//             their offsets within the bucket;
//   subhist   each bucket is cut into work items of up to `item` sources (at
//             least a bucket block's chunk; sized so that the items run as one
//             round of resident blocks, or RAGNAR_SUBHIST_ITEM); each item
//             walks its runs and reduces the sources into M + 1 sub-cells in
//             LDS (ds_add_f64);
//   prefix    per bucket: the items in order, then prefix sums over sub-cells;
//   evaluate  one block per photon bin, fixed-order reduction over k.
// Adjoint (a gather): dL/dS is folded into a table U[k][j] = (sum_e g'_e C,
// sum_e g'_e sigma), built per k as a base plus prefix sums of per-rank
// switches from segment m to m - 1, and every particle reads one entry.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "emission.hpp"
#include "rgn_core.hpp"

namespace rgn {

namespace {
constexpr int kLutCells = 2048;  // sub-cell lookup cells in a sub-histogram block
constexpr int kLog2Tab = 256;    // entries of the factorised passes' log2 table
}  // namespace

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
  std::vector<int> rcount(plan.M + 1, 0);
  for (int e = 0; e < plan.nbins; ++e) {
    if (!plan.live[e]) {
      bi[e] = BinInfo{0.0, 0, -1};
      continue;
    }
    const double a = std::floor(plan.A[e]);
    const double al = plan.A[e] - a;
    const int r = static_cast<int>(std::lower_bound(alphas.begin(), alphas.end(), al) - alphas.begin());
    bi[e] = BinInfo{plan.A[e], static_cast<int>(a), r};
    ++rcount[r];
  }
  // live bins grouped by alpha rank, ascending e within a rank
  std::vector<int> rptr(plan.M + 1, 0), rlist;
  for (int r = 0; r < plan.M; ++r) rptr[r + 1] = rptr[r] + rcount[r];
  rlist.resize(std::max(1, rptr[plan.M]));
  {
    std::vector<int> cur(rptr.begin(), rptr.end() - 1);
    for (int e = 0; e < plan.nbins; ++e) {
      if (bi[e].r >= 0) rlist[cur[bi[e].r]++] = e;
    }
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
  // sub-cell lookup grid: jlut[q] = number of alphas below q / JQ (about 8 grid
  // cells per alpha, so the final scan is short)
  int JQ = 64;
  while (JQ < 8 * plan.M && JQ < 8192) JQ *= 2;
  plan.JQ = JQ;
  std::vector<unsigned short> jlut(JQ);
  for (int q = 0; q < JQ; ++q) {
    const double edge = static_cast<double>(q) / JQ;
    jlut[q] = static_cast<unsigned short>(std::lower_bound(alphas.begin(), alphas.end(), edge) - alphas.begin());
  }
  plan.binned_ok = plan.M <= kBinnedMaxM && plan.NK <= kBinnedMaxNK;

  plan.alpha.reset(std::max<size_t>(1, alphas.size()) * sizeof(double));
  plan.athr.reset(std::max<size_t>(1, alphas.size()) * sizeof(float));
  std::vector<float> athr(alphas.size());
  for (size_t i = 0; i < alphas.size(); ++i) {
    float f = static_cast<float>(alphas[i]);
    if (static_cast<double>(f) <= alphas[i]) f = std::nextafter(f, 2.0f);
    athr[i] = f;
  }
  if (!alphas.empty()) {
    RGN_HIP(hipMemcpy(plan.alpha.ptr, alphas.data(), alphas.size() * sizeof(double), hipMemcpyHostToDevice));
    RGN_HIP(hipMemcpy(plan.athr.ptr, athr.data(), athr.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  plan.jlut.reset(jlut.size() * sizeof(unsigned short));
  RGN_HIP(hipMemcpy(plan.jlut.ptr, jlut.data(), plan.jlut.bytes, hipMemcpyHostToDevice));
  {
    // the sub-histogram's lookup cells c of [c / JQ2, (c + 1) / JQ2): the alphas
    // below the cell, a flag for cells holding two alphas or more, and the
    // first threshold (as float bits)
    const int JQ2 = std::min(JQ, kLutCells), r = JQ / JQ2;
    std::vector<uint2> lut(JQ2);
    for (int c = 0; c < JQ2; ++c) {
      const unsigned j0 = jlut[c * r];
      const unsigned jn = c + 1 < JQ2 ? jlut[(c + 1) * r] : static_cast<unsigned>(plan.M);
      const float thr = j0 < static_cast<unsigned>(plan.M) ? athr[j0] : std::numeric_limits<float>::infinity();
      unsigned bits;
      std::memcpy(&bits, &thr, sizeof(bits));
      lut[c] = make_uint2(j0 | (jn - j0 >= 2 ? 0x10000u : 0u), bits);
    }
    plan.sublut.reset(lut.size() * sizeof(uint2));
    RGN_HIP(hipMemcpy(plan.sublut.ptr, lut.data(), plan.sublut.bytes, hipMemcpyHostToDevice));
  }
  plan.rptr.reset(rptr.size() * sizeof(int));
  RGN_HIP(hipMemcpy(plan.rptr.ptr, rptr.data(), plan.rptr.bytes, hipMemcpyHostToDevice));
  plan.rlist.reset(rlist.size() * sizeof(int));
  RGN_HIP(hipMemcpy(plan.rlist.ptr, rlist.data(), plan.rlist.bytes, hipMemcpyHostToDevice));
  plan.binfo.reset(bi.size() * sizeof(BinInfo));
  RGN_HIP(hipMemcpy(plan.binfo.ptr, bi.data(), bi.size() * sizeof(BinInfo), hipMemcpyHostToDevice));
  plan.seg.reset(seg.size() * sizeof(SegInfo));
  RGN_HIP(hipMemcpy(plan.seg.ptr, seg.data(), seg.size() * sizeof(SegInfo), hipMemcpyHostToDevice));
  {
    // log2 table (log2_tab): entry i holds 2 inv_i and -log2 inv_i, inv_i = RN(1 / c_i),
    // c_i = 1 + (i + 1/2) / 256, the logarithm in long double
    std::vector<double2> l2(kLog2Tab);
    for (int i = 0; i < kLog2Tab; ++i) {
      const double inv = 1.0 / (1.0 + (i + 0.5) / kLog2Tab);
      l2[i] = make_double2(2.0 * inv, static_cast<double>(-std::log2(static_cast<long double>(inv))));
    }
    plan.l2tab.reset(l2.size() * sizeof(double2));
    RGN_HIP(hipMemcpy(plan.l2tab.ptr, l2.data(), plan.l2tab.bytes, hipMemcpyHostToDevice));
  }
}

namespace {

constexpr int kThreads = 256;
constexpr int kBucketThreads = 1024;       // threads of a bucket block (two per CU at 64 VGPRs)
constexpr int kPer = 8;                    // particles per thread in the bucket pass
constexpr int kChunk = kBucketThreads * kPer;  // particles per bucket block
static_assert(kChunk <= 32768, "run descriptors pack start and length in 16 bits each");
constexpr size_t kLdsPerCU = 160 * 1024;   // LDS of a gfx950 CU
constexpr int kSubThreads = 512;           // threads of a sub-histogram work item
constexpr int kWin = 2048;                 // runs per LDS window of a work item
constexpr int kPrefThreads = 768;          // threads of a prefix block (two blocks per CU)
constexpr int kRowThreads = 1024;          // one row of run descriptors per scan block

struct BinnedDev {
  const double* alpha;
  const float* athr;
  const unsigned short* jlut;
  const uint2* sublut;
  const int* rptr;
  const int* rlist;
  const BinInfo* binfo;
  const SegInfo* seg;
  const float* bine;
  const double2* l2tab;
  int M, JQ, T, kmin, NK, nbins;
  EmitConst em;
  double K;
};

BinnedDev binned_view(const SyncPlan& p) {
  return BinnedDev{p.alpha.as<double>(), p.athr.as<float>(), p.jlut.as<unsigned short>(), p.sublut.as<uint2>(), p.rptr.as<int>(), p.rlist.as<int>(),
                   p.binfo.as<BinInfo>(), p.seg.as<SegInfo>(), p.bine.as<float>(), p.l2tab.as<double2>(),
                   p.M, p.JQ, p.T, p.kmin, p.NK, p.nbins, emit_const(p), p.K};
}

// log2 of a positive finite float to ~1e-11 absolute, from a table (in LDS):
// x = m 2^E with m in [1, 2); the top eight fraction bits i of m pick
// c_i = 1 + (i + 1/2) / 256 and the table's entry (2 inv_i, L_i), inv_i =
// RN(1 / c_i), L_i = -log2 inv_i; then log2 x = E + L_i + log2(1 + r) with
// r = m inv_i - 1 (one rounding, |r| < 2^-9) and ln(1 + r) through r^3
// (truncation below 6e-12).  Its three FMAs take constants as multiplicands
// or inline operands, and no division or reciprocal is needed.
__device__ __forceinline__ double log2_tab(float xf, const double2* __restrict__ tab) {
  int E;
  const float mf = frexpf(xf, &E);  // [0.5, 1), m = 2 mf
  const double2 t = tab[(__float_as_uint(mf) >> 15) & (kLog2Tab - 1)];
  const double r = fma(static_cast<double>(mf), t.x, -1.0);
  const double q = r * fma(r, fma(r, 1.0 / 3.0, -0.5), 1.0);  // ln(1 + r)
  return static_cast<double>(E - 1) + fma(q, 1.4426950408889634, t.y);
}

// W and B of a particle from its OmegaSync terms (l2: the log2 table in LDS);
// false when it contributes nowhere.
__device__ __forceinline__ bool particle_wb(const OmegaTerms& o, const RawParticle& r, const BinnedDev& P,
                                            const double2* l2, float& W, double& B) {
  W = o.chiR * r.w;
  if (!(o.e_peak > 0.0f && o.e_peak < INFINITY)) return false;  // synchrotron.hpp:162; NaN fails too
  B = log2_tab(o.e_peak, l2) * P.K;
  return B >= static_cast<double>(P.kmin) && B < static_cast<double>(P.kmin + P.NK);
}

// Number of alphas strictly below phi (0 <= phi < 1): the lookup grid gives
// the count below the grid cell's left edge, a short scan finishes it.
__device__ __forceinline__ int subcell(const double* __restrict__ alpha, const unsigned short* __restrict__ jlut,
                                       int M, int JQ, double phi) {
  int j = jlut[min(static_cast<int>(phi * JQ), JQ - 1)];
  while (j < M && alpha[j] < phi) ++j;
  return j;
}

// Exclusive prefix of per-thread values over an NT-thread block: a shuffle
// scan inside each wave, then the wave totals (wsum: NT / 64 words of LDS),
// added in a fixed order by every thread.  Sets total; every thread calls it.
template <class T, int NT>
__device__ __forceinline__ T block_prefix(T v, T& total, T* wsum) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T t = __shfl_up(x, d);
    if (lane >= d) x += t;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  T before = T(0), tot = T(0);
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const T sw = wsum[w];
    if (w < wave) before += sw;
    tot += sw;
  }
  total = tot;
  __syncthreads();  // wsum may be reused
  return before + (x - v);
}

// Block-wide exclusive scan of a[0..N) into o[0..N) (o may alias a); returns
// the total.  part: NT / 64 words of LDS.  Every thread of the NT-thread block
// calls it.
template <class T, int NT = kThreads, class PA, class PO>
__device__ T block_scan_exclusive(PA a, PO o, int N, T* part) {
  const int per = (N + NT - 1) / NT;
  const int i0 = min(N, static_cast<int>(threadIdx.x) * per), i1 = min(N, i0 + per);
  T s = 0;
  for (int i = i0; i < i1; ++i) s += a[i];
  T total;
  T acc = block_prefix<T, NT>(s, total, part);
  for (int i = i0; i < i1; ++i) {
    const T v = a[i];
    o[i] = acc;
    acc += v;
  }
  __syncthreads();
  return total;
}

// Run descriptor of bucket k in bucket block b: start << 16 | length (both at
// most kChunk = 8192).  Only non-empty runs are written (each scattered 4-byte
// store costs a 32-byte write-back); the scan pass zeroes what it reads, so the
// table stays clear between calls and fill_zero runs only on a new buffer.
__device__ __forceinline__ unsigned run_len(unsigned r) { return r & 0xffffu; }
__device__ __forceinline__ unsigned run_start(unsigned r) { return r >> 16; }

__global__ void __launch_bounds__(kThreads) fill_zero(unsigned* __restrict__ p, size_t n) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * kThreads) {
    p[i] = 0u;
  }
}

// Pass 1: per block of kChunk particles, sources sorted by bucket k.
//   sortWP[b * kChunk + i]: the block's sources (W, phi) in bucket order;
//   runs[k * G + b]: run descriptor of bucket k in block b.
// Two blocks per CU without E fields (64 VGPRs), one with them.  NT: the
// particles are read with non-temporal loads (see sync_forward_binned).
template <bool HAS_E, bool NT>
__global__ void __launch_bounds__(kBucketThreads) __attribute__((amdgpu_waves_per_eu(HAS_E ? 4 : 8)))
binned_bucket(ParticlesDev D, unsigned n, BinnedDev P, float2* __restrict__ sortWP, unsigned* __restrict__ runs) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NK = P.NK;
  double2* l2 = reinterpret_cast<double2*>(smem);
  unsigned* hist = reinterpret_cast<unsigned*>(l2 + kLog2Tab);
  unsigned* off = hist + NK;
  unsigned* part = off + NK;
  float2* sWP = reinterpret_cast<float2*>(part + kBucketThreads / 64);  // after an even word count: 8-byte aligned
  for (int i = threadIdx.x; i < NK; i += kBucketThreads) hist[i] = 0u;
  if (threadIdx.x < kLog2Tab) l2[threadIdx.x] = P.l2tab[threadIdx.x];
  __syncthreads();
  const unsigned G = gridDim.x;
  const unsigned base = blockIdx.x * static_cast<unsigned>(kChunk);
  int kk[kPer];
  float Wr[kPer], ph[kPer];
  unsigned rk[kPer];
  unsigned p = base + threadIdx.x;
  RawParticle raw{};
  if (p < n) raw = load_particle32<NT>(D, p);
#pragma unroll
  for (int i = 0; i < kPer; ++i, p += kBucketThreads) {
    const RawParticle cur = raw;
    if (i + 1 < kPer && p + kBucketThreads < n) raw = load_particle32<NT>(D, p + kBucketThreads);
    kk[i] = -1;
    Wr[i] = 0.0f;
    ph[i] = 0.0f;
    rk[i] = 0u;
    float W;
    double B;
    // OmegaSync with refined reciprocals (about one double ulp): chiR and
    // e_peak round to the reference's floats except within ~2 double ulps of
    // a float rounding boundary, where one float ulp moves W and B; the
    // binned sum never depended on them bit for bit
    if (p < n && particle_wb(omega_terms_approx<HAS_E>(cur, P.em), cur, P, l2, W, B)) {
      const double fl = floor(B);
      kk[i] = static_cast<int>(fl) - P.kmin;
      Wr[i] = W;
      ph[i] = fminf(static_cast<float>(B - fl), 0x1.fffffep-1f);
      rk[i] = atomicAdd(hist + kk[i], 1u);
    }
  }
  __syncthreads();
  block_scan_exclusive<unsigned, kBucketThreads>(hist, off, NK, part);
  for (int k = threadIdx.x; k < NK; k += kBucketThreads) {
    if (hist[k]) runs[static_cast<size_t>(k) * G + blockIdx.x] = off[k] << 16 | hist[k];
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    if (kk[i] >= 0) {
      sWP[off[kk[i]] + rk[i]] = make_float2(Wr[i], ph[i]);
    }
  }
  __syncthreads();
  const unsigned total = off[NK - 1] + hist[NK - 1];
  for (unsigned i = threadIdx.x; i < total; i += kBucketThreads) st32(sortWP, base + i, sWP[i]);
}

// Pass 2a: per bucket k, its non-empty runs in block order, compacted:
//   cvoff[k * (G + 1) + i]: virtual offset of run i within the bucket
//                           (cvoff[k * (G + 1) + nrun[k]] = totals[k]);
//   cidx[k * G + i]:        physical index of the run's first source;
//   irun[k * IS + t]:       the run holding source t * item of the bucket, the
//                           first source of its work item t (a run is at most
//                           one chunk, not longer than an item, so it holds at
//                           most one such source).
// One exclusive scan over blocks of (run length, non-empty flag) packed in 64 bits.
// The descriptors read are zeroed again, so the table needs no clearing pass
// before the next call.
__global__ void __launch_bounds__(kRowThreads) __attribute__((amdgpu_waves_per_eu(8))) binned_scan_rows(unsigned* __restrict__ runs, int G,
                                                             unsigned* __restrict__ cvoff,
                                                             unsigned* __restrict__ cidx,
                                                             unsigned* __restrict__ nrun,
                                                             unsigned* __restrict__ totals,
                                                             unsigned* __restrict__ irun, int IS,
                                                             unsigned item) {
  __shared__ unsigned long long part[kRowThreads / 64];
  const int k = blockIdx.x;
  unsigned* row = runs + static_cast<size_t>(k) * G;
  unsigned* vo = cvoff + static_cast<size_t>(k) * (G + 1);
  unsigned* ci = cidx + static_cast<size_t>(k) * G;
  const int per = (G + kRowThreads - 1) / kRowThreads;
  const int i0 = min(G, static_cast<int>(threadIdx.x) * per), i1 = min(G, i0 + per);
  // up to kRowCache descriptors per thread stay in registers between the passes
  // (G <= 4096 blocks, 2^25 particles; 64 VGPRs, so two blocks share a CU)
  constexpr int kRowCache = 4;
  unsigned cache[kRowCache];
  unsigned long long acc = 0;
  if (per <= kRowCache) {
#pragma unroll
    for (int q = 0; q < kRowCache; ++q) {
      cache[q] = i0 + q < i1 ? row[i0 + q] : 0u;
      const unsigned len = run_len(cache[q]);
      acc += len + (len ? (1ull << 32) : 0ull);
    }
  } else {
    for (int i = i0; i < i1; ++i) {
      const unsigned len = run_len(row[i]);
      acc += len + (len ? (1ull << 32) : 0ull);
    }
  }
  unsigned long long all;
  acc = block_prefix<unsigned long long, kRowThreads>(acc, all, part);
  auto emit = [&](int i, unsigned r) {
    const unsigned len = run_len(r);
    if (len) {
      const unsigned c = static_cast<unsigned>(acc >> 32);
      const unsigned v = static_cast<unsigned>(acc);
      vo[c] = v;
      ci[c] = static_cast<unsigned>(i) * kChunk + run_start(r);
      const unsigned t = (v + item - 1) / item;  // first item start at or after v
      if (t * item < v + len) irun[static_cast<size_t>(k) * IS + t] = c;
      acc += len + (1ull << 32);
      row[i] = 0u;  // leave the run table zeroed for the next call
    }
  };
  if (per <= kRowCache) {
#pragma unroll
    for (int q = 0; q < kRowCache; ++q) {
      if (i0 + q < i1) emit(i0 + q, cache[q]);
    }
  } else {
    for (int i = i0; i < i1; ++i) emit(i, row[i]);
  }
  if (threadIdx.x == kRowThreads - 1) {
    const unsigned nr = static_cast<unsigned>(all >> 32);
    const unsigned total = static_cast<unsigned>(all);
    vo[nr] = total;
    nrun[k] = nr;
    totals[k] = total;
  }
}

// Work items: bucket k is cut into ceil(totals[k] / item) items of `item`
// sources, numbered bucket after bucket.  Both helpers are block-wide (every
// thread of the NT-thread block calls them; wsum: NT / 64 words of LDS).
__device__ __forceinline__ unsigned items_of(unsigned total, unsigned item) { return (total + item - 1) / item; }

// Bucket k of item `it` and the bucket's first item; false past the last item
// (the same answer in every thread).
template <int NT>
__device__ __forceinline__ bool item_bucket(const unsigned* __restrict__ totals, int NK, unsigned it, unsigned item,
                                            int& k, unsigned& first, unsigned* wsum, unsigned* res) {
  const int per = (NK + NT - 1) / NT;
  const int q0 = min(NK, static_cast<int>(threadIdx.x) * per), q1 = min(NK, q0 + per);
  unsigned c = 0;
  for (int q = q0; q < q1; ++q) c += items_of(totals[q], item);
  unsigned all;
  unsigned acc = block_prefix<unsigned, NT>(c, all, wsum);
  for (int q = q0; q < q1; ++q) {
    const unsigned m = items_of(totals[q], item);
    if (acc <= it && it < acc + m) {
      res[0] = static_cast<unsigned>(q);
      res[1] = acc;
    }
    acc += m;
  }
  __syncthreads();
  if (it >= all) return false;
  k = static_cast<int>(res[0]);
  first = res[1];
  return true;
}

// First item of bucket k (the number of items of buckets before it).
template <int NT>
__device__ __forceinline__ unsigned items_before(const unsigned* __restrict__ totals, int k, unsigned item,
                                                 unsigned* wsum) {
  unsigned c = 0;
  for (int q = threadIdx.x; q < k; q += NT) c += items_of(totals[q], item);
  unsigned all;
  block_prefix<unsigned, NT>(c, all, wsum);
  return all;
}

// Pass 3: sub-histogram of one work item: sources [v0, v1) of bucket k, found
// as runs of the bucket blocks' sorted chunks.  The item's runs are copied to
// LDS; each wave walks a contiguous share of the positions, 64 per step (one
// per lane), four loads in flight per lane, and reduces the sources into the
// M + 1 sub-cells with LDS atomics.
template <int kBatch>
__global__ void __launch_bounds__(kSubThreads) binned_subhist(BinnedDev P, int G, const unsigned* __restrict__ totals,
                                                              const unsigned* __restrict__ cvoff,
                                                              const unsigned* __restrict__ cidx,
                                                              const unsigned* __restrict__ nrun,
                                                              const unsigned* __restrict__ irun, int IS,
                                                              unsigned item, const float2* __restrict__ sortWP,
                                                              double* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int kWaves = kSubThreads / 64;
  const int M1 = P.M + 1;
  double* m0 = reinterpret_cast<double*>(smem);
  double* m1 = m0 + M1;
  const int JQ2 = min(P.JQ, kLutCells);
  uint2* lut = reinterpret_cast<uint2*>(m1 + M1);            // JQ2
  float* athr = reinterpret_cast<float*>(lut + JQ2);
  unsigned* wro = reinterpret_cast<unsigned*>(athr + P.M);   // kWin + 65
  unsigned* wlo = wro + kWin + 65;                             // kWin
  __shared__ unsigned s_item[2], s_wsum[kWaves];
  // per wave: run starts marked by position within a 64-position step (see step)
  __shared__ unsigned s_mark[kWaves][65];
  const unsigned it = blockIdx.x;
  int k;
  unsigned first;
  if (!item_bucket<kSubThreads>(totals, P.NK, it, item, k, first, s_wsum, s_item)) return;  // uniform per block
  for (int i = threadIdx.x; i < M1; i += kSubThreads) {
    m0[i] = 0.0;
    m1[i] = 0.0;
  }
  for (int i = threadIdx.x; i < P.M; i += kSubThreads) athr[i] = P.athr[i];
  for (int c = threadIdx.x; c < JQ2; c += kSubThreads) lut[c] = P.sublut[c];
  s_mark[threadIdx.x >> 6][threadIdx.x & 63] = 0u;  // entry 64 is only written
  __syncthreads();
  const unsigned v0 = (it - first) * item;
  const unsigned v1 = min(totals[k], v0 + item);
  const unsigned* vo = cvoff + static_cast<size_t>(k) * (G + 1);
  const unsigned* ci = cidx + static_cast<size_t>(k) * G;
  // first and last run of the item: the runs holding sources v0 and v1 - 1.
  // The scan pass recorded the run holding each item's first source; the
  // next item's first run holds v1 - 1 too unless it starts exactly at v1.
  const unsigned* ir = irun + static_cast<size_t>(k) * IS + (it - first);
  const int rlo = static_cast<int>(ir[0]);
  int rhi;
  if (v1 == totals[k]) {
    rhi = static_cast<int>(nrun[k]) - 1;
  } else {
    const int r1 = static_cast<int>(ir[1]);
    rhi = vo[r1] == v1 ? r1 - 1 : r1;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int ws = rlo; ws <= rhi; ws += kWin) {
    const int nc = min(rhi - ws + 1, kWin);
    // the window's runs clipped to [v0, v1): run c covers virtual
    // [crs[c], crs[c + 1]), and position v of it is physical index v + dlt[c];
    // 64 entries of ~0 follow crs[nc], so a step reads starts without a check
    unsigned* crs = wro;   // nc + 65 entries
    unsigned* dlt = wlo;   // nc entries
    for (int c = threadIdx.x; c < nc; c += kSubThreads) {
      const unsigned r0 = vo[ws + c];
      const unsigned ra = max(r0, v0);
      crs[c] = ra;
      dlt[c] = ci[ws + c] - r0;  // modulo 2^32
    }
    if (threadIdx.x == 0) crs[nc] = min(vo[ws + nc], v1);
    if (threadIdx.x < 64) crs[nc + 1 + threadIdx.x] = 0xffffffffu;
    __syncthreads();
    if (nc > 0) {
      // wave w walks positions [pa, pb) of the window, lane-interleaved
      const unsigned va = crs[0], vb = crs[nc];
      const unsigned per = ((vb - va + kWaves - 1) / kWaves + 63) & ~63u;
      const unsigned pa = min(vb, va + per * wave), pb = min(vb, pa + per);
      // The wave's positions go 64 at a time, lane l taking v00 + l.  c0 is
      // the run holding v00 (wave-uniform).  Each lane reads one run start
      // past c0; those inside the next 64 positions mark their offset in the
      // wave's LDS row (with a tag unique to the step, so the row is never
      // cleared), a ballot of the row gives the position bitmap, and a
      // lane's run is c0 plus the starts at or below its position: no
      // per-lane search and no loop over the starts.
      int c0 = 0;
      if (pa < pb) {
        int lo = 0, hi = nc - 1;  // last run starting at or before pa
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (crs[mid] <= pa) lo = mid; else hi = mid - 1;
        }
        c0 = lo;
      }
      unsigned* mark = s_mark[wave];
      const unsigned long long below = (2ull << lane) - 1;  // lanes 0 .. lane (all for lane 63)
      unsigned sc = crs[c0 + 1 + lane];  // the lane's start past c0 (> every position of run c0)
      // every lane calls it; returns the lane's run (past the window end, nc,
      // only for positions >= pb)
      auto step = [&](unsigned v00) {
        const int L = __popcll(__ballot(sc <= v00 + 64));  // starts in (v00, v00 + 64]
        int run = c0;
        if (L) {  // wave-uniform; none in most steps of the long runs
          const unsigned tag = v00 + 1;  // positions only grow: unique per step, never 0
          // starts at v00 + 1 .. v00 + 63 mark their offset; the others write
          // the unread entry 64
          const unsigned d = sc - v00;
          mark[d - 1 < 63u ? d : 64u] = tag;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          run += __popcll(__ballot(mark[lane] == tag) & below);
          c0 += L;
          sc = crs[c0 + 1 + lane];  // c0 <= nc
        }
        return run;
      };
      auto accumulate = [&](float2 wp) {  // 0 <= phi < 1: the cell index is below JQ2
        const uint2 e = lut[static_cast<int>(wp.y * JQ2)];
        int j = static_cast<int>(e.x & 0xffffu) + (__uint_as_float(e.y) <= wp.y);
        if (e.x >> 16) {  // rare: a cell with several alphas
          while (j < P.M && athr[j] <= wp.y) ++j;
        }
        const double wd = wp.x;
        atomicAdd(m0 + j, wd);
        atomicAdd(m1 + j, wd * static_cast<double>(wp.y));
      };
      unsigned v00 = pa;
      for (; v00 + 64 * kBatch <= pb; v00 += 64 * kBatch) {
        unsigned idx[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
          const unsigned v = v00 + q * 64 + lane;
          const int r = step(v00 + q * 64);
          idx[q] = v + dlt[r];
        }
        float2 wp[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) wp[q] = ld32(sortWP, idx[q]);
#pragma unroll
        for (int q = 0; q < kBatch; ++q) accumulate(wp[q]);
      }
      if (v00 < pb) {  // the last, partial batch
        unsigned idx[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
          const unsigned v = v00 + q * 64 + lane;
          const int r = min(step(v00 + q * 64), nc - 1);
          idx[q] = v < pb ? v + dlt[r] : ~0u;
        }
        float2 wp[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) wp[q] = idx[q] != ~0u ? ld32(sortWP, idx[q]) : make_float2(0.0f, 0.0f);
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
          if (idx[q] != ~0u) accumulate(wp[q]);
        }
      }
    }
    __syncthreads();
  }
  double* out = partial + static_cast<size_t>(it) * 2 * M1;
  for (int i = threadIdx.x; i < M1; i += kSubThreads) {
    out[i] = m0[i];
    out[M1 + i] = m1[i];
  }
}

// Pass 4: per bucket, its items summed (interleaved item groups, combined in a
// fixed order), then inclusive prefix sums over sub-cells:
// P0[k][j] = sum_{j' <= j} m0, P1 likewise.  The block's threads form as many
// groups of C = M + 1 rounded up to 64 columns as fit (two of 320 for 256
// bins), so a pass covers every sub-cell and each group sums a share of the
// items.
__global__ void __launch_bounds__(kPrefThreads) binned_prefix(int M, unsigned item, const unsigned* __restrict__ totals,
                                                              const double* __restrict__ partial,
                                                              double* __restrict__ P0, double* __restrict__ P1) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double part[kPrefThreads / 64];
  __shared__ double red[2][kPrefThreads];  // [m0, m1][group * C + column]
  const int M1 = M + 1;
  double* s0 = reinterpret_cast<double*>(smem);
  double* s1 = s0 + M1;
  const int k = blockIdx.x;
  __shared__ unsigned s_wsum[kPrefThreads / 64];
  const unsigned i0 = items_before<kPrefThreads>(totals, k, item, s_wsum), i1 = i0 + items_of(totals[k], item);
  const int C = min((M1 + 63) & ~63, kPrefThreads);
  const int groups = kPrefThreads / C;
  const int grp = static_cast<int>(threadIdx.x) / C, col = static_cast<int>(threadIdx.x) - grp * C;
  for (int j0 = 0; j0 < M1; j0 += C) {
    const int j = j0 + col;
    // this group's items it = i0 + grp + groups q, four of them per pass with
    // their loads in flight together, summed in a fixed order
    double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
    if (grp < groups && j < M1) {
      unsigned it = i0 + grp;
      for (; it + 3 * groups < i1; it += 4 * groups) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          a[q] += partial[static_cast<size_t>(it + groups * q) * 2 * M1 + j];
          b[q] += partial[static_cast<size_t>(it + groups * q) * 2 * M1 + M1 + j];
        }
      }
      for (int q = 0; it < i1; it += groups, ++q) {  // at most three left
        a[q] += partial[static_cast<size_t>(it) * 2 * M1 + j];
        b[q] += partial[static_cast<size_t>(it) * 2 * M1 + M1 + j];
      }
    }
    if (grp < groups) {
      red[0][threadIdx.x] = (a[0] + a[1]) + (a[2] + a[3]);
      red[1][threadIdx.x] = (b[0] + b[1]) + (b[2] + b[3]);
    }
    __syncthreads();
    if (grp == 0 && j < M1) {
      double t0 = red[0][col], t1 = red[1][col];
      for (int g = 1; g < groups; ++g) {
        t0 += red[0][g * C + col];
        t1 += red[1][g * C + col];
      }
      s0[j] = t0;
      s1[j] = t1;
    }
    __syncthreads();
  }
  double* r0 = P0 + static_cast<size_t>(k) * M1;
  double* r1 = P1 + static_cast<size_t>(k) * M1;
  const double t0 = block_scan_exclusive<double, kPrefThreads>(s0, s0, M1, part);
  const double t1 = block_scan_exclusive<double, kPrefThreads>(s1, s1, M1, part);
  // inclusive = exclusive of the next entry; the last one is the total
  for (int j = threadIdx.x; j < M1; j += kPrefThreads) {
    r0[j] = j + 1 < M1 ? s0[j + 1] : t0;
    r1[j] = j + 1 < M1 ? s1[j + 1] : t1;
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

// Pass 5: S_e = e_e sum_k (...), one block per bin (a thread per bucket),
// fixed-order reduction.
__global__ void __launch_bounds__(kThreads) binned_evaluate(BinnedDev P, const double* __restrict__ P0,
                                                            const double* __restrict__ P1,
                                                            float* __restrict__ out) {
  const int e = blockIdx.x;
  const BinInfo b = P.binfo[e];
  __shared__ double red[kThreads / 64];
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
  // a fixed shuffle tree per wave, then the four wave sums in order
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) acc += __shfl_xor(acc, w);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int w = 1; w < kThreads / 64; ++w) t += red[w];
    const float ee = P.bine[e];
    out[e] = (b.r < 0 || ee == 0.0f) ? 0.0f : static_cast<float>(static_cast<double>(ee) * t);
  }
}

// ---------------------------------------------------------------------------
// Adjoint.  A source in bucket k, sub-cell j sees bin e on segment m = a_e - k
// when j <= r_e and on m - 1 otherwise, so
//   U[k][j] = sum_e g'_e line_m(e, k) + sum_{r < j} D_k[r],
//   D_k[r]  = sum_{e: r_e = r} g'_e (line_{m-1} - line_m)(e, k),
// with line = (C, sigma) and g'_e = dL/dS_e * e_e.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) binned_adj_table(BinnedDev P, const float* __restrict__ gbar,
                                                             double2* __restrict__ U) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* d0 = reinterpret_cast<double*>(smem);
  double* d1 = d0 + P.M;
  __shared__ double part[kThreads / 64];
  __shared__ double base0[kThreads], base1[kThreads];
  const int kk = blockIdx.x;
  const int k = P.kmin + kk;
  double b0 = 0.0, b1 = 0.0;
  for (int r = threadIdx.x; r < P.M; r += kThreads) {
    double x0 = 0.0, x1 = 0.0;
    for (int q = P.rptr[r]; q < P.rptr[r + 1]; ++q) {
      const int e = P.rlist[q];
      const BinInfo b = P.binfo[e];
      const double g = static_cast<double>(gbar[e]) * P.bine[e];
      double Cp, sp, Cm, sm;
      seg_line(P, b, k, b.a - k, Cp, sp);
      seg_line(P, b, k, b.a - k - 1, Cm, sm);
      b0 += g * Cp;
      b1 += g * sp;
      x0 += g * (Cm - Cp);
      x1 += g * (sm - sp);
    }
    d0[r] = x0;
    d1[r] = x1;
  }
  base0[threadIdx.x] = b0;
  base1[threadIdx.x] = b1;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      base0[threadIdx.x] += base0[threadIdx.x + w];
      base1[threadIdx.x] += base1[threadIdx.x + w];
    }
    __syncthreads();
  }
  const double B0 = base0[0], B1 = base1[0];
  const double t0 = block_scan_exclusive<double>(d0, d0, P.M, part);
  const double t1 = block_scan_exclusive<double>(d1, d1, P.M, part);
  double2* row = U + static_cast<size_t>(kk) * (P.M + 1);
  for (int j = threadIdx.x; j <= P.M; j += kThreads) {
    row[j] = j < P.M ? make_double2(B0 + d0[j], B1 + d1[j]) : make_double2(B0 + t0, B1 + t1);
  }
}

struct GradPtrs {
  float* gw;
  float* gu[3];
  float* gb[3];
  float* ge[3];
};

// One U entry per particle.  With STAGE, a persistent block first bins the
// buckets of its first kGather particles, then copies the U rows of the
// kStage-row window holding most of them into LDS (a power-law population
// puts most particles in a few consecutive buckets); particles in the window
// read their entry from LDS, the rest from the table in memory (L2).  The
// random 16-byte reads from L2 otherwise move a whole cache line each.  One
// 1024-thread block per CU with 120 KB of rows (29 at 256 bins) measured
// 197 us at 2^24 particles, two 512-thread blocks with 60 KB each 205 us.
// HAS_E = false: the store has no E and dL/dE is not wanted (the E terms
// drop out of OmegaSync and of the gradients).
constexpr int kGather = 1024;                // threads of a staged gather block
constexpr size_t kStageBytes = 120 * 1024;   // LDS for the staged U rows

template <bool HAS_E, bool NT, bool STAGE>
__global__ void __launch_bounds__(STAGE ? kGather : kThreads) binned_adj_gather(ParticlesDev D, size_t n, BinnedDev P,
                                                                              const double2* __restrict__ U,
                                                                              GradPtrs G, int kstage) {
  constexpr int NT_ = STAGE ? kGather : kThreads;
  extern __shared__ __align__(16) unsigned char smem[];
  double2* urows = reinterpret_cast<double2*>(smem);  // STAGE: kstage rows of M + 1 (or the histogram)
  double2* l2 = reinterpret_cast<double2*>(smem + (STAGE ? kStageBytes : 0));
  double* alpha = reinterpret_cast<double*>(l2 + kLog2Tab);
  unsigned short* jlut = reinterpret_cast<unsigned short*>(alpha + P.M);
  __shared__ int s_w0;
  for (int i = threadIdx.x; i < kLog2Tab; i += NT_) l2[i] = P.l2tab[i];
  for (int i = threadIdx.x; i < P.M; i += NT_) alpha[i] = P.alpha[i];
  for (int i = threadIdx.x; i < P.JQ; i += NT_) jlut[i] = P.jlut[i];
  const double dlnx = P.K / 0.69314718055994530942;  // dt / dln x
  // n < 2^29 (the factorised form's limit): 32-bit indices and byte offsets
  const unsigned nn = static_cast<unsigned>(n);
  const unsigned stride = gridDim.x * NT_;
  const size_t M1 = static_cast<size_t>(P.M) + 1;
  int w0 = 0, w1 = 0;  // staged buckets [w0, w1)
  if (STAGE) {
    unsigned* hist = reinterpret_cast<unsigned*>(smem);
    for (int i = threadIdx.x; i < P.NK; i += NT_) hist[i] = 0u;
    __syncthreads();
    const unsigned p = blockIdx.x * NT_ + threadIdx.x;
    if (p < nn) {
      const RawParticle r = load_particle32<false>(D, p);
      float W;
      double B;
      if (particle_wb(omega_terms_approx<HAS_E>(r, P.em), r, P, l2, W, B)) {
        atomicAdd(hist + (static_cast<int>(floor(B)) - P.kmin), 1u);
      }
    }
    __syncthreads();
    // the window of kstage buckets with the most sampled particles (lowest start on ties)
    const int nw = max(1, P.NK - kstage + 1);
    unsigned best = 0;
    int bw = 0x7fffffff;
    for (int w = threadIdx.x; w < nw; w += NT_) {
      unsigned c = 0;
      for (int q = w; q < min(P.NK, w + kstage); ++q) c += hist[q];
      if (c > best || (c == best && w < bw)) {
        best = c;
        bw = w;
      }
    }
    // block arg-max: pack (count, -start) into 64 bits
    unsigned long long key = (static_cast<unsigned long long>(best) << 32) |
                             static_cast<unsigned>(0x7fffffff - min(bw, 0x7fffffff));
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
      const unsigned long long o = __shfl_xor(key, m);
      key = o > key ? o : key;
    }
    __shared__ unsigned long long s_key[NT_ / 64];
    if ((threadIdx.x & 63) == 0) s_key[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long k2 = s_key[0];
      for (int i = 1; i < NT_ / 64; ++i) k2 = s_key[i] > k2 ? s_key[i] : k2;
      s_w0 = 0x7fffffff - static_cast<int>(k2 & 0xffffffffu);
    }
    __syncthreads();  // the histogram is read: its LDS takes the rows
    w0 = min(s_w0, nw - 1);
    w1 = min(P.NK, w0 + kstage);
    const size_t nrow = static_cast<size_t>(w1 - w0) * M1;
    const double2* src = U + static_cast<size_t>(w0) * M1;
    for (size_t i = threadIdx.x; i < nrow; i += NT_) urows[i] = src[i];
  }
  __syncthreads();
  unsigned p = blockIdx.x * NT_ + threadIdx.x;
  // NT: particles and gradients stream through once (non-temporal), so that
  // the table lookups keep the caches
  // two particles ahead in flight per thread (the gather is limited by memory
  // latency at two blocks per CU)
  RawParticle r1{}, r2{};
  if (p < nn) r1 = load_particle32<NT>(D, p);
  if (p + stride < nn) r2 = load_particle32<NT>(D, p + stride);
  for (; p < nn; p += stride) {
    const RawParticle cur = r1;
    r1 = r2;
    if (p + 2 * stride < nn) r2 = load_particle32<NT>(D, p + 2 * stride);
    float W;
    double B;
    ParticleGrads g{};
    const OmegaTerms o = omega_terms_approx<HAS_E>(cur, P.em);
    if (particle_wb(o, cur, P, l2, W, B)) {
      const double fl = floor(B);
      const double phi = B - fl;
      const int kk = static_cast<int>(fl) - P.kmin;
      const int j = subcell(alpha, jlut, P.M, P.JQ, phi);
      const double2 u = (STAGE && kk >= w0 && kk < w1) ? urows[static_cast<size_t>(kk - w0) * M1 + j]
                                                       : U[static_cast<size_t>(kk) * M1 + j];
      // sum_e g'_e F and sum_e g'_e dF/dln x at this particle
      g = emission_grads<HAS_E>(o, cur, u.x - u.y * phi, dlnx * u.y, P.em);
    }
    const double gw = g.gw, gux = g.gu[0], guy = g.gu[1], guz = g.gu[2];
    const double gbx = g.gb[0], gby = g.gb[1], gbz = g.gb[2];
    const double gex = g.ge[0], gey = g.ge[1], gez = g.ge[2];
    if (G.gw) st32<NT>(G.gw, p, static_cast<float>(gw));
    if (G.gu[0]) st32<NT>(G.gu[0], p, static_cast<float>(gux));
    if (G.gu[1]) st32<NT>(G.gu[1], p, static_cast<float>(guy));
    if (G.gu[2]) st32<NT>(G.gu[2], p, static_cast<float>(guz));
    if (G.gb[0]) st32<NT>(G.gb[0], p, static_cast<float>(gbx));
    if (G.gb[1]) st32<NT>(G.gb[1], p, static_cast<float>(gby));
    if (G.gb[2]) st32<NT>(G.gb[2], p, static_cast<float>(gbz));
    if (G.ge[0]) st32<NT>(G.ge[0], p, static_cast<float>(gex));
    if (G.ge[1]) st32<NT>(G.ge[1], p, static_cast<float>(gey));
    if (G.ge[2]) st32<NT>(G.ge[2], p, static_cast<float>(gez));
  }
}

}  // namespace

// LDS bytes of the bucket pass for a plan (log2 table, hist, offsets, scan
// partials and the block's sorted sources).
static size_t bucket_smem(const SyncPlan& plan) {
  return sizeof(double2) * kLog2Tab + sizeof(unsigned) * (2 * static_cast<size_t>(plan.NK) + kBucketThreads / 64) +
         2 * sizeof(float) * kChunk;
}

static size_t subhist_smem(const SyncPlan& plan) {
  const size_t M1 = plan.M + 1;
  return sizeof(double) * 2 * M1 + sizeof(uint2) * std::min(plan.JQ, kLutCells) + sizeof(float) * plan.M +
         sizeof(unsigned) * (2 * kWin + 65);
}

void sync_forward_binned(const SyncPlan& plan, const ParticlesDev& D, size_t n, float* out, hipStream_t st) {
  RGN_REQUIRE(plan.binned_ok && n < kBinnedMaxParticles, RGN_ERR_UNSUPPORTED, "plan outside the binned limits");
  const BinnedDev P = binned_view(plan);
  if (plan.NK == 0 || n == 0) {
    // a fill kernel rather than a memset node, so the call stays safe to capture
    if (plan.nbins > 0) {
      fill_zero<<<static_cast<unsigned>(ceil_div(plan.nbins, kThreads)), kThreads, 0, st>>>(
        reinterpret_cast<unsigned*>(out), static_cast<size_t>(plan.nbins));
      RGN_HIP(hipGetLastError());
    }
    return;
  }
  const int G = static_cast<int>(ceil_div(n, kChunk));
  const size_t NK = plan.NK;
  const size_t M1 = plan.M + 1;
  // Work-item size (sources), at least a bucket block's chunk.  The items
  // run as one round of resident sub-histogram blocks; the full items should
  // take about two thirds of that round's slots (the rest hold the buckets'
  // last, partial items): measured fastest at 2^24 and at 2^26 particles
  // (52.8 against 61.1 µs for 16384-source items at 2^24, 190 against 215 µs
  // for 65536 at 2^26).  RAGNAR_SUBHIST_ITEM overrides it.
  // + the static LDS (marks, item and scan words, with alignment)
  const size_t sub_lds = subhist_smem(plan) + sizeof(unsigned) * (kSubThreads / 64 * 66 + 2) + 64;
  const size_t sub_slots = static_cast<size_t>(num_cus()) * std::max<size_t>(1, std::min<size_t>(4, kLdsPerCU / sub_lds));
  const size_t auto_item = std::max<size_t>(kChunk, ceil_div(3 * n, 2 * sub_slots));
  const unsigned item = static_cast<unsigned>(std::max<size_t>(
    static_cast<size_t>(env_int("RAGNAR_SUBHIST_ITEM", static_cast<int>(std::min<size_t>(auto_item, 1u << 30)))),
    kChunk));
  // items: sum over k of ceil(totals_k / item) <= n / item + NK
  const size_t max_items = n / item + NK + 1;
  float2* sortWP = static_cast<float2*>(scratch(sizeof(float2) * G * size_t{kChunk}, 6));
  bool fresh = false;
  unsigned* runs = static_cast<unsigned*>(scratch(sizeof(unsigned) * NK * G, 8, &fresh));
  unsigned* cvoff = static_cast<unsigned*>(scratch(sizeof(unsigned) * NK * (G + 1), 10));
  unsigned* cidx = static_cast<unsigned*>(scratch(sizeof(unsigned) * NK * G, 9));
  unsigned* small = static_cast<unsigned*>(scratch(sizeof(unsigned) * (2 * (NK + 1) + 64), 11));
  // first runs of the work items: a bucket holds at most ceil(G kChunk / item) items
  const int IS = static_cast<int>(ceil_div(size_t{kChunk} * G, item));
  unsigned* irun = static_cast<unsigned*>(scratch(sizeof(unsigned) * NK * IS, 14));
  unsigned* totals = small;
  unsigned* nrun = small + (NK + 1);
  double* partial = static_cast<double*>(scratch(sizeof(double) * 2 * M1 * max_items, 12));
  double* pref = static_cast<double*>(scratch(sizeof(double) * 2 * M1 * NK, 13));
  double* P0 = pref;
  double* P1 = pref + M1 * NK;

  // Leading entries of the run table known to be zero: every call's scan pass
  // re-zeroes the descriptors its bucket pass wrote, so a fill is needed only
  // for a new buffer or a larger table (reset before the bucket pass, so that a
  // call interrupted in between leaves the next one to fill).
  static size_t runs_clean = 0;
  if (fresh) runs_clean = 0;
  if (NK * G > runs_clean) {
    fill_zero<<<static_cast<unsigned>(std::min<size_t>(ceil_div(NK * G, kThreads), 4096)), kThreads, 0, st>>>(
      runs, NK * G);
    RGN_HIP(hipGetLastError());
    stage_mark(st, "fill_zero");
  }
  const size_t clean_after = std::max(runs_clean, NK * G);
  runs_clean = 0;
  const bool has_e = D.e[0] || D.e[1] || D.e[2];
  // Particles larger than the device's last-level cache (256 MiB on MI355X)
  // are read non-temporally: they would not be found there on the next call
  // anyway, and the bucket pass ran 107 -> 100 us at 2^24 particles (403 MB).
  const size_t in_bytes = sizeof(float) * n * (6 + (has_e ? 3 : 0) + (D.w ? 1 : 0));
  const bool nt = in_bytes > (size_t{256} << 20);
  const unsigned nn = static_cast<unsigned>(n);
  const size_t bsm = bucket_smem(plan);
  if (has_e) {
    if (nt) binned_bucket<true, true><<<G, kBucketThreads, bsm, st>>>(D, nn, P, sortWP, runs);
    else binned_bucket<true, false><<<G, kBucketThreads, bsm, st>>>(D, nn, P, sortWP, runs);
  } else {
    if (nt) binned_bucket<false, true><<<G, kBucketThreads, bsm, st>>>(D, nn, P, sortWP, runs);
    else binned_bucket<false, false><<<G, kBucketThreads, bsm, st>>>(D, nn, P, sortWP, runs);
  }
  RGN_HIP(hipGetLastError());
  stage_mark(st, "binned_bucket");
  binned_scan_rows<<<plan.NK, kRowThreads, 0, st>>>(runs, G, cvoff, cidx, nrun, totals, irun, IS, item);
  RGN_HIP(hipGetLastError());
  runs_clean = clean_after;
  stage_mark(st, "binned_scan_rows");
  // sources per lane with their loads in flight together
  if (env_int("RAGNAR_SUBHIST_BATCH", 4) == 8) {
    binned_subhist<8><<<static_cast<unsigned>(max_items), kSubThreads, subhist_smem(plan), st>>>(
      P, G, totals, cvoff, cidx, nrun, irun, IS, item, sortWP, partial);
  } else {
    binned_subhist<4><<<static_cast<unsigned>(max_items), kSubThreads, subhist_smem(plan), st>>>(
      P, G, totals, cvoff, cidx, nrun, irun, IS, item, sortWP, partial);
  }
  RGN_HIP(hipGetLastError());
  stage_mark(st, "binned_subhist");
  binned_prefix<<<plan.NK, kPrefThreads, sizeof(double) * 2 * M1, st>>>(plan.M, item, totals, partial, P0, P1);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "binned_prefix");
  binned_evaluate<<<plan.nbins, kThreads, 0, st>>>(P, P0, P1, out);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "binned_evaluate");
}

void sync_adjoint_binned(const SyncPlan& plan, const ParticlesDev& D, size_t n, const float* gbar, float* gw,
                         float* const gu[3], float* const gb[3], float* const ge[3], hipStream_t st) {
  RGN_REQUIRE(plan.binned_ok && n < kBinnedMaxParticles, RGN_ERR_UNSUPPORTED, "plan outside the binned limits");
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
    binned_adj_table<<<plan.NK, kThreads, 2 * sizeof(double) * std::max(plan.M, 1), st>>>(P, gbar, U);
    RGN_HIP(hipGetLastError());
    stage_mark(st, "binned_adj_table");
  }
  BinnedDev Pg = P;
  if (plan.NK == 0) Pg.NK = 0;  // every particle falls outside: gradients are zero
  const size_t tables = sizeof(double2) * kLog2Tab + sizeof(double) * std::max(plan.M, 1) +
                        sizeof(unsigned short) * plan.JQ;
  // The unstaged form is the general one also for stores without E: the
  // shorter one measured ~4% slower there (more resident waves, same memory
  // traffic).  The staged form drops the E terms when the store has no E and
  // dL/dE is not wanted (35 of the 253 vector instructions of the compiled
  // loop body, 195 -> 188 us at 2^24 particles).  Non-temporal
  // streams pay off once a call's particles and gradients outgrow the
  // device's last-level cache (256 MiB on MI355X): below that, inputs written
  // just before and outputs read just after are served from it.
  const size_t streams = 6 + (D.e[0] ? 3 : 0) + (D.w ? 1 : 0) + (G.gw ? 1 : 0) + (G.gu[0] ? 3 : 0) +
                         (G.gb[0] ? 3 : 0) + (G.ge[0] ? 3 : 0);
  const bool nt = sizeof(float) * streams * n > (size_t{256} << 20);
  // staged rows: as many as kStageBytes holds (and the histogram of NK buckets fits)
  const int kstage = static_cast<int>(std::min<size_t>(plan.NK, kStageBytes / (sizeof(double2) * (plan.M + 1))));
  // RAGNAR_GATHER_STAGE=0 reads every U entry from the table in memory
  const char* stage_env = std::getenv("RAGNAR_GATHER_STAGE");
  const bool stage = Pg.NK > 0 && kstage >= 1 && sizeof(unsigned) * plan.NK <= kStageBytes &&
                     !(stage_env && std::strcmp(stage_env, "0") == 0);
  if (stage) {
    const int grid = static_cast<int>(std::max<size_t>(
      1, std::min<size_t>(ceil_div(n, kGather), static_cast<size_t>(num_cus()) * env_int("RAGNAR_GATHER_SBPC", 1))));
    const size_t smem = kStageBytes + tables;
    const bool e_terms = D.e[0] || D.e[1] || D.e[2] || G.ge[0] || G.ge[1] || G.ge[2];
    if (e_terms) {
      if (nt) binned_adj_gather<true, true, true><<<grid, kGather, smem, st>>>(D, n, Pg, U, G, kstage);
      else binned_adj_gather<true, false, true><<<grid, kGather, smem, st>>>(D, n, Pg, U, G, kstage);
    } else {
      if (nt) binned_adj_gather<false, true, true><<<grid, kGather, smem, st>>>(D, n, Pg, U, G, kstage);
      else binned_adj_gather<false, false, true><<<grid, kGather, smem, st>>>(D, n, Pg, U, G, kstage);
    }
  } else {
    const size_t bpc = static_cast<size_t>(env_int("RAGNAR_GATHER_BPC", 8));  // blocks per CU
    const int grid = static_cast<int>(std::max<size_t>(1, std::min<size_t>(ceil_div(n, kThreads),
                                                                              static_cast<size_t>(num_cus()) * bpc)));
    if (nt) {
      binned_adj_gather<true, true, false><<<grid, kThreads, tables, st>>>(D, n, Pg, U, G, 0);
    } else {
      binned_adj_gather<true, false, false><<<grid, kThreads, tables, st>>>(D, n, Pg, U, G, 0);
    }
  }
  RGN_HIP(hipGetLastError());
  stage_mark(st, "binned_adj_gather");
}

}  // namespace rgn
