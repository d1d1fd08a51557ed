// Particles<D>::energyDistribution (particles.cpp:189-260) on gfx950.
//
// Index function.  The reference bins an energy E (double) as
//   E < e_min -> 0,  E >= e_max -> n - 1,
//   else min(n - 1, (size_t)((float)(n - 1) * |log10(E / e_min)| / log10f(e_max / e_min)))
// (particles.cpp:233-245), with e_min / e_max the float MinMax of the bins.
// That function of E is monotone and piecewise constant, so the plan finds its
// n - 1 jump points E_k (the smallest double E with index >= k) on the host by
// bisection over the ordered bit patterns of positive doubles, evaluating the
// reference expression with the host libm.  The kernel takes a float log2
// estimate of the index and corrects it against E_k and E_{k+1}: the result
// equals the reference expression bit for bit, whatever the device log does.
//
// Accumulation.  Each wave owns a private LDS histogram (uint32 counts and FP64
// sums of the per-particle float terms); blocks write per-block partials that a
// second kernel combines in a fixed order.  Counts are exact integers; sums are
// the FP64 sum of the reference's float terms, rounded to float once.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>

#include "rgn_core.hpp"

struct rgn_hist_plan {
  int nb = 0;
  int log_bins = 0;
  float emin = 0.f, emax = 0.f;
  float approx_scale = 0.f;  // (nb - 1) / log2(emax / emin)
  std::vector<double> edges;  // nb + 1 entries: edges[k] = E_k for 1 <= k < nb
  rgn::DeviceBuffer d_edges;
};

namespace rgn {
namespace {

constexpr int kThreads = 256;

// particles.cpp:239-243, evaluated exactly as the reference does on the host.
size_t reference_index(double energy, float emin, float emax, size_t n) {
  if (energy < emin) return 0;
  if (energy >= emax) return n - 1;
  const auto ei = static_cast<std::size_t>(static_cast<float>(n - 1) *
                                           std::fabs(std::log10(energy / emin)) /
                                           std::log10(emax / emin));
  return ei > n - 1 ? n - 1 : ei;
}

uint64_t bits_of(double d) {
  uint64_t b;
  std::memcpy(&b, &d, sizeof b);
  return b;
}

double from_bits(uint64_t b) {
  double d;
  std::memcpy(&d, &b, sizeof d);
  return d;
}

void build_edges(rgn_hist_plan& plan) {
  const size_t n = plan.nb;
  plan.edges.assign(n + 1, std::numeric_limits<double>::infinity());
  plan.edges[0] = -std::numeric_limits<double>::infinity();
  if (n < 2) return;
  // All positive doubles in [emin, emax) are ordered like their bit patterns.
  const uint64_t lo = bits_of(static_cast<double>(plan.emin));
  const uint64_t hi = bits_of(static_cast<double>(plan.emax));  // exclusive
  for (size_t k = 1; k < n; ++k) {
    // smallest E in [emin, emax) with index >= k, or emax if none
    uint64_t a = lo, b = hi;
    while (a < b) {
      const uint64_t mid = a + (b - a) / 2;
      if (reference_index(from_bits(mid), plan.emin, plan.emax, n) >= k) {
        b = mid;
      } else {
        a = mid + 1;
      }
    }
    // E >= emax maps to n - 1 regardless; an edge at emax covers that.
    plan.edges[k] = from_bits(a);
  }
}

struct HistParams {
  const double* edges;  // in LDS
  int nb;
  float emin, emax, scale;
  float inv_emin;       // 1 / emin for the bin estimate (corrected against the edges)
  int log_bins, fourvel;
};

// 1 / x to about one ulp for finite x > 0 (hardware estimate, two Newton steps).
__device__ __forceinline__ double rcp_hist(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  return fma(y, fma(-x, y, 1.0), y);
}

// particles.cpp:228-252 for one particle: bin index and the float term.
__device__ __forceinline__ void bin_particle(float x, float y, float z, const HistParams& H,
                                             double* __restrict__ my_sum,
                                             unsigned* __restrict__ my_cnt) {
  const float usqr = x * x + y * y + z * z;  // float, as the reference
  const double energy = H.fourvel ? static_cast<double>(sqrtf(usqr))
                                  : sqrt(1.0 + static_cast<double>(usqr));
  int idx;
  if (energy < H.emin) {
    idx = 0;
  } else if (energy >= H.emax) {
    idx = H.nb - 1;
  } else if (energy == energy) {
    // a float estimate of the index; the edges make it exact.  Both edges of
    // the estimated bin come in one LDS read; only an estimate off by a bin
    // (energies within float rounding of an edge) walks further.
    const float est = H.scale * __builtin_amdgcn_logf(static_cast<float>(energy) * H.inv_emin);
    idx = static_cast<int>(fminf(fmaxf(est, 0.0f), static_cast<float>(H.nb - 1)));
    const double lo = H.edges[idx], hi = H.edges[idx + 1];
    if (!(lo <= energy && energy < hi)) {
      while (idx < H.nb - 1 && energy >= H.edges[idx + 1]) ++idx;
      while (idx > 0 && energy < H.edges[idx]) --idx;
    }
  } else {
    idx = H.nb - 1;  // NaN energy: the x86 reference's size_t conversion lands here
  }
  // particles.cpp:247-252: (float)(1 / e) for log-spaced bins.  A refined
  // reciprocal (about one double ulp) rounds to the same float except within
  // one double ulp of a float rounding boundary (~1e-9 of the terms); the
  // sums carry a 1e-6 tolerance, the counts do not depend on it.  The special
  // values follow 1.0 / e: 0 -> inf, inf -> 0, NaN -> NaN.
  float term = 1.0f;
  if (H.log_bins) {
    const bool normal = energy > 0.0 && energy < INFINITY;
    term = normal ? static_cast<float>(rcp_hist(energy))
                  : (energy == 0.0 ? INFINITY : (energy == INFINITY ? 0.0f : __builtin_nanf("")));
  }
  atomicAdd(my_cnt + idx, 1u);
  atomicAdd(my_sum + idx, static_cast<double>(term));
}

// Each thread keeps two 16-byte loads per component in flight (8 particles)
// when the three component arrays are 16-byte aligned; the remainder (and
// unaligned inputs) go through the scalar loop.
__global__ void __launch_bounds__(kThreads) hist_kernel(const float* __restrict__ ux,
                                                        const float* __restrict__ uy,
                                                        const float* __restrict__ uz, size_t n,
                                                        const double* __restrict__ g_edges, int nb,
                                                        float emin, float emax, float scale,
                                                        int log_bins, int fourvel, int vec, int qper,
                                                        unsigned* __restrict__ cpart,
                                                        double* __restrict__ partials) {
  extern __shared__ __align__(16) unsigned char smem[];
  // 4 * qper private sub-histograms per block: one per wave and lane group of
  // 64 / qper lanes (fewer same-address atomics where energies pile up in a
  // few bins); pitch nb + 1 staggers the copies over the LDS banks
  const int copies = 4 * qper, pitch = nb + 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* lsum = reinterpret_cast<double*>(smem);                                 // [copies][pitch]
  double* ledges = lsum + copies * pitch;                                         // [nb + 1]
  unsigned* lcnt = reinterpret_cast<unsigned*>(ledges + nb + 1);                  // [copies][pitch]
  for (int i = threadIdx.x; i < copies * pitch; i += kThreads) {
    lsum[i] = 0.0;
    lcnt[i] = 0u;
  }
  for (int i = threadIdx.x; i <= nb; i += kThreads) ledges[i] = g_edges[i];
  __syncthreads();
  const HistParams H{ledges, nb, emin, emax, scale, 1.0f / emin, log_bins, fourvel};
  const int copy = wave * qper + lane / (64 / qper);
  double* my_sum = lsum + copy * pitch;
  unsigned* my_cnt = lcnt + copy * pitch;
  const size_t stride = static_cast<size_t>(gridDim.x) * kThreads;
  const size_t tid = blockIdx.x * static_cast<size_t>(kThreads) + threadIdx.x;
  size_t scalar_from = 0;
  if (vec) {
    const size_t nq = n / 4;
    const float4* x4 = reinterpret_cast<const float4*>(ux);
    const float4* y4 = reinterpret_cast<const float4*>(uy);
    const float4* z4 = reinterpret_cast<const float4*>(uz);
    for (size_t q = tid; q < nq; q += 2 * stride) {
      const bool two = q + stride < nq;
      const float4 a0 = x4[q], b0 = y4[q], c0 = z4[q];
      float4 a1 = a0, b1 = b0, c1 = c0;
      if (two) {
        a1 = x4[q + stride];
        b1 = y4[q + stride];
        c1 = z4[q + stride];
      }
      bin_particle(a0.x, b0.x, c0.x, H, my_sum, my_cnt);
      bin_particle(a0.y, b0.y, c0.y, H, my_sum, my_cnt);
      bin_particle(a0.z, b0.z, c0.z, H, my_sum, my_cnt);
      bin_particle(a0.w, b0.w, c0.w, H, my_sum, my_cnt);
      if (two) {
        bin_particle(a1.x, b1.x, c1.x, H, my_sum, my_cnt);
        bin_particle(a1.y, b1.y, c1.y, H, my_sum, my_cnt);
        bin_particle(a1.z, b1.z, c1.z, H, my_sum, my_cnt);
        bin_particle(a1.w, b1.w, c1.w, H, my_sum, my_cnt);
      }
    }
    scalar_from = 4 * nq;
  }
  for (size_t p = scalar_from + tid; p < n; p += stride) {
    bin_particle(ux[p], uy[p], uz[p], H, my_sum, my_cnt);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += kThreads) {
    unsigned c = 0;
    double s = 0.0;
    for (int q = 0; q < copies; ++q) {  // fixed order
      c += lcnt[q * pitch + i];
      s += lsum[q * pitch + i];
    }
    cpart[static_cast<size_t>(i) * gridDim.x + blockIdx.x] = c;
    partials[static_cast<size_t>(i) * gridDim.x + blockIdx.x] = s;
  }
}

// Fixed-order sums of the per-block partials: FP64 sums and (when counts is
// not null) uint64 counts.  No global atomics.
__global__ void __launch_bounds__(kThreads) hist_combine_kernel(const double* __restrict__ partials,
                                                                const unsigned* __restrict__ cpart,
                                                                int nblocks, int nb,
                                                                float* __restrict__ out,
                                                                unsigned long long* __restrict__ counts) {
  const int i = blockIdx.x;
  if (i >= nb) return;
  double s = 0.0;
  unsigned long long c = 0;
  for (int b = threadIdx.x; b < nblocks; b += kThreads) {
    s += partials[static_cast<size_t>(i) * nblocks + b];
    c += cpart[static_cast<size_t>(i) * nblocks + b];
  }
  __shared__ double red[kThreads];
  __shared__ unsigned long long cred[kThreads];
  red[threadIdx.x] = s;
  cred[threadIdx.x] = c;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
      red[threadIdx.x] += red[threadIdx.x + w];
      cred[threadIdx.x] += cred[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[i] = static_cast<float>(red[0]);
    if (counts) counts[i] = cred[0];
  }
}

// Launch shape of one histogram call.
struct HistShape {
  int grid = 1;
  size_t smem = 0;
  int qper = 1;
  int vec = 0;
};

HistShape hist_shape(const rgn_hist_plan& plan, const float* const d_u[3], size_t n) {
  HistShape sh;
  const int nb = plan.nb;
  auto aligned16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  sh.vec = aligned16(d_u[0]) && aligned16(d_u[1]) && aligned16(d_u[2]);
  const size_t per_thread = sh.vec ? 8 : 1;
  const size_t tiles = ceil_div(std::max<size_t>(n, 1), kThreads * per_thread);
  // lane groups per wave: as many private copies as fit in 24 KB of LDS (at
  // 200 bins two per wave, and seven blocks share a CU: 51 -> 50 us at 2^24
  // against four per wave in 48 KB)
  const size_t per_copy = static_cast<size_t>(nb + 1) * (sizeof(double) + sizeof(unsigned));
  sh.qper = 4;
  while (sh.qper > 1 && 4 * sh.qper * per_copy + (nb + 1) * sizeof(double) > 24 * 1024) sh.qper /= 2;
  sh.smem = 4 * sh.qper * per_copy + (nb + 1) * sizeof(double);
  // one round of blocks: as many per CU as the LDS holds (at most 8), so the
  // per-block clearing and reduction of the private copies is paid once
  const size_t per_cu = std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / sh.smem));
  sh.grid = static_cast<int>(std::min<size_t>(tiles, static_cast<size_t>(num_cus()) * per_cu));
  return sh;
}

// Bytes of per-block partials a launch of `grid` blocks writes (FP64 sums and
// uint32 counts per bin and block).  Every entry is written before it is read.
size_t hist_partial_bytes(int nb, int grid) {
  return (sizeof(double) + sizeof(unsigned)) * static_cast<size_t>(nb) * grid;
}

void launch_hist(const rgn_hist_plan& plan, const float* const d_u[3], size_t n, int fourvel,
                 float* d_hist, uint64_t* d_counts, const HistShape& sh, double* partials,
                 hipStream_t st) {
  const int nb = plan.nb;
  unsigned* cpart = reinterpret_cast<unsigned*>(partials + static_cast<size_t>(nb) * sh.grid);
  hist_kernel<<<sh.grid, kThreads, sh.smem, st>>>(
    d_u[0], d_u[1], d_u[2], n, plan.d_edges.as<double>(), nb, plan.emin, plan.emax,
    plan.approx_scale, plan.log_bins, fourvel ? 1 : 0, sh.vec, sh.qper, cpart, partials);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "hist_kernel");
  hist_combine_kernel<<<nb, kThreads, 0, st>>>(partials, cpart, sh.grid, nb, d_hist,
                                               reinterpret_cast<unsigned long long*>(d_counts));
  RGN_HIP(hipGetLastError());
  stage_mark(st, "hist_combine_kernel");
}

}  // namespace
}  // namespace rgn

// Per-block partials of rgn_energy_hist_noalloc: a fixed device array (each
// device holds its own copy until the process ends), used instead of the
// library's allocated scratch.  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device (noalloc_partials below).
constexpr size_t kHistNoAllocBytes = size_t{8} << 20;
static __device__ __align__(256) unsigned char scratch[kHistNoAllocBytes];

namespace {
std::mutex g_noalloc_mutex;
constexpr int kMaxDevices = 64;
void* g_noalloc_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device and kept.
void* noalloc_partials() {
  int dev = 0;
  RGN_HIP(hipGetDevice(&dev));
  RGN_REQUIRE(dev >= 0 && dev < kMaxDevices, RGN_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mutex);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    RGN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    RGN_REQUIRE(p != nullptr, RGN_ERR_HIP, "no device address for the histogram's fixed scratch");
    g_noalloc_ptr[dev] = p;
  }
  return g_noalloc_ptr[dev];
}
}  // namespace

using namespace rgn;

extern "C" {

int rgn_hist_plan_create(const float* h_bins, int nb, int log_bins, rgn_hist_plan** out) {
  return guarded([&] {
    RGN_REQUIRE(out != nullptr && h_bins != nullptr, RGN_ERR_INVALID, "null argument");
    RGN_REQUIRE(nb >= 1, RGN_ERR_INVALID, "no energy bins");
    // LDS of the smallest configuration: 4 private copies (pitch nb + 1) and the edges
    RGN_REQUIRE(static_cast<size_t>(nb + 1) * (4 * (sizeof(double) + sizeof(unsigned)) + sizeof(double)) <=
                  150 * 1024,
                RGN_ERR_UNSUPPORTED, "too many energy bins for the LDS histogram (at most 2741)");
    auto* plan = new rgn_hist_plan();
    plan->nb = nb;
    plan->log_bins = log_bins ? 1 : 0;
    // particles.cpp:200-215 ("EnergyBinsMinMax")
    float lo = h_bins[0], hi = h_bins[0];
    for (int i = 0; i < nb; ++i) {
      lo = h_bins[i] < lo ? h_bins[i] : lo;
      hi = h_bins[i] > hi ? h_bins[i] : hi;
    }
    plan->emin = lo;
    plan->emax = hi;
    plan->approx_scale =
      (lo > 0.0f && hi > lo) ? static_cast<float>((nb - 1) / std::log2(static_cast<double>(hi) / lo))
                             : 0.0f;
    try {
      if (lo > 0.0f && hi > lo) {
        build_edges(*plan);
      } else {
        // degenerate bins: every finite energy lands in bin 0 or n - 1
        plan->edges.assign(nb + 1, std::numeric_limits<double>::infinity());
        plan->edges[0] = -std::numeric_limits<double>::infinity();
      }
      plan->d_edges.reset(plan->edges.size() * sizeof(double));
      RGN_HIP(hipMemcpy(plan->d_edges.ptr, plan->edges.data(), plan->d_edges.bytes,
                        hipMemcpyHostToDevice));
    } catch (...) {
      delete plan;
      throw;
    }
    *out = plan;
  });
}

int rgn_hist_plan_destroy(rgn_hist_plan* plan) {
  delete plan;
  return RGN_OK;
}

int rgn_energy_hist(const rgn_hist_plan* plan, const float* const d_u[3], size_t n, int fourvel,
                    float* d_hist, uint64_t* d_counts, rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    RGN_REQUIRE(d_u && d_u[0] && d_u[1] && d_u[2] && d_hist, RGN_ERR_INVALID, "null argument");
    hipStream_t st = as_stream(stream);
    ScratchUse order(st);
    const HistShape sh = hist_shape(*plan, d_u, n);
    double* partials = static_cast<double*>(rgn::scratch(hist_partial_bytes(plan->nb, sh.grid), 2));
    launch_hist(*plan, d_u, n, fourvel, d_hist, d_counts, sh, partials, st);
  });
}

int rgn_energy_hist_noalloc(const rgn_hist_plan* plan, const float* const d_u[3], size_t n,
                            int fourvel, float* d_hist, uint64_t* d_counts, rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    RGN_REQUIRE(d_u && d_u[0] && d_u[1] && d_u[2] && d_hist, RGN_ERR_INVALID, "null argument");
    HistShape sh = hist_shape(*plan, d_u, n);
    // the grid is cut to what the fixed array holds: bins x blocks x 12 bytes
    const size_t per_block = hist_partial_bytes(plan->nb, 1);
    RGN_REQUIRE(per_block <= kHistNoAllocBytes, RGN_ERR_UNSUPPORTED,
                "energy bins beyond the fixed scratch of rgn_energy_hist_noalloc");
    sh.grid = static_cast<int>(std::min<size_t>(sh.grid, kHistNoAllocBytes / per_block));
    double* partials = static_cast<double*>(noalloc_partials());
    launch_hist(*plan, d_u, n, fourvel, d_hist, d_counts, sh, partials, as_stream(stream));
  });
}

// Jump points of the index function (nb + 1 doubles), for tests.
int rgn_hist_plan_edges(const rgn_hist_plan* plan, double* h_edges) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr && h_edges != nullptr, RGN_ERR_INVALID, "null argument");
    std::copy(plan->edges.begin(), plan->edges.end(), h_edges);
  });
}

}  // extern "C"
