// Spatially resolved spectra: S[c, e] for ncells particle slices (CSR offsets)
// in one persistent launch plus a fixed-order combine, captured in a hipGraph.
//
// Balance.  The particle range [off[0], off[ncells]) is cut into G equal chunks
// of Q particles, one per block, regardless of cell boundaries, so ragged cell
// sizes (lognormal counts) cost nothing in balance.  Block b walks the cell
// segments that intersect its chunk; at the end of each segment it reduces its
// four waves and writes one record, at index c + b (the segments of block b
// are cells c_first(b)..c_last(b) and c_first(b+1) >= c_last(b), so c + b never
// collides).  The combine kernel sums, for cell c, the records of blocks
// b_lo(c)..b_hi(c) in order: the result is deterministic and equals the
// per-cell spectrum.
//
// Tuning.  The launch shape (blocks per CU in {1, 2, 4, 8, 16}, bins split
// over one or two waves) is chosen by timing every candidate on first use per
// shape class (device, bins, log2 of the cell and particle counts) and cached
// in a JSON file (RAGNAR_TUNE_CACHE, default ~/.cache/ragnar_amd/tune.json;
// RAGNAR_TUNE=0 disables tuning: 4 blocks per CU, no split).
// RAGNAR_CELLS_GRAPH=0 launches without the graph.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "emission.hpp"
#include "rgn_core.hpp"

namespace rgn {
namespace {

constexpr int kThreads = 256;

// SPLIT = 1: each wave takes 64 of a tile's 256 particles and all 64 NB bins
// (NB per lane).  SPLIT = 2: waves 2 g + h take particles [128 g, 128 g + 128)
// and bin half h (NB / 2 per lane): half the accumulators per lane, twice the
// particle reads per bin.
template <int NB, int SPLIT>
__global__ void __launch_bounds__(kThreads) cells_kernel(ParticlesDev D, PlanDev P,
                                                         const int64_t* __restrict__ off, int ncells,
                                                         int64_t Q, double* __restrict__ records) {
  static_assert(NB % SPLIT == 0, "bins per lane must split evenly");
  constexpr int NBL = NB / SPLIT;              // bins per lane
  constexpr int kPW = 64 * SPLIT;              // particles per wave and tile
  constexpr int kGroups = (kThreads / 64) / SPLIT;
  extern __shared__ __align__(16) unsigned char smem[];
  float4* tile = reinterpret_cast<float4*>(smem);
  double* red = reinterpret_cast<double*>(smem + kThreads * sizeof(float4));
  float2* ltab = reinterpret_cast<float2*>(smem + kThreads * sizeof(float4) + 4 * 64 * NB * sizeof(double));
  for (int j = threadIdx.x; j < P.TP; j += kThreads) ltab[j] = P.table[j];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int h = wave % SPLIT, grp = wave / SPLIT;
  float c[NBL], S[NBL];
#pragma unroll
  for (int k = 0; k < NBL; ++k) {
    const float2 cs = P.bincs[lane + 64 * (h * NBL + k)];
    c[k] = cs.x;
    S[k] = cs.y;
  }
  const int64_t first = off[0], last = off[ncells];
  const int64_t p0 = first + static_cast<int64_t>(blockIdx.x) * Q;
  const int64_t p1 = min(p0 + Q, last);
  if (p0 >= p1) return;  // whole block, uniform
  // last cell whose start is <= p0
  int lo = 0, hi = ncells - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= p0) lo = mid; else hi = mid - 1;
  }
  for (int cell = lo; cell < ncells && off[cell] < p1; ++cell) {
    const int64_t s0 = max(p0, off[cell]), s1 = min(p1, off[cell + 1]);
    if (s0 >= s1) continue;  // empty cell: no record
    double dacc[NBL];
#pragma unroll
    for (int k = 0; k < NBL; ++k) dacc[k] = 0.0;
    for (int64_t base = s0; base < s1; base += kThreads) {
      const int64_t p = base + threadIdx.x;
      Source s{0.0f, P.v_safe, 0.0f};
      if (p < s1) s = particle_source(D, static_cast<size_t>(p), P);
      __syncthreads();
      tile[threadIdx.x] = make_float4(s.W, s.vhi, s.vlo, 0.0f);
      __syncthreads();
      float acc[NBL];
#pragma unroll
      for (int k = 0; k < NBL; ++k) acc[k] = 0.0f;
      const float4* src = tile + grp * kPW;
#pragma unroll 4
      for (int j = 0; j < kPW; ++j) {
        const float4 q = src[j];
#pragma unroll
        for (int k = 0; k < NBL; ++k) {
          const float u = fmaf(c[k], S[k], q.y) + q.z;
          const float2 cs = ltab[cvt_sat_u32(u)];
          acc[k] = fmaf(q.x, fmaf(u, cs.y, cs.x), acc[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < NBL; ++k) dacc[k] += static_cast<double>(acc[k]);
    }
    // red[(g * NB + bin block) * 64 + lane]: the partial of particle group g
#pragma unroll
    for (int k = 0; k < NBL; ++k) red[(grp * NB + h * NBL + k) * 64 + lane] = dacc[k];
    __syncthreads();
    double* rec = records + static_cast<size_t>(cell + blockIdx.x) * (64 * NB);
    for (int e = threadIdx.x; e < 64 * NB; e += kThreads) {
      double t = red[e];
#pragma unroll
      for (int g = 1; g < kGroups; ++g) t += red[g * NB * 64 + e];  // fixed order
      rec[e] = t;
    }
    __syncthreads();  // red is reused by the next segment
  }
}

// One block per cell: out[c][e] = e_e * sum_{b_lo..b_hi} records[c + b][e].
__global__ void __launch_bounds__(kThreads) cells_combine_kernel(const double* __restrict__ records,
                                                                 const int64_t* __restrict__ off,
                                                                 int64_t Q, const float* __restrict__ bine,
                                                                 int nbins, int nbins_pad,
                                                                 float* __restrict__ out) {
  const int cell = blockIdx.x;
  const int64_t first = off[0], a = off[cell], b = off[cell + 1];
  for (int e = threadIdx.x; e < nbins; e += kThreads) {
    double s = 0.0;
    if (b > a) {
      const int64_t blo = (a - first) / Q, bhi = (b - 1 - first) / Q;
      for (int64_t blk = blo; blk <= bhi; ++blk) {
        s += records[static_cast<size_t>(cell + blk) * nbins_pad + e];
      }
    }
    const float ee = bine[e];
    out[static_cast<size_t>(cell) * nbins + e] =
      (ee == 0.0f || b <= a) ? 0.0f : static_cast<float>(static_cast<double>(ee) * s);
  }
}

size_t cells_smem(const SyncPlan& plan) {
  return kThreads * sizeof(float4) + 4 * 64 * plan.NB * sizeof(double) + plan.TP * sizeof(float2);
}

void launch_cells(const SyncPlan& plan, const ParticlesDev& D, const int64_t* d_off, int ncells,
                  int64_t Q, int grid, int split, double* records, float* out, hipStream_t st) {
  const PlanDev P = device_view(plan);
  const size_t smem = cells_smem(plan);
  RGN_REQUIRE(split == 1 || (split == 2 && plan.NB % 2 == 0), RGN_ERR_INVALID, "bin split must divide the bins");
  switch (plan.NB * 4 + split) {
#define RGN_CELLS_CASE(NB_, SP_)                                                             \
  case NB_ * 4 + SP_:                                                                        \
    cells_kernel<NB_, SP_><<<grid, kThreads, smem, st>>>(D, P, d_off, ncells, Q, records);  \
    break;
    RGN_CELLS_CASE(1, 1)
    RGN_CELLS_CASE(2, 1)
    RGN_CELLS_CASE(2, 2)
    RGN_CELLS_CASE(3, 1)
    RGN_CELLS_CASE(4, 1)
    RGN_CELLS_CASE(4, 2)
    RGN_CELLS_CASE(5, 1)
    RGN_CELLS_CASE(6, 1)
    RGN_CELLS_CASE(6, 2)
    RGN_CELLS_CASE(7, 1)
    RGN_CELLS_CASE(8, 1)
    RGN_CELLS_CASE(8, 2)
#undef RGN_CELLS_CASE
    default:
      throw Error(RGN_ERR_UNSUPPORTED, "at most 512 photon bins per plan");
  }
  RGN_HIP(hipGetLastError());
  cells_combine_kernel<<<ncells, kThreads, 0, st>>>(records, d_off, Q, plan.bine.as<float>(), plan.nbins,
                                                    plan.nbins_pad, out);
  RGN_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Autotuner: launch shape (blocks per CU, bin split) per shape class, cached
// in a JSON file of "key": value lines.  The key names the device (architecture
// and CU count); values are checked against the candidates when read, so a
// damaged or foreign file costs a re-tune, never a bad launch.  The file is
// replaced by rename(), so a concurrent reader sees the old or the new file.
// ---------------------------------------------------------------------------
std::mutex g_tune_mutex;
std::map<std::string, int> g_tune;
bool g_tune_loaded = false;
constexpr int kBpcCandidates[] = {1, 2, 4, 8, 16};

std::string tune_path() {
  if (const char* p = std::getenv("RAGNAR_TUNE_CACHE")) return p;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/ragnar_amd/tune.json";
}

void tune_load() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  std::ifstream in(tune_path());
  std::string line;
  while (std::getline(in, line)) {
    const auto q1 = line.find('"');
    if (q1 == std::string::npos) continue;
    const auto q2 = line.find('"', q1 + 1);
    if (q2 == std::string::npos) continue;
    const auto colon = line.find(':', q2);
    if (colon == std::string::npos) continue;
    g_tune[line.substr(q1 + 1, q2 - q1 - 1)] = std::atoi(line.c_str() + colon + 1);
  }
}

void tune_store() {
  const std::string path = tune_path();
  const auto slash = path.rfind('/');
  if (slash != std::string::npos) {
    std::string dir = path.substr(0, slash);
    for (size_t i = 1; i <= dir.size(); ++i) {
      if (i == dir.size() || dir[i] == '/') ::mkdir(dir.substr(0, i).c_str(), 0755);
    }
  }
  const std::string tmp = path + ".tmp" + std::to_string(static_cast<long>(::getpid()));
  {
    std::ofstream out(tmp);
    out << "{\n";
    size_t i = 0;
    for (const auto& kv : g_tune) {
      out << "  \"" << kv.first << "\": " << kv.second << (++i < g_tune.size() ? ",\n" : "\n");
    }
    out << "}\n";
    if (!out) return;  // a cache that cannot be written only costs a re-tune
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

// The cached shape of a class, if both entries are valid candidates.
bool tune_lookup(const std::string& key, int nb_per_lane, int& bpc, int& split) {
  const auto b = g_tune.find(key + "/bpc"), s = g_tune.find(key + "/split");
  if (b == g_tune.end() || s == g_tune.end()) return false;
  const bool bpc_ok = std::find(std::begin(kBpcCandidates), std::end(kBpcCandidates), b->second) !=
                      std::end(kBpcCandidates);
  const bool split_ok = s->second == 1 || (s->second == 2 && nb_per_lane % 2 == 0);
  if (!bpc_ok || !split_ok) return false;
  bpc = b->second;
  split = s->second;
  return true;
}

int log2_bucket(int64_t v) {
  int b = 0;
  while (v > 1) {
    v >>= 1;
    ++b;
  }
  return b;
}

// ---------------------------------------------------------------------------
// Graph cache keyed by every launch argument.
// ---------------------------------------------------------------------------
struct GraphKey {
  const void* plan;
  const void* ptrs[10];
  const void* off;
  const void* records;
  const void* out;
  int ncells;
  int64_t Q;
  int grid;
  int split;
  bool operator<(const GraphKey& o) const {
    return std::lexicographical_compare(reinterpret_cast<const char*>(this),
                                        reinterpret_cast<const char*>(this) + sizeof(GraphKey),
                                        reinterpret_cast<const char*>(&o),
                                        reinterpret_cast<const char*>(&o) + sizeof(GraphKey));
  }
};

std::mutex g_graph_mutex;
std::map<GraphKey, hipGraphExec_t> g_graphs;

}  // namespace

static void run_cells(const SyncPlan& plan, const ParticlesDev& D, const int64_t* h_off, int ncells,
                      float* out, hipStream_t st) {
  RGN_REQUIRE(ncells > 0, RGN_ERR_INVALID, "no cells");
  RGN_REQUIRE(h_off != nullptr && out != nullptr, RGN_ERR_INVALID, "null argument");
  for (int c = 0; c < ncells; ++c) {
    RGN_REQUIRE(h_off[c] <= h_off[c + 1], RGN_ERR_INVALID, "cell offsets must be non-decreasing");
  }
  RGN_REQUIRE(h_off[0] >= 0, RGN_ERR_INVALID, "negative cell offset");
  const int64_t total = h_off[ncells] - h_off[0];
  // offsets to a scratch buffer (outside any graph), then the launch
  int64_t* d_off = static_cast<int64_t*>(scratch(sizeof(int64_t) * (ncells + 1), 4));
  RGN_HIP(hipMemcpyAsync(d_off, h_off, sizeof(int64_t) * (ncells + 1), hipMemcpyHostToDevice, st));

  // tuned launch shape for this shape class
  const char* tune_env = std::getenv("RAGNAR_TUNE");
  const bool tune = !(tune_env && std::string(tune_env) == "0");
  std::ostringstream key;
  key << "cells:" << device_arch() << ":cu" << num_cus() << ":nb" << plan.nbins_pad << ":c"
      << log2_bucket(ncells) << ":n" << log2_bucket(total + 1);
  int bpc = 4, split = 1;
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    tune_load();
    cached = tune_lookup(key.str(), plan.NB, bpc, split);
  }
  auto plan_grid = [&](int blocks_per_cu, int64_t& Q) {
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>(static_cast<int64_t>(num_cus()) * blocks_per_cu,
                                                             ceil_div(std::max<int64_t>(total, 1), kThreads)));
    Q = std::max<int64_t>(1, static_cast<int64_t>(ceil_div(std::max<int64_t>(total, 1), g)));
    return static_cast<int>(ceil_div(std::max<int64_t>(total, 1), Q));
  };
  const size_t rec_bytes_max = sizeof(double) * plan.nbins_pad * (static_cast<size_t>(ncells) + num_cus() * 16 + 1);
  double* records = static_cast<double*>(scratch(rec_bytes_max, 3));

  if (tune && !cached && total > 0) {
    hipEvent_t e0, e1;
    RGN_HIP(hipEventCreate(&e0));
    RGN_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    for (int sp = 1; sp <= (plan.NB % 2 == 0 ? 2 : 1); ++sp) {
      for (int cand : kBpcCandidates) {
        int64_t Q;
        const int grid = plan_grid(cand, Q);
        launch_cells(plan, D, d_off, ncells, Q, grid, sp, records, out, st);  // warm
        RGN_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < 3; ++r) launch_cells(plan, D, d_off, ncells, Q, grid, sp, records, out, st);
        RGN_HIP(hipEventRecord(e1, st));
        RGN_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        RGN_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
          best_ms = ms;
          bpc = cand;
          split = sp;
        }
      }
    }
    RGN_HIP(hipEventDestroy(e0));
    RGN_HIP(hipEventDestroy(e1));
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    g_tune[key.str() + "/bpc"] = bpc;
    g_tune[key.str() + "/split"] = split;
    tune_store();
  }

  int64_t Q;
  const int grid = plan_grid(bpc, Q);
  const char* graph_env = std::getenv("RAGNAR_CELLS_GRAPH");
  const bool use_graph = !(graph_env && std::string(graph_env) == "0");
  if (!use_graph) {
    launch_cells(plan, D, d_off, ncells, Q, grid, split, records, out, st);
    return;
  }
  GraphKey gk{};
  gk.plan = &plan;
  for (int k = 0; k < 3; ++k) {
    gk.ptrs[k] = D.u[k];
    gk.ptrs[3 + k] = D.e[k];
    gk.ptrs[6 + k] = D.b[k];
  }
  gk.ptrs[9] = D.w;
  gk.off = d_off;
  gk.records = records;
  gk.out = out;
  gk.ncells = ncells;
  gk.Q = Q;
  gk.grid = grid;
  gk.split = split;
  std::lock_guard<std::mutex> lk(g_graph_mutex);
  auto it = g_graphs.find(gk);
  if (it == g_graphs.end()) {
    hipStream_t cap;
    RGN_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipGraph_t graph;
    RGN_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    launch_cells(plan, D, d_off, ncells, Q, grid, split, records, out, cap);
    RGN_HIP(hipStreamEndCapture(cap, &graph));
    hipGraphExec_t exec;
    RGN_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    RGN_HIP(hipGraphDestroy(graph));
    RGN_HIP(hipStreamDestroy(cap));
    if (g_graphs.size() >= 16) {
      RGN_HIP(hipGraphExecDestroy(g_graphs.begin()->second));
      g_graphs.erase(g_graphs.begin());
    }
    it = g_graphs.emplace(gk, exec).first;
  }
  RGN_HIP(hipGraphLaunch(it->second, st));
}

// Drops cached graphs of a plan that is being destroyed.
void forget_plan_graphs(const void* plan) {
  std::lock_guard<std::mutex> lk(g_graph_mutex);
  for (auto it = g_graphs.begin(); it != g_graphs.end();) {
    if (it->first.plan == plan) {
      (void)hipGraphExecDestroy(it->second);
      it = g_graphs.erase(it);
    } else {
      ++it;
    }
  }
}

}  // namespace rgn

using namespace rgn;

extern "C" int rgn_sync_cells(const rgn_sync_plan* plan, const float* const d_u[3],
                              const float* const d_e[3], const float* const d_b[3], const float* d_w,
                              const int64_t* h_offsets, int ncells, float* d_out, rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    RGN_REQUIRE(h_offsets != nullptr, RGN_ERR_INVALID, "null offsets");
    const size_t n = ncells > 0 ? static_cast<size_t>(h_offsets[ncells]) : 0;
    const ParticlesDev D = make_particles_checked(d_u, d_e, d_b, d_w, n);
    ScratchUse order(as_stream(stream));
    run_cells(plan->p, D, h_offsets, ncells, d_out, as_stream(stream));
  });
}
