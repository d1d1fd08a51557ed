// Host-side plan of the emission kernels (shared by emission.hip and cells.hip).
#pragma once

#include <cmath>
#include <vector>

#include "emission_common.hpp"
#include "rgn_core.hpp"

namespace rgn {

struct SyncPlan {
  int nbins = 0;
  int nbins_pad = 0;  // 64 * NB
  int NB = 0;         // photon bins per lane
  int Z = 0;          // zero-tail origin of the u coordinate
  int TP = 0;         // padded table entries
  float B0 = 1.f, g_syn = 1.f, e_at = 1.f;
  double K = 0.0;          // table units per unit of log2 x
  double log2_xmin = 0.0;  // log2 of the table's smallest node
  double Amin = 0.0, Amax = 0.0;
  double win_lo = 0.0, win_hi = 0.0;
  float v_safe = 0.f;
  std::vector<float> bins;  // host copy of the photon bins
  DeviceBuffer table;       // float2[TP]
  DeviceBuffer bincs;       // float2[nbins_pad]
  DeviceBuffer bine;        // float[nbins_pad]

  // Factorised ("binned") form, binned.hip.  With B = k + phi (k integer) and
  // A_e = a_e + alpha_e, a source's value in bin e depends on k and on which
  // side of alpha_e its phi lies; the sorted distinct alpha_e cut [0, 1) into
  // M + 1 sub-cells.
  int T = 0;                       // table nodes
  int kmin = 0, NK = 0;            // integer parts of B that can contribute
  int M = 0;                       // distinct alpha_e
  int JQ = 0;                      // cells of the sub-cell lookup grid on [0, 1)
  bool binned_ok = false;          // the plan fits the factorised kernels' limits
  std::vector<double> A;           // A_e of live bins
  std::vector<char> live;          // bin e has e_e > 0
  DeviceBuffer alpha;              // double[M], sorted
  DeviceBuffer athr;               // float[M]: smallest float above alpha[i] (so, for a float
                                   //   phi, alpha[i] < phi exactly when athr[i] <= phi)
  DeviceBuffer jlut;               // uint16[JQ]: number of alphas below q / JQ (JQ >= 8 M)
  DeviceBuffer sublut;             // uint2[min(JQ, 2048)]: the sub-histogram's lookup cells
  DeviceBuffer rptr;               // int[M + 1]: live bins with alpha rank r are
  DeviceBuffer rlist;              //   rlist[rptr[r] .. rptr[r + 1]) (ascending e)
  DeviceBuffer binfo;              // BinInfo[nbins]
  DeviceBuffer seg;                // SegInfo[T - 1]
  DeviceBuffer l2tab;              // double2[256]: the factorised passes' log2 table
};

// The plan's emission constants with the correctly rounded reciprocals of the
// fixed divisors B0 and float(g_syn * g_syn) (host division: RN(1 / d)).
inline EmitConst emit_const(const SyncPlan& p) {
  EmitConst c{};
  c.B0 = p.B0;
  c.g_syn = p.g_syn;
  c.e_at = p.e_at;
  const float g2 = p.g_syn * p.g_syn;
  c.Dg = static_cast<double>(g2);
  c.rB0 = 1.0 / static_cast<double>(p.B0);
  c.rDg = 1.0 / c.Dg;
  c.fast = std::isnormal(static_cast<double>(p.B0)) && std::isnormal(c.Dg) && std::isnormal(c.rB0) &&
           std::isnormal(c.rDg);
  return c;
}

// Limits of the factorised kernels (LDS-resident per-bucket and per-sub-cell
// arrays, each kernel within 64 KB of LDS); plans beyond them use the direct
// kernels.
constexpr int kBinnedMaxM = 1024;
constexpr int kBinnedMaxNK = 3072;
constexpr size_t kBinnedMaxParticles = size_t{1} << 29;  // 32-bit byte offsets of 8-byte records

// Per photon bin of the factorised form.
struct BinInfo {
  double A;   // A_e
  int a;      // floor(A_e)
  int r;      // index of alpha_e in the sorted distinct list; -1 for dead bins
};

// Per table segment s (nodes s, s + 1 at their true coordinates tau):
// F(t) = y + (t - tau) * sigma on [tau_s, tau_{s+1}].
struct SegInfo {
  double tau, y, sigma;
};

void build_binned_tables(SyncPlan& plan, const std::vector<float>& xs, const std::vector<float>& ys);
void sync_forward_binned(const SyncPlan& plan, const ParticlesDev& D, size_t n, float* out,
                         hipStream_t st);
void sync_adjoint_binned(const SyncPlan& plan, const ParticlesDev& D, size_t n, const float* gbar,
                         float* gw, float* const gu[3], float* const gb[3], float* const ge[3],
                         hipStream_t st);
int resolve_method(const SyncPlan& plan, size_t n, int method, bool adjoint);

PlanDev device_view(const SyncPlan& plan);
size_t forward_smem(const SyncPlan& plan);
int forward_grid(size_t n, int blocks_per_cu);
void launch_combine(const SyncPlan& plan, const double* partials, int grid, float* out,
                    hipStream_t st);

// Particle pointers of the C ABI; they may be null only when n == 0.
ParticlesDev make_particles_checked(const float* const d_u[3], const float* const d_e[3],
                                    const float* const d_b[3], const float* d_w, size_t n);

// Drops the cached cell graphs that reference a plan being destroyed.
void forget_plan_graphs(const void* plan);

}  // namespace rgn

struct rgn_sync_plan {
  rgn::SyncPlan p;
};
