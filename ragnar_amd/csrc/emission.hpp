// Host-side plan of the emission kernels (shared by emission.hip and cells.hip).
#pragma once

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
  std::vector<double> A;           // A_e of live bins
  std::vector<char> live;          // bin e has e_e > 0
  DeviceBuffer alpha;              // double[M], sorted
  DeviceBuffer binfo;              // BinInfo[nbins]
  DeviceBuffer seg;                // SegInfo[T - 1]
};

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
int resolve_method(const SyncPlan& plan, size_t n, int method);

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
