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
};

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
