// Inverse-Compton spectrum of a tabulated particle distribution scattering a
// tabulated soft-photon field (ICSpectrum, src/physics/ic.cpp:15-46, with the
// per-triple kernel of src/physics/ic.hpp:28-86).
//
// For every output energy e_ic and every (gamma, e_soft) pair the reference adds
//     f_g (f_s / e_s) e_ic^2 KN(Gamma, q) / g      (/ g^2 for linear gamma bins)
// with Gamma = 4 g e_s, q = (e_ic / g) / (Gamma (1 - e_ic / g)), skipped above
// the kinematic limit e_ic > g Gamma / (1 + Gamma).  Each term is computed here in
// the reference's float arithmetic and operation order (KN's second part is
// double in the reference: its 0.5 literal promotes it), and the terms of a bin
// are summed in FP64 in a fixed order, so results are reproducible bit for bit.
//
// Layout: one workgroup per (output bin, chunk of gamma bins), enough of them
// to fill the device; the threads take the soft-photon bins (the two
// distributions are a few KB and stay in cache); a fixed-order tree per block
// and a second pass adding the chunks in order.  No table or plan.  The kernel
// is bound by its VALU work (IEEE float and double divides), so the quotients
// that depend on one operand only -- e_ic / g per gamma bin, f_s / e_s per
// soft-photon bin -- are formed once (the same float values the reference
// forms per triple) and the triples loop over gamma bins staged in LDS.
#include <algorithm>
#include <cmath>

#include "rgn_core.hpp"

namespace rgn {
namespace {

constexpr int kIcThreads = 256;
constexpr int kIcStage = 1024;  // gamma bins staged in LDS at a time

// ic.hpp:20-25.  2 * q * log(q) is float; the 0.5 literal makes the rest double.
__device__ __forceinline__ float kn_func(float Gamma, float q) {
  const float Gq = Gamma * q;
  const float first = 2.0f * q * logf(q);
  const double rest = static_cast<double>(1.0f - q) *
                      (static_cast<double>(1.0f + 2.0f * q) +
                       0.5 * static_cast<double>(Gq) * static_cast<double>(Gq) / static_cast<double>(1.0f + Gq));
  return static_cast<float>(static_cast<double>(first) + rest);
}

// One workgroup per (output bin, chunk of gamma bins): its threads take the
// soft-photon bins, and each thread runs its bins over the chunk's gamma bins,
// staged kIcStage at a time; the block's FP64 partial goes to
// partial[chunk][bin].  Summation order: per thread, stage by stage, soft-photon
// bin by bin, gamma bin by gamma bin; then a fixed tree over the threads.
__global__ void __launch_bounds__(kIcThreads) ic_kernel(const float* __restrict__ g, const float* __restrict__ fg,
                                                        int ng, int g_log, const float* __restrict__ es,
                                                        const float* __restrict__ fes, int nes,
                                                        const float* __restrict__ eic, int gper,
                                                        double* __restrict__ partial) {
  __shared__ double red[kIcThreads];
  __shared__ float s_g[kIcStage], s_f[kIcStage], s_x[kIcStage];
  const int e = blockIdx.x, neic = gridDim.x;
  const float e_ic = eic[e];
  const int g0 = blockIdx.y * gper, g1 = min(ng, g0 + gper);
  double acc = 0.0;
  for (int gs = g0; gs < g1; gs += kIcStage) {
    const int m = min(kIcStage, g1 - gs);
    __syncthreads();  // the previous stage has been read
    for (int i = threadIdx.x; i < m; i += kIcThreads) {
      const float gp = g[gs + i];
      s_g[i] = gp;
      s_f[i] = fg[gs + i];
      s_x[i] = e_ic / gp;
    }
    __syncthreads();
    for (int si = threadIdx.x; si < nes; si += kIcThreads) {
      // ic.hpp:62-85, float arithmetic in the reference's order
      const float e_soft = es[si];
      const float fs_es = fes[si] / e_soft;
      for (int i = 0; i < m; ++i) {
        const float g_prtls = s_g[i], f_prtls = s_f[i], x = s_x[i];
        const float Gamma = 4.0f * g_prtls * e_soft;
        if (e_ic > g_prtls * Gamma / (1.0f + Gamma)) continue;
        const float q = x / (Gamma * (1.0f - x));
        const float kn = kn_func(Gamma, q);
        const float term = g_log ? f_prtls * fs_es * e_ic * e_ic * kn / g_prtls
                                 : f_prtls * fs_es * e_ic * e_ic * kn / (g_prtls * g_prtls);
        acc += static_cast<double>(term);
      }
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kIcThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[static_cast<size_t>(blockIdx.y) * neic + e] = red[0];
}

// out[e] = the chunk partials of bin e, summed in chunk order.
__global__ void __launch_bounds__(kIcThreads) ic_combine(const double* __restrict__ partial, int nchunk, int neic,
                                                         float* __restrict__ out) {
  const int e = blockIdx.x * kIcThreads + threadIdx.x;
  if (e >= neic) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += partial[static_cast<size_t>(c) * neic + e];
  out[e] = static_cast<float>(s);
}

}  // namespace
}  // namespace rgn

using namespace rgn;

extern "C" {

int rgn_ic_spectrum(const float* d_g, const float* d_fg, int ng, int g_log, const float* d_es,
                    const float* d_fes, int nes, const float* d_eic, int neic, float* d_out,
                    rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(ng >= 0 && nes >= 0 && neic >= 0, RGN_ERR_INVALID, "negative size");
    if (neic == 0) return;
    RGN_REQUIRE(d_eic && d_out, RGN_ERR_INVALID, "null argument");
    RGN_REQUIRE((ng == 0 || (d_g && d_fg)) && (nes == 0 || (d_es && d_fes)), RGN_ERR_INVALID,
                "null argument");
    hipStream_t st = as_stream(stream);
    ScratchUse order(st);
    // enough workgroups to fill the device: gamma bins split into chunks
    const int chunks = std::min(std::max(1, (8 * num_cus() + neic - 1) / neic), std::max(ng, 1));
    const int gper = std::max(1, (ng + chunks - 1) / chunks);
    const int nchunk = std::max(1, (ng + gper - 1) / gper);  // at most chunks <= 8 CUs + 1
    double* partial = static_cast<double*>(scratch(sizeof(double) * nchunk * static_cast<size_t>(neic), 17));
    ic_kernel<<<dim3(neic, nchunk), kIcThreads, 0, st>>>(d_g, d_fg, ng, g_log ? 1 : 0, d_es, d_fes, nes, d_eic,
                                                         gper, partial);
    RGN_HIP(hipGetLastError());
    stage_mark(st, "ic_kernel");
    ic_combine<<<(neic + kIcThreads - 1) / kIcThreads, kIcThreads, 0, st>>>(partial, nchunk, neic, d_out);
    RGN_HIP(hipGetLastError());
    stage_mark(st, "ic_combine");
  });
}

}  // extern "C"
