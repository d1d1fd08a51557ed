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
// Layout: one workgroup per output bin; its threads stride over the
// ng x nes pairs (the two distributions are a few KB and stay in cache), then
// a fixed-order tree reduction.  The pair count per bin is what the reference
// launches per bin; there is no table or plan.
#include <cmath>

#include "rgn_core.hpp"

namespace rgn {
namespace {

constexpr int kIcThreads = 256;

// ic.hpp:20-25.  2 * q * log(q) is float; the 0.5 literal makes the rest double.
__device__ __forceinline__ float kn_func(float Gamma, float q) {
  const float Gq = Gamma * q;
  const float first = 2.0f * q * logf(q);
  const double rest = static_cast<double>(1.0f - q) *
                      (static_cast<double>(1.0f + 2.0f * q) +
                       0.5 * static_cast<double>(Gq) * static_cast<double>(Gq) / static_cast<double>(1.0f + Gq));
  return static_cast<float>(static_cast<double>(first) + rest);
}

__global__ void __launch_bounds__(kIcThreads) ic_kernel(const float* __restrict__ g, const float* __restrict__ fg,
                                                        int ng, int g_log, const float* __restrict__ es,
                                                        const float* __restrict__ fes, int nes,
                                                        const float* __restrict__ eic, float* __restrict__ out) {
  __shared__ double red[kIcThreads];
  const int e = blockIdx.x;
  const float e_ic = eic[e];
  const long long npair = static_cast<long long>(ng) * nes;
  double acc = 0.0;
  for (long long p = threadIdx.x; p < npair; p += kIcThreads) {
    const int gi = static_cast<int>(p / nes), si = static_cast<int>(p - static_cast<long long>(gi) * nes);
    // ic.hpp:62-85, float arithmetic in the reference's order
    const float g_prtls = g[gi], f_prtls = fg[gi];
    const float e_soft = es[si], f_soft = fes[si];
    const float Gamma = 4.0f * g_prtls * e_soft;
    if (e_ic > g_prtls * Gamma / (1.0f + Gamma)) continue;
    const float q = (e_ic / g_prtls) / (Gamma * (1.0f - (e_ic / g_prtls)));
    const float kn = kn_func(Gamma, q);
    const float term = g_log ? f_prtls * (f_soft / e_soft) * e_ic * e_ic * kn / g_prtls
                             : f_prtls * (f_soft / e_soft) * e_ic * e_ic * kn / (g_prtls * g_prtls);
    acc += static_cast<double>(term);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kIcThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[e] = static_cast<float>(red[0]);
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
    ic_kernel<<<neic, kIcThreads, 0, st>>>(d_g, d_fg, ng, g_log ? 1 : 0, d_es, d_fes, nes, d_eic, d_out);
    RGN_HIP(hipGetLastError());
    stage_mark(st, "ic_kernel");
  });
}

}  // extern "C"
