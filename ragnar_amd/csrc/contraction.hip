// Batched distribution -> spectrum contraction S[d, e] = sum_g f[d, g] K[g, e]
// for nd distributions on one gbeta grid: KernelFromDist (synchrotron.hpp:72-95)
// applied to every row, written as a GEMM.
//
// K[g, e] = e_e * (gbeta_g if the distribution bins are log spaced) *
// F(e_e / eps_g), eps_g = e_at gbeta_g^2 / g_syn^2, is formed on the GPU with
// the same interpolation coordinate and segment table as the emission kernels
// (dist_kernel_value).  The GEMM then runs on the matrix cores:
//   RGN_PREC_FP32  rocBLAS SGEMM on K built in full (kernel_matrix; a plain
//                  library GEMM, FP32 throughout);
//   RGN_PREC_F16X3 / _F16X1 / _BF16X6: hand-written MFMA kernels
//                  (contraction_f16.hip), whose packing kernels form K's
//                  entries themselves, FP32 accumulation.
#include <rocblas/rocblas.h>

#include <mutex>

#include "emission.hpp"
#include "rgn_core.hpp"

namespace rgn {

// Split-FP16 MFMA form, contraction_f16.hip.
void contraction_f16(const SyncPlan& plan, const float* d_gbeta, int ng, int dist_log, const float* d_f, size_t nd,
                     int nprod, float* d_out, hipStream_t st);
// Split-BF16 form: rows from d_f, or generated when d_index is not null.
void contraction_bf16(const SyncPlan& plan, const float* d_gbeta, int ng, int dist_log, const float* d_f,
                      const float* d_index, const float* d_cutoff, size_t nd, float* d_out, hipStream_t st);
// The same with the rows generated from (index, cutoff) inside the kernel.
void contraction_f16_plaw(const SyncPlan& plan, const float* d_gbeta, int ng, int dist_log, const float* d_index,
                          const float* d_cutoff, size_t nd, int nprod, float* d_out, hipStream_t st);

namespace {

__global__ void __launch_bounds__(256) kernel_matrix(PlanDev P, const float* __restrict__ gbeta,
                                                     int ng, int dist_log, float* __restrict__ K) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  const int g = blockIdx.y;
  if (e >= P.nbins) return;
  K[static_cast<size_t>(g) * P.nbins + e] = dist_kernel_value(dist_source(gbeta[g], dist_log, P), e, P);
}

std::mutex g_blas_mutex;
rocblas_handle g_blas = nullptr;

rocblas_handle blas() {
  if (!g_blas) {
    if (rocblas_create_handle(&g_blas) != rocblas_status_success) {
      throw Error(RGN_ERR_HIP, "rocblas_create_handle failed");
    }
  }
  return g_blas;
}

}  // namespace

void build_kernel_matrix(const SyncPlan& plan, const float* d_gbeta, int ng, int dist_log,
                         float* d_k, hipStream_t st) {
  const dim3 grid(static_cast<unsigned>(ceil_div(plan.nbins, 256)), static_cast<unsigned>(ng));
  kernel_matrix<<<grid, 256, 0, st>>>(device_view(plan), d_gbeta, ng, dist_log, d_k);
  RGN_HIP(hipGetLastError());
}

}  // namespace rgn

using namespace rgn;

extern "C" int rgn_sync_fromdist_batched(const rgn_sync_plan* plan, const float* d_gbeta, int ng,
                                         int dist_log, const float* d_f, size_t nd, int precision,
                                         float* d_out, rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    RGN_REQUIRE(d_gbeta && (nd == 0 || (d_f && d_out)), RGN_ERR_INVALID, "null argument");
    RGN_REQUIRE(ng > 0, RGN_ERR_INVALID, "empty distribution grid");
    const SyncPlan& p = plan->p;
    hipStream_t st = as_stream(stream);
    ScratchUse order(st);
    if (precision == RGN_PREC_FP32) {
      float* d_k = static_cast<float*>(scratch(sizeof(float) * static_cast<size_t>(ng) * p.nbins, 5));
      build_kernel_matrix(p, d_gbeta, ng, dist_log, d_k, st);
      if (nd == 0) return;
      std::lock_guard<std::mutex> lk(g_blas_mutex);
      rocblas_handle h = blas();
      RGN_REQUIRE(rocblas_set_stream(h, st) == rocblas_status_success, RGN_ERR_HIP,
                  "rocblas_set_stream failed");
      const float one = 1.0f, zero = 0.0f;
      // row-major out[nd][nb] = f[nd][ng] K[ng][nb]  <=>  column-major out^T = K^T f^T
      const rocblas_status s = rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none,
                                             p.nbins, static_cast<rocblas_int>(nd), ng, &one, d_k,
                                             p.nbins, d_f, ng, &zero, d_out, p.nbins);
      RGN_REQUIRE(s == rocblas_status_success, RGN_ERR_HIP, "rocblas_sgemm failed");
    } else if (precision == RGN_PREC_F16X3 || precision == RGN_PREC_F16X1) {
      contraction_f16(p, d_gbeta, ng, dist_log, d_f, nd, precision == RGN_PREC_F16X3 ? 3 : 1, d_out, st);
    } else if (precision == RGN_PREC_BF16X6) {
      contraction_bf16(p, d_gbeta, ng, dist_log, d_f, nullptr, nullptr, nd, d_out, st);
    } else {
      throw Error(RGN_ERR_INVALID, "unknown precision mode");
    }
  });
}

extern "C" int rgn_sync_fromdist_batched_plaw(const rgn_sync_plan* plan, const float* d_gbeta, int ng,
                                              int dist_log, const float* d_index, const float* d_cutoff,
                                              size_t nd, int precision, float* d_out, rgn_stream_t stream) {
  return guarded([&] {
    RGN_REQUIRE(plan != nullptr, RGN_ERR_INVALID, "null plan");
    RGN_REQUIRE(d_gbeta && (nd == 0 || (d_out && d_index && d_cutoff)), RGN_ERR_INVALID, "null argument");
    RGN_REQUIRE(ng >= 2, RGN_ERR_INVALID, "a generated distribution needs at least two grid points");
    RGN_REQUIRE(ng <= 2048, RGN_ERR_RANGE, "generated distributions take at most 2048 grid points");
    RGN_REQUIRE(precision == RGN_PREC_F16X3 || precision == RGN_PREC_F16X1 || precision == RGN_PREC_BF16X6,
                RGN_ERR_UNSUPPORTED,
                "generated rows run on the matrix-core kernels only (RGN_PREC_F16X3, _F16X1 or _BF16X6)");
    const SyncPlan& p = plan->p;
    hipStream_t st = as_stream(stream);
    ScratchUse order(st);
    if (precision == RGN_PREC_BF16X6) {
      contraction_bf16(p, d_gbeta, ng, dist_log, nullptr, d_index, d_cutoff, nd, d_out, st);
    } else {
      contraction_f16_plaw(p, d_gbeta, ng, dist_log, d_index, d_cutoff, nd, precision == RGN_PREC_F16X3 ? 3 : 1,
                           d_out, st);
    }
  });
}
