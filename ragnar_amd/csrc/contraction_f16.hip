// Split-FP16 matrix-core form of the batched contraction
//     S[d, e] = sum_g f[d, g] K[g, e]
// on gfx950 (v_mfma_f32_32x32x16_f16, FP32 accumulation).
//
// Precision.  Each operand is split per block of 64 consecutive g into
//     x = 2^-s (hi + lo),  hi = fp16(x 2^s),  lo = fp16(x 2^s - hi),
// with one power-of-two exponent s per (row of f, g block) and per (g block,
// column of K) chosen so that the block's largest |x| 2^s lies in
// [2^14, 2^15): hi and lo together carry 22 significant bits for every element
// within 2^18 of its block maximum, and fp16 range is never exceeded.  The
// three-product form (F16X3) accumulates hi*hi + hi*lo + lo*hi per block in
// FP32 (every fp16 x fp16 product is exact in FP32), then adds the block's
// partial tile scaled by 2^-(s_row + s_col) into the FP32 result; the dropped
// lo*lo term is ~2^-22 relative.  F16X1 uses hi*hi only (~2^-11).
//
// Tiling.  A workgroup computes a 128 x 256 tile of S with eight waves, each a
// 64 x 64 quadrant held as 2 x 2 MFMA 32x32 tiles (f is read from HBM
// nbins / 256 times).  Per g block, the f tile (128 rows x 64 g, 16-byte
// loads, four lanes per row) is split and parked in LDS as hi/lo; K is split
// once per call into a packed [g block][column][64 g] layout whose tile is one
// contiguous range, copied into LDS.  The next block's f and K are loaded into
// registers while the matrix cores work on the current one.  An MFMA operand fragment is 8 consecutive g of
// one row (A) or one column (B): both operands use the same lane -> g
// assignment, so the sum over g is the same whatever order the hardware pairs
// the 16 g of one instruction.
#include <hip/hip_runtime.h>

#include "emission.hpp"
#include "rgn_core.hpp"

namespace rgn {

namespace {

constexpr int kTM = 128;        // rows (distributions) per workgroup
constexpr int kTN = 256;        // columns (photon bins) per workgroup
constexpr int kKB = 64;         // g per block (one scale exponent)
constexpr int kLd = kKB + 8;    // LDS row pitch in halves (bank spread)
constexpr int kThreadsC = 512;  // eight waves: 2 x 4 quadrants of 64 x 64
constexpr int kBChunks = kTN * kKB / 8 / kThreadsC;  // 16-byte K chunks per thread per half

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));

// Exponent s with m 2^s in [2^14, 2^15) (0 for m == 0).
__device__ __forceinline__ int block_exp(float m) { return m > 0.0f ? 14 - ilogbf(m) : 0; }

__device__ __forceinline__ void split2(float x, int s, _Float16& hi, _Float16& lo) {
  const float xs = ldexpf(x, s);
  hi = static_cast<_Float16>(xs);
  lo = static_cast<_Float16>(xs - static_cast<float>(hi));  // the difference is exact
}

// K[g][e] (ng x nb, FP32) -> Bh/Bl[(kb * nbp + c) * 64 + i] (g = 64 kb + i), and
// the exponent of each (kb, c).  One thread per (column, g block).
__global__ void __launch_bounds__(256) pack_kernel_matrix(const float* __restrict__ K, int ng, int nb, int nbp,
                                                          _Float16* __restrict__ Bh, _Float16* __restrict__ Bl,
                                                          int* __restrict__ sB) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int kb = blockIdx.y;
  if (c >= nbp) return;
  float x[kKB];
  float m = 0.0f;
#pragma unroll
  for (int i = 0; i < kKB; ++i) {
    const int g = kb * kKB + i;
    x[i] = (c < nb && g < ng) ? K[static_cast<size_t>(g) * nb + c] : 0.0f;
    m = fmaxf(m, fabsf(x[i]));
  }
  const int s = block_exp(m);
  const size_t base = (static_cast<size_t>(kb) * nbp + c) * kKB;
#pragma unroll
  for (int i = 0; i < kKB; ++i) {
    _Float16 hi, lo;
    split2(x[i], s, hi, lo);
    Bh[base + i] = hi;
    Bl[base + i] = lo;
  }
  sB[static_cast<size_t>(kb) * nbp + c] = s;
}

template <int NPROD, bool VEC>
__global__ void __launch_bounds__(kThreadsC) mfma_contract(const float* __restrict__ A, int nd, int ng,
                                                           const _Float16* __restrict__ Bh,
                                                           const _Float16* __restrict__ Bl,
                                                           const int* __restrict__ sB, int nb, int nbp, int nkb,
                                                           float* __restrict__ out) {
  __shared__ __align__(16) _Float16 Ah[kTM][kLd];
  __shared__ __align__(16) _Float16 Al[kTM][kLd];
  __shared__ __align__(16) _Float16 Bsh[kTN][kLd];
  __shared__ __align__(16) _Float16 Bsl[kTN][kLd];
  __shared__ int eA[kTM];
  __shared__ int eB[kTN];
  const int d0 = blockIdx.x * kTM;
  const int c0 = blockIdx.y * kTN;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wr = (wave / (kTN / 64)) * 64, wc = (wave % (kTN / 64)) * 64;
  float16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
    }
  }
  // f loader: four threads per row, 16 consecutive g each (one 256-byte row
  // segment per four lanes); K loader: the tile's packed hi/lo block is one
  // contiguous 32 KB range per half, read as 16-byte chunks
  const int ar = threadIdx.x >> 2, aq = threadIdx.x & 3;
  const int d = d0 + ar;
  float ra[16];
  half8 rbh[kBChunks], rbl[kBChunks];
  int reB = 0;
  auto load_block = [&](int kb) {
    const int g0 = kb * kKB + aq * 16;
    if (VEC) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (d < nd && g0 + 4 * q < ng) v = *reinterpret_cast<const float4*>(A + static_cast<size_t>(d) * ng + g0 + 4 * q);
        ra[4 * q] = v.x;
        ra[4 * q + 1] = v.y;
        ra[4 * q + 2] = v.z;
        ra[4 * q + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; ++i) ra[i] = (d < nd && g0 + i < ng) ? A[static_cast<size_t>(d) * ng + g0 + i] : 0.0f;
    }
    const size_t bbase = (static_cast<size_t>(kb) * nbp + c0) * kKB;
#pragma unroll
    for (int q = 0; q < kBChunks; ++q) {
      const int c = q * kThreadsC + threadIdx.x;  // 16-byte chunk of the tile
      rbh[q] = *reinterpret_cast<const half8*>(Bh + bbase + 8 * c);
      rbl[q] = *reinterpret_cast<const half8*>(Bl + bbase + 8 * c);
    }
    if (threadIdx.x < kTN) reB = sB[static_cast<size_t>(kb) * nbp + c0 + threadIdx.x];
  };
  load_block(0);
  for (int kb = 0; kb < nkb; ++kb) {
    __syncthreads();  // the previous block's fragments have been read
    {
      float m = 0.0f;
#pragma unroll
      for (int i = 0; i < 16; ++i) m = fmaxf(m, fabsf(ra[i]));
      m = fmaxf(m, __shfl_xor(m, 1));
      m = fmaxf(m, __shfl_xor(m, 2));
      const int s = block_exp(m);
      if (aq == 0) eA[ar] = s;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        half8 h, l;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          _Float16 hi, lo;
          split2(ra[q * 8 + i], s, hi, lo);
          h[i] = hi;
          l[i] = lo;
        }
        *reinterpret_cast<half8*>(&Ah[ar][aq * 16 + q * 8]) = h;
        *reinterpret_cast<half8*>(&Al[ar][aq * 16 + q * 8]) = l;
      }
#pragma unroll
      for (int q = 0; q < kBChunks; ++q) {
        const int c = q * kThreadsC + threadIdx.x;
        *reinterpret_cast<half8*>(&Bsh[c >> 3][(c & 7) * 8]) = rbh[q];
        *reinterpret_cast<half8*>(&Bsl[c >> 3][(c & 7) * 8]) = rbl[q];
      }
      if (threadIdx.x < kTN) eB[threadIdx.x] = reB;
    }
    if (kb + 1 < nkb) load_block(kb + 1);  // in flight during the matrix work
    __syncthreads();
    float16v tmp[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int v = 0; v < 16; ++v) tmp[i][j][v] = 0.0f;
      }
    }
#pragma unroll
    for (int ks = 0; ks < kKB / 16; ++ks) {
      const int kk = ks * 16 + (lane >> 5) * 8;
      half8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ah[i] = *reinterpret_cast<const half8*>(&Ah[wr + i * 32 + (lane & 31)][kk]);
        if (NPROD == 3) al[i] = *reinterpret_cast<const half8*>(&Al[wr + i * 32 + (lane & 31)][kk]);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        bh[j] = *reinterpret_cast<const half8*>(&Bsh[wc + j * 32 + (lane & 31)][kk]);
        if (NPROD == 3) bl[j] = *reinterpret_cast<const half8*>(&Bsl[wc + j * 32 + (lane & 31)][kk]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          tmp[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bh[j], tmp[i][j], 0, 0, 0);
          if (NPROD == 3) {
            tmp[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bl[j], tmp[i][j], 0, 0, 0);
            tmp[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[i], bh[j], tmp[i][j], 0, 0, 0);
          }
        }
      }
    }
    // acc += tmp 2^-(s_row + s_col); output element v of a lane sits at
    // row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32 of its 32x32 tile
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int sc = eB[wc + j * 32 + (lane & 31)];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = wr + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
          acc[i][j][v] += ldexpf(tmp[i][j][v], -(eA[row] + sc));
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = c0 + wc + j * 32 + (lane & 31);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int dd = d0 + wr + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (dd < nd && c < nb) out[static_cast<size_t>(dd) * nb + c] = acc[i][j][v];
      }
    }
  }
}

}  // namespace

void contraction_f16(const SyncPlan& plan, const float* d_k, int ng, const float* d_f, size_t nd, int nprod,
                     float* d_out, hipStream_t st) {
  RGN_REQUIRE(nprod == 1 || nprod == 3, RGN_ERR_INVALID, "split-FP16 products must be 1 or 3");
  RGN_REQUIRE(nd <= static_cast<size_t>(INT32_MAX) - kTM, RGN_ERR_RANGE, "too many distributions");
  const int nb = plan.nbins;
  const int nbp = static_cast<int>(ceil_div(nb, kTN)) * kTN;
  const int nkb = static_cast<int>(ceil_div(ng, kKB));
  const size_t packed = static_cast<size_t>(nkb) * nbp * kKB;
  _Float16* Bh = static_cast<_Float16*>(scratch(2 * sizeof(_Float16) * packed + sizeof(int) * nkb * nbp, 16));
  _Float16* Bl = Bh + packed;
  int* sB = reinterpret_cast<int*>(Bl + packed);
  pack_kernel_matrix<<<dim3(nbp / 256 + (nbp % 256 != 0), nkb), 256, 0, st>>>(d_k, ng, nb, nbp, Bh, Bl, sB);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "pack_kernel_matrix");
  const dim3 grid(static_cast<unsigned>(ceil_div(nd, kTM)), static_cast<unsigned>(nbp / kTN));
  // 16-byte loads of f when every row starts 16-byte aligned
  const bool vec = ng % 4 == 0 && reinterpret_cast<uintptr_t>(d_f) % 16 == 0;
  const int nd_i = static_cast<int>(nd);
#define RGN_CONTRACT(P_, V_) \
  mfma_contract<P_, V_><<<grid, kThreadsC, 0, st>>>(d_f, nd_i, ng, Bh, Bl, sB, nb, nbp, nkb, d_out)
  if (nprod == 3) {
    if (vec) RGN_CONTRACT(3, true); else RGN_CONTRACT(3, false);
  } else {
    if (vec) RGN_CONTRACT(1, true); else RGN_CONTRACT(1, false);
  }
#undef RGN_CONTRACT
  RGN_HIP(hipGetLastError());
  stage_mark(st, "mfma_contract");
}

}  // namespace rgn
