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

// Exponent s with m 2^s in [2^14, 2^15) (0 for m == 0), at most 126 so that
// the rescale factor 2^-s is a normal float (a block whose largest value is
// below 2^-112 keeps fewer bits: such values are at float underflow anyway).
__device__ __forceinline__ int block_exp(float m) { return m > 0.0f ? min(14 - ilogbf(m), 126) : 0; }

__device__ __forceinline__ void split2(float x, int s, _Float16& hi, _Float16& lo) {
  const float xs = ldexpf(x, s);
  hi = static_cast<_Float16>(xs);
  lo = static_cast<_Float16>(xs - static_cast<float>(hi));  // the difference is exact
}

// K[g][e] (ng x nb, FP32) -> Bh/Bl[(kb * nbp + c) * 64 + i] (g = 64 kb + i), and
// the exponent of each (kb, c).  One thread per (column, g block).
// Packed index of (g block kb, column c, i = g - 64 kb).  Block-major
// (ksmajor = 0): [kb][c][64], a tile's block one contiguous range (copied
// through LDS).  Step-major (ksmajor = 1): [kb][i / 16][c][16], so the 16 g of
// one MFMA step of one column are 32 contiguous bytes and a wave's fragment
// load (32 columns, two halves of 8 g) reads 1 KB contiguously.
__device__ __forceinline__ size_t packed_index(int kb, int c, int i, int nbp, int ksmajor) {
  return ksmajor ? ((static_cast<size_t>(kb) * (kKB / 16) + i / 16) * nbp + c) * 16 + i % 16
                 : (static_cast<size_t>(kb) * nbp + c) * kKB + i;
}

__global__ void __launch_bounds__(256) pack_kernel_matrix(const float* __restrict__ K, int ng, int nb, int nbp,
                                                          _Float16* __restrict__ Bh, _Float16* __restrict__ Bl,
                                                          int* __restrict__ sB, float* __restrict__ rB,
                                                          int ksmajor) {
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
#pragma unroll
  for (int i = 0; i < kKB; ++i) {
    _Float16 hi, lo;
    split2(x[i], s, hi, lo);
    const size_t at = packed_index(kb, c, i, nbp, ksmajor);
    Bh[at] = hi;
    Bl[at] = lo;
  }
  sB[static_cast<size_t>(kb) * nbp + c] = s;
  rB[static_cast<size_t>(kb) * nbp + c] = ldexpf(1.0f, -s);
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

// Second form: a 64 x 256 tile per 4-wave workgroup, two workgroups per CU.
// Each wave owns 64 columns (2 x 2 MFMA tiles over all 64 rows).  Only the
// f tile is shared by the waves, so only it goes through LDS, double
// buffered: while the matrix cores work on block kb, the waves split block
// kb + 1 (already in registers) into the other buffer and load block kb + 2,
// and one barrier per block orders the buffers.  K fragments are read by each
// wave straight from the packed layout (1 MB for 512 x 512, resident in L2),
// one 16-g step ahead.  The two workgroups of a CU meet at their own
// barriers only, so one's rescale overlaps the other's matrix work.  Row and
// column scales are floats (2^-s), so the rescale is one multiply and one FMA
// per element.  Workgroups are numbered so that the column tiles of a row
// block run on one XCD (one L2) at about the same time, and f is fetched from
// HBM once.
constexpr int kTM2 = 64;
constexpr int kThreads2 = 256;

// Rows of f given as terms instead of values (GEN): f[d, g] = 2^t with
// t = p_d L_g - c_d x_g + o_d, L_g = log2 x_g, c_d = log2(e) / x_cut,d and
// o_d = -log2 of the row's trapezoid norm (plaw_rows); t is formed in FP64
// and 2^t as 2^n exp2f(t - n), n = floor(t), so f is within ~2 float ulps of
// the same expression evaluated in FP64 and rounded.
struct GenRows {
  const double* rows;  // [nd][3]: p_d, c_d, o_d
  const double* L;     // [ng] log2 x_g
  const double* x;     // [ng] x_g
};

__device__ __forceinline__ float gen_value(double p, double c, double o, double L, double x) {
  const double t = fma(p, L, fma(-c, x, o));
  // no branch: elements interleave; exponents below -1000 are clamped, and
  // ldexpf flushes anything below the float range to zero
  const double n = fmax(floor(t), -1000.0);
  return ldexpf(__builtin_amdgcn_exp2f(static_cast<float>(t - n)), static_cast<int>(n));
}

template <int NPROD, bool VEC, bool GEN = false>
__global__ void __launch_bounds__(kThreads2, 2) mfma_contract2(const float* __restrict__ A, int nd, int ng,
                                                               const _Float16* __restrict__ Bh,
                                                               const _Float16* __restrict__ Bl,
                                                               const float* __restrict__ rB, int nb, int nbp,
                                                               int nkb, int ntn, float* __restrict__ out,
                                                               GenRows gen = GenRows{}) {
  __shared__ __align__(16) _Float16 Ah[2][kTM2][kLd];
  __shared__ __align__(16) _Float16 Al[2][kTM2][kLd];
  __shared__ __align__(16) float rA[2][kTM2];
  // tile of this workgroup: the ntn column tiles of a row block on one XCD
  const unsigned T = gridDim.x;
  unsigned q = blockIdx.x;
  if (T % 8 == 0) q = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
  const int d0 = static_cast<int>(q / ntn) * kTM2;
  const int c0 = static_cast<int>(q % ntn) * kTN;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wc = wave * 64;
  // f producer: four threads per row, 16 consecutive g each
  const int ar = threadIdx.x >> 2, aq = threadIdx.x & 3;
  const int d = d0 + ar;
  float ra[16];
  double gp = 0.0, gc = 0.0, go = 0.0;
  // generated rows: the grid terms L_g, x_g in LDS (read by every row)
  extern __shared__ __align__(16) double gterms[];  // GEN: [2][ng]
  if (GEN) {
    for (int i = threadIdx.x; i < ng; i += blockDim.x) {
      gterms[i] = gen.L[i];
      gterms[ng + i] = gen.x[i];
    }
    __syncthreads();
    if (d < nd) {
      gp = gen.rows[3 * static_cast<size_t>(d)];
      gc = gen.rows[3 * static_cast<size_t>(d) + 1];
      go = gen.rows[3 * static_cast<size_t>(d) + 2];
    }
  }
  auto load_a = [&](int kb) {
    const int g0 = kb * kKB + aq * 16;
    if (GEN) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        ra[i] = (d < nd && g0 + i < ng) ? gen_value(gp, gc, go, gterms[g0 + i], gterms[ng + g0 + i]) : 0.0f;
      }
    } else if (VEC) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (d < nd && g0 + 4 * k < ng) v = *reinterpret_cast<const float4*>(A + static_cast<size_t>(d) * ng + g0 + 4 * k);
        ra[4 * k] = v.x;
        ra[4 * k + 1] = v.y;
        ra[4 * k + 2] = v.z;
        ra[4 * k + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; ++i) ra[i] = (d < nd && g0 + i < ng) ? A[static_cast<size_t>(d) * ng + g0 + i] : 0.0f;
    }
  };
  auto split_a = [&](int buf) {
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) m = fmaxf(m, fabsf(ra[i]));
    m = fmaxf(m, __shfl_xor(m, 1));
    m = fmaxf(m, __shfl_xor(m, 2));
    const int sc = block_exp(m);
    if (aq == 0) rA[buf][ar] = ldexpf(1.0f, -sc);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      half8 h, l;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        _Float16 hi, lo;
        split2(ra[k * 8 + i], sc, hi, lo);
        h[i] = hi;
        l[i] = lo;
      }
      *reinterpret_cast<half8*>(&Ah[buf][ar][aq * 16 + k * 8]) = h;
      if (NPROD == 3) *reinterpret_cast<half8*>(&Al[buf][ar][aq * 16 + k * 8]) = l;
    }
  };
  // K fragments of this wave (step-major layout): column wc + 32 j + lane % 32,
  // g = 16 ks + 8 (lane / 32) .. + 8
  const int kofs = (lane >> 5) * 8;
  auto b_addr = [&](int kb, int j, int ks) {
    return packed_index(kb, c0 + wc + j * 32 + (lane & 31), ks * 16 + kofs, nbp, 1);
  };
  float16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
    }
  }
  load_a(0);
  split_a(0);
  if (nkb > 1) load_a(1);
  half8 bh[2], bl[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    bh[j] = *reinterpret_cast<const half8*>(Bh + b_addr(0, j, 0));
    if (NPROD == 3) bl[j] = *reinterpret_cast<const half8*>(Bl + b_addr(0, j, 0));
  }
  __syncthreads();
  for (int kb = 0; kb < nkb; ++kb) {
    const int cur = kb & 1;
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
      // next step's K fragments (the next block's first step after the last)
      half8 nbh[2], nbl[2];
      const int nks = (ks + 1) % (kKB / 16);
      const int nkbb = ks + 1 < kKB / 16 ? kb : kb + 1;
      if (nkbb < nkb) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          nbh[j] = *reinterpret_cast<const half8*>(Bh + b_addr(nkbb, j, nks));
          if (NPROD == 3) nbl[j] = *reinterpret_cast<const half8*>(Bl + b_addr(nkbb, j, nks));
        }
      }
      const int kk = ks * 16 + kofs;
      half8 ah[2], al[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ah[i] = *reinterpret_cast<const half8*>(&Ah[cur][i * 32 + (lane & 31)][kk]);
        if (NPROD == 3) al[i] = *reinterpret_cast<const half8*>(&Al[cur][i * 32 + (lane & 31)][kk]);
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
      if (nkbb < nkb) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          bh[j] = nbh[j];
          if (NPROD == 3) bl[j] = nbl[j];
        }
      }
      // the next block's f: split into the other buffer, then fetch the one after
      if (ks == 1 && kb + 1 < nkb) {
        split_a(cur ^ 1);
        if (kb + 2 < nkb) load_a(kb + 2);
      }
    }
    // acc += tmp 2^-(s_row + s_col); output element v of a lane sits at
    // row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32 of its 32x32 tile
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const float sc = rB[static_cast<size_t>(kb) * nbp + c0 + wc + j * 32 + (lane & 31)];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float4 sr = *reinterpret_cast<const float4*>(&rA[cur][i * 32 + 8 * r + 4 * (lane >> 5)]);
          acc[i][j][4 * r + 0] = fmaf(tmp[i][j][4 * r + 0] * sr.x, sc, acc[i][j][4 * r + 0]);
          acc[i][j][4 * r + 1] = fmaf(tmp[i][j][4 * r + 1] * sr.y, sc, acc[i][j][4 * r + 1]);
          acc[i][j][4 * r + 2] = fmaf(tmp[i][j][4 * r + 2] * sr.z, sc, acc[i][j][4 * r + 2]);
          acc[i][j][4 * r + 3] = fmaf(tmp[i][j][4 * r + 3] * sr.w, sc, acc[i][j][4 * r + 3]);
        }
      }
    }
    __syncthreads();  // buffer cur is free; buffer cur ^ 1 is written
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = c0 + wc + j * 32 + (lane & 31);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int dd = d0 + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (dd < nd && c < nb) __builtin_nontemporal_store(acc[i][j][v], out + static_cast<size_t>(dd) * nb + c);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Split-BF16 form (RGN_PREC_BF16X6).  bf16 has float's exponent range, so an
// operand needs no scaling: x = hi + mid + lo exactly for a normal float x
// (hi = bf16(x); mid = bf16(x - hi) and lo = bf16(x - hi - mid), each
// difference exact), and the six products hi*hi, hi*mid, mid*hi, hi*lo,
// lo*hi, mid*mid accumulated in FP32 leave out terms of at most ~2^-24 of the
// product.  With no per-block exponents there is no partial tile and no
// rescale: every product goes straight into the accumulator, so the loop's
// vector work is the split of f (about 7 instructions per element) and the
// kernel is bound by the matrix cores (six products against three for the
// split-FP16 form, which spends more vector work than matrix work).
// Tiling as mfma_contract2: 64 x 256 per 4-wave workgroup, two per CU, f
// double buffered in LDS, K fragments from the packed layout in L2.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = static_cast<__bf16>(x);
  const float r1 = x - static_cast<float>(h);  // exact
  m = static_cast<__bf16>(r1);
  l = static_cast<__bf16>(r1 - static_cast<float>(m));  // exact difference
}

// K[g][e] -> B{h,m,l} in the step-major packed layout.
__global__ void __launch_bounds__(256) pack_kernel_matrix_bf16(const float* __restrict__ K, int ng, int nb, int nbp,
                                                               __bf16* __restrict__ Bh, __bf16* __restrict__ Bm,
                                                               __bf16* __restrict__ Bl) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int kb = blockIdx.y;
  if (c >= nbp) return;
  for (int i = 0; i < kKB; ++i) {
    const int g = kb * kKB + i;
    const float x = (c < nb && g < ng) ? K[static_cast<size_t>(g) * nb + c] : 0.0f;
    __bf16 h, m, l;
    split3(x, h, m, l);
    const size_t at = packed_index(kb, c, i, nbp, 1);
    Bh[at] = h;
    Bm[at] = m;
    Bl[at] = l;
  }
}

template <bool VEC, bool GEN>
__global__ void __launch_bounds__(kThreads2, 2) mfma_contract_bf16(const float* __restrict__ A, int nd, int ng,
                                                                   const __bf16* __restrict__ Bh,
                                                                   const __bf16* __restrict__ Bm,
                                                                   const __bf16* __restrict__ Bl, int nb, int nbp,
                                                                   int nkb, int ntn, float* __restrict__ out,
                                                                   GenRows gen) {
  __shared__ __align__(16) __bf16 As[2][3][kTM2][kLd];
  const unsigned T = gridDim.x;
  unsigned q = blockIdx.x;
  if (T % 8 == 0) q = (blockIdx.x % 8) * (T / 8) + blockIdx.x / 8;
  const int d0 = static_cast<int>(q / ntn) * kTM2;
  const int c0 = static_cast<int>(q % ntn) * kTN;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wc = wave * 64;
  const int ar = threadIdx.x >> 2, aq = threadIdx.x & 3;
  const int d = d0 + ar;
  float ra[16];
  double gp = 0.0, gc = 0.0, go = 0.0;
  // generated rows: the grid terms L_g, x_g in LDS (read by every row)
  extern __shared__ __align__(16) double gterms[];  // GEN: [2][ng]
  if (GEN) {
    for (int i = threadIdx.x; i < ng; i += blockDim.x) {
      gterms[i] = gen.L[i];
      gterms[ng + i] = gen.x[i];
    }
    __syncthreads();
    if (d < nd) {
      gp = gen.rows[3 * static_cast<size_t>(d)];
      gc = gen.rows[3 * static_cast<size_t>(d) + 1];
      go = gen.rows[3 * static_cast<size_t>(d) + 2];
    }
  }
  auto load_a = [&](int kb) {
    const int g0 = kb * kKB + aq * 16;
    if (GEN) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        ra[i] = (d < nd && g0 + i < ng) ? gen_value(gp, gc, go, gterms[g0 + i], gterms[ng + g0 + i]) : 0.0f;
      }
    } else if (VEC) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (d < nd && g0 + 4 * k < ng) v = *reinterpret_cast<const float4*>(A + static_cast<size_t>(d) * ng + g0 + 4 * k);
        ra[4 * k] = v.x;
        ra[4 * k + 1] = v.y;
        ra[4 * k + 2] = v.z;
        ra[4 * k + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; ++i) ra[i] = (d < nd && g0 + i < ng) ? A[static_cast<size_t>(d) * ng + g0 + i] : 0.0f;
    }
  };
  auto split_a = [&](int buf) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      bf16x8 h, m, l;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        __bf16 a, b, c;
        split3(ra[k * 8 + i], a, b, c);
        h[i] = a;
        m[i] = b;
        l[i] = c;
      }
      *reinterpret_cast<bf16x8*>(&As[buf][0][ar][aq * 16 + k * 8]) = h;
      *reinterpret_cast<bf16x8*>(&As[buf][1][ar][aq * 16 + k * 8]) = m;
      *reinterpret_cast<bf16x8*>(&As[buf][2][ar][aq * 16 + k * 8]) = l;
    }
  };
  const int kofs = (lane >> 5) * 8;
  auto b_addr = [&](int kb, int j, int ks) {
    return packed_index(kb, c0 + wc + j * 32 + (lane & 31), ks * 16 + kofs, nbp, 1);
  };
  float16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
    }
  }
  load_a(0);
  split_a(0);
  if (nkb > 1) load_a(1);
  // K fragments two 16-g steps ahead: b? holds step s, n1? step s + 1
  bf16x8 bh[2], bm[2], bl[2], n1h[2], n1m[2], n1l[2];
  auto load_b = [&](int step, bf16x8* h, bf16x8* m, bf16x8* l) {
    const int kb = step / (kKB / 16), ks = step % (kKB / 16);
    if (kb < nkb) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        h[j] = *reinterpret_cast<const bf16x8*>(Bh + b_addr(kb, j, ks));
        m[j] = *reinterpret_cast<const bf16x8*>(Bm + b_addr(kb, j, ks));
        l[j] = *reinterpret_cast<const bf16x8*>(Bl + b_addr(kb, j, ks));
      }
    }
  };
  load_b(0, bh, bm, bl);
  load_b(1, n1h, n1m, n1l);
  __syncthreads();
  for (int kb = 0; kb < nkb; ++kb) {
    const int cur = kb & 1;
#pragma unroll
    for (int ks = 0; ks < kKB / 16; ++ks) {
      bf16x8 nbh[2], nbm[2], nbl[2];
      load_b(kb * (kKB / 16) + ks + 2, nbh, nbm, nbl);
      const int kk = ks * 16 + kofs;
      bf16x8 ah[2], am[2], al[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ah[i] = *reinterpret_cast<const bf16x8*>(&As[cur][0][i * 32 + (lane & 31)][kk]);
        am[i] = *reinterpret_cast<const bf16x8*>(&As[cur][1][i * 32 + (lane & 31)][kk]);
        al[i] = *reinterpret_cast<const bf16x8*>(&As[cur][2][i * 32 + (lane & 31)][kk]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          // small terms first
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[i], bm[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bm[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        bh[j] = n1h[j];
        bm[j] = n1m[j];
        bl[j] = n1l[j];
        n1h[j] = nbh[j];
        n1m[j] = nbm[j];
        n1l[j] = nbl[j];
      }
      if (ks == 1 && kb + 1 < nkb) {
        split_a(cur ^ 1);
        if (kb + 2 < nkb) load_a(kb + 2);
      }
    }
    __syncthreads();  // buffer cur is free; buffer cur ^ 1 is written
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = c0 + wc + j * 32 + (lane & 31);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int dd = d0 + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (dd < nd && c < nb) __builtin_nontemporal_store(acc[i][j][v], out + static_cast<size_t>(dd) * nb + c);
      }
    }
  }
}

// L_g = log2 x_g and x_g in FP64 for the generated rows.
__global__ void __launch_bounds__(256) grid_terms(const float* __restrict__ gbeta, int ng, double* __restrict__ L,
                                                  double* __restrict__ x) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= ng) return;
  const double v = gbeta[g];
  x[g] = v;
  L[g] = log2(v);
}

// Per-row terms (p_d, c_d, o_d) of f[d, g] = x_g^p_d e^(-x_g / x_cut,d) / N_d,
// N_d = sum_g (f_g + f_g+1) (x_g+1 - x_g) / 2 (trapezoid over the grid).  One
// wave per row: each lane takes a run of intervals; the maximum exponent is
// found first so that the summed terms 2^(t - t_max) never overflow.
__global__ void __launch_bounds__(256) plaw_rows(const float* __restrict__ index, const float* __restrict__ cutoff,
                                                 size_t nd, const double* __restrict__ L,
                                                 const double* __restrict__ x, int ng, double* __restrict__ rows) {
  const size_t d = static_cast<size_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (d >= nd) return;  // uniform per wave
  const double p = index[d];
  const double c = 1.4426950408889634 / static_cast<double>(cutoff[d]);  // log2(e) / x_cut
  const int per = (ng - 1 + 63) / 64;
  const int i0 = min(ng - 1, lane * per), i1 = min(ng - 1, i0 + per);
  double tmax = -1e300;
  for (int g = i0; g <= i1 && g < ng; ++g) tmax = fmax(tmax, fma(p, L[g], -c * x[g]));
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) tmax = fmax(tmax, __shfl_xor(tmax, m));
  double sum = 0.0;
  double prev = (i0 < i1) ? gen_value(p, c, -tmax, L[i0], x[i0]) : 0.0;
  for (int g = i0; g < i1; ++g) {
    const double next = gen_value(p, c, -tmax, L[g + 1], x[g + 1]);
    sum += 0.5 * (prev + next) * (x[g + 1] - x[g]);
    prev = next;
  }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) sum += __shfl_xor(sum, m);
  if (lane == 0) {
    rows[3 * d] = p;
    rows[3 * d + 1] = c;
    rows[3 * d + 2] = -(tmax + log2(sum));
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
  _Float16* Bh = static_cast<_Float16*>(
    scratch(2 * sizeof(_Float16) * packed + (sizeof(int) + sizeof(float)) * nkb * nbp, 16));
  _Float16* Bl = Bh + packed;
  int* sB = reinterpret_cast<int*>(Bl + packed);
  float* rB = reinterpret_cast<float*>(sB + static_cast<size_t>(nkb) * nbp);
  const bool form1 = env_int("RAGNAR_CONTRACT_FORM", 2) == 1;
  pack_kernel_matrix<<<dim3(nbp / 256 + (nbp % 256 != 0), nkb), 256, 0, st>>>(d_k, ng, nb, nbp, Bh, Bl, sB, rB,
                                                                              form1 ? 0 : 1);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "pack_kernel_matrix");
  // 16-byte loads of f when every row starts 16-byte aligned
  const bool vec = ng % 4 == 0 && reinterpret_cast<uintptr_t>(d_f) % 16 == 0;
  const int nd_i = static_cast<int>(nd);
  if (form1) {
    const dim3 grid(static_cast<unsigned>(ceil_div(nd, kTM)), static_cast<unsigned>(nbp / kTN));
#define RGN_CONTRACT(P_, V_) \
  mfma_contract<P_, V_><<<grid, kThreadsC, 0, st>>>(d_f, nd_i, ng, Bh, Bl, sB, nb, nbp, nkb, d_out)
    if (nprod == 3) {
      if (vec) RGN_CONTRACT(3, true); else RGN_CONTRACT(3, false);
    } else {
      if (vec) RGN_CONTRACT(1, true); else RGN_CONTRACT(1, false);
    }
#undef RGN_CONTRACT
  } else {
    const int ntn = nbp / kTN;
    const unsigned grid = static_cast<unsigned>(ceil_div(nd, kTM2) * ntn);
#define RGN_CONTRACT(P_, V_) \
  mfma_contract2<P_, V_><<<grid, kThreads2, 0, st>>>(d_f, nd_i, ng, Bh, Bl, rB, nb, nbp, nkb, ntn, d_out)
    if (nprod == 3) {
      if (vec) RGN_CONTRACT(3, true); else RGN_CONTRACT(3, false);
    } else {
      if (vec) RGN_CONTRACT(1, true); else RGN_CONTRACT(1, false);
    }
#undef RGN_CONTRACT
  }
  RGN_HIP(hipGetLastError());
  stage_mark(st, "mfma_contract");
}

// Grid and row terms of generated rows (scratch slot 18): L_g, x_g and
// (p_d, c_d, o_d).
static GenRows prepare_gen_rows(const float* d_gbeta, int ng, const float* d_index, const float* d_cutoff,
                                size_t nd, hipStream_t st) {
  double* terms = static_cast<double*>(scratch(sizeof(double) * (2 * static_cast<size_t>(ng) + 3 * nd), 18));
  double* L = terms;
  double* x = terms + ng;
  double* rows = terms + 2 * static_cast<size_t>(ng);
  grid_terms<<<static_cast<unsigned>(ceil_div(ng, 256)), 256, 0, st>>>(d_gbeta, ng, L, x);
  RGN_HIP(hipGetLastError());
  if (nd > 0) {
    plaw_rows<<<static_cast<unsigned>(ceil_div(nd, 4)), 256, 0, st>>>(d_index, d_cutoff, nd, L, x, ng, rows);
    RGN_HIP(hipGetLastError());
    stage_mark(st, "plaw_rows");
  }
  return GenRows{rows, L, x};
}

void contraction_f16_plaw(const SyncPlan& plan, const float* d_k, int ng, const float* d_gbeta, const float* d_index,
                          const float* d_cutoff, size_t nd, int nprod, float* d_out, hipStream_t st) {
  RGN_REQUIRE(nprod == 1 || nprod == 3, RGN_ERR_INVALID, "split-FP16 products must be 1 or 3");
  RGN_REQUIRE(ng >= 2, RGN_ERR_INVALID, "a generated distribution needs at least two grid points");
  RGN_REQUIRE(nd <= static_cast<size_t>(INT32_MAX) - kTM, RGN_ERR_RANGE, "too many distributions");
  const int nb = plan.nbins;
  const int nbp = static_cast<int>(ceil_div(nb, kTN)) * kTN;
  const int nkb = static_cast<int>(ceil_div(ng, kKB));
  const size_t packed = static_cast<size_t>(nkb) * nbp * kKB;
  _Float16* Bh = static_cast<_Float16*>(
    scratch(2 * sizeof(_Float16) * packed + (sizeof(int) + sizeof(float)) * nkb * nbp, 16));
  _Float16* Bl = Bh + packed;
  int* sB = reinterpret_cast<int*>(Bl + packed);
  float* rB = reinterpret_cast<float*>(sB + static_cast<size_t>(nkb) * nbp);
  pack_kernel_matrix<<<dim3(nbp / 256 + (nbp % 256 != 0), nkb), 256, 0, st>>>(d_k, ng, nb, nbp, Bh, Bl, sB, rB, 1);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "pack_kernel_matrix");
  const GenRows gen = prepare_gen_rows(d_gbeta, ng, d_index, d_cutoff, nd, st);
  const size_t gsmem = 2 * sizeof(double) * static_cast<size_t>(ng);
  if (nd == 0) return;
  const int ntn = nbp / kTN;
  const unsigned grid = static_cast<unsigned>(ceil_div(nd, kTM2) * ntn);
  const int nd_i = static_cast<int>(nd);
  if (nprod == 3) {
    mfma_contract2<3, false, true><<<grid, kThreads2, gsmem, st>>>(nullptr, nd_i, ng, Bh, Bl, rB, nb, nbp, nkb,
                                                                   ntn, d_out, gen);
  } else {
    mfma_contract2<1, false, true><<<grid, kThreads2, gsmem, st>>>(nullptr, nd_i, ng, Bh, Bl, rB, nb, nbp, nkb,
                                                                   ntn, d_out, gen);
  }
  RGN_HIP(hipGetLastError());
  stage_mark(st, "mfma_contract");
}

// Split-BF16 form.  Rows come from d_f (nd x ng, row-major) or, when d_index is
// not null, are generated from (index, cutoff) over the grid d_gbeta.
void contraction_bf16(const SyncPlan& plan, const float* d_k, int ng, const float* d_f, const float* d_gbeta,
                      const float* d_index, const float* d_cutoff, size_t nd, float* d_out, hipStream_t st) {
  RGN_REQUIRE(nd <= static_cast<size_t>(INT32_MAX) - kTM, RGN_ERR_RANGE, "too many distributions");
  const bool generated = d_index != nullptr;
  RGN_REQUIRE(!generated || ng >= 2, RGN_ERR_INVALID, "a generated distribution needs at least two grid points");
  const int nb = plan.nbins;
  const int nbp = static_cast<int>(ceil_div(nb, kTN)) * kTN;
  const int nkb = static_cast<int>(ceil_div(ng, kKB));
  const size_t packed = static_cast<size_t>(nkb) * nbp * kKB;
  __bf16* Bh = static_cast<__bf16*>(scratch(3 * sizeof(__bf16) * packed, 19));
  __bf16* Bm = Bh + packed;
  __bf16* Bl = Bm + packed;
  pack_kernel_matrix_bf16<<<dim3(nbp / 256 + (nbp % 256 != 0), nkb), 256, 0, st>>>(d_k, ng, nb, nbp, Bh, Bm, Bl);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "pack_kernel_matrix_bf16");
  GenRows gen{};
  if (generated) gen = prepare_gen_rows(d_gbeta, ng, d_index, d_cutoff, nd, st);
  if (nd == 0) return;
  const int ntn = nbp / kTN;
  const unsigned grid = static_cast<unsigned>(ceil_div(nd, kTM2) * ntn);
  const int nd_i = static_cast<int>(nd);
  if (generated) {
    mfma_contract_bf16<false, true><<<grid, kThreads2, 2 * sizeof(double) * static_cast<size_t>(ng), st>>>(
      nullptr, nd_i, ng, Bh, Bm, Bl, nb, nbp, nkb, ntn, d_out, gen);
  } else if (ng % 4 == 0 && reinterpret_cast<uintptr_t>(d_f) % 16 == 0) {
    mfma_contract_bf16<true, false><<<grid, kThreads2, 0, st>>>(d_f, nd_i, ng, Bh, Bm, Bl, nb, nbp, nkb, ntn, d_out,
                                                                gen);
  } else {
    mfma_contract_bf16<false, false><<<grid, kThreads2, 0, st>>>(d_f, nd_i, ng, Bh, Bm, Bl, nb, nbp, nkb, ntn, d_out,
                                                                 gen);
  }
  RGN_HIP(hipGetLastError());
  stage_mark(st, "mfma_contract_bf16");
}

}  // namespace rgn
