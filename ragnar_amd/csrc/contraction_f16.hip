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
// K is formed and split once per call into a packed step-major layout
// (pack_kernel_matrix); the matrix kernel (mfma_contract2) splits f as it
// goes.  An MFMA operand fragment is 8 consecutive g of one row (A) or one
// column (B): both operands use the same lane -> g assignment, so the sum over
// g is the same whatever order the hardware pairs the 16 g of one
// instruction.  A split-BF16 form (six products, no scaling) follows.
#include <hip/hip_runtime.h>

#include "emission.hpp"
#include "rgn_core.hpp"

namespace rgn {

namespace {

constexpr int kTN = 256;        // columns (photon bins) per workgroup
constexpr int kKB = 64;         // g per block (one scale exponent)
constexpr int kLd = kKB + 8;    // LDS row pitch in halves (bank spread)

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

// Packed index of (g block kb, column c, i = g - 64 kb), step-major:
// [kb][i / 16][c][16], so the 16 g of one MFMA step of one column are 32
// contiguous bytes and a wave's fragment load (32 columns, two halves of 8 g)
// reads 1 KB contiguously.
__device__ __forceinline__ size_t packed_index(int kb, int c, int i, int nbp) {
  return ((static_cast<size_t>(kb) * (kKB / 16) + i / 16) * nbp + c) * 16 + i % 16;
}

// The 64 x 16 tile of K for g block kb and columns c0 .. c0 + 15, formed in
// LDS: the block's 64 sources first (one FP64 log2 each), then the entries
// (dist_kernel_value, the same float operations as kernel_matrix, so K is
// the same bit for bit).  Entries past ng or nb are 0.
constexpr int kPackCols = 16;

__device__ __forceinline__ void form_k_tile(const PlanDev& P, const float* __restrict__ gbeta, int ng, int dist_log,
                                            int kb, int c0, float (&tile)[kKB][kPackCols + 1], Source (&src)[kKB]) {
  const int t = threadIdx.x;
  if (t < kKB) {
    const int g = kb * kKB + t;
    src[t] = g < ng ? dist_source(gbeta[g], dist_log, P) : Source{0.0f, P.v_safe, 0.0f};
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kKB * kPackCols / 256; ++r) {
    const int idx = r * 256 + t, i = idx / kPackCols, c = idx % kPackCols;
    tile[i][c] = c0 + c < P.nbins ? dist_kernel_value(src[i], c0 + c, P) : 0.0f;
  }
  __syncthreads();
}

// Split K into the packed hi/lo halves.  One block per (16 columns, g
// block): the tile is formed in LDS, each column's block maximum sets its
// exponent (16 partial maxima per column, then their maximum), and the split
// halves are written in packed order, consecutive threads to consecutive
// addresses.
__global__ void __launch_bounds__(256) pack_kernel_matrix(PlanDev P, const float* __restrict__ gbeta, int ng,
                                                          int dist_log, int nbp, _Float16* __restrict__ Bh,
                                                          _Float16* __restrict__ Bl, float* __restrict__ rB) {
  __shared__ float tile[kKB][kPackCols + 1];
  __shared__ Source src[kKB];
  __shared__ float part[16][kPackCols];
  __shared__ int sexp[kPackCols];
  const int c0 = blockIdx.x * kPackCols, kb = blockIdx.y;
  const int t = threadIdx.x;
  form_k_tile(P, gbeta, ng, dist_log, kb, c0, tile, src);
  {
    const int c = t % kPackCols, q = t / kPackCols;  // rows 4q .. 4q + 3
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) m = fmaxf(m, fabsf(tile[4 * q + i][c]));
    part[q][c] = m;
  }
  __syncthreads();
  if (t < kPackCols) {
    float m = 0.0f;
#pragma unroll
    for (int q = 0; q < 16; ++q) m = fmaxf(m, part[q][t]);
    const int sc = block_exp(m);
    sexp[t] = sc;
    rB[static_cast<size_t>(kb) * nbp + c0 + t] = ldexpf(1.0f, -sc);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kKB * kPackCols / 256; ++r) {
    const int idx = r * 256 + t;
    // packed order: (16-g step, column, g within the step)
    const int c = (idx / 16) % kPackCols;
    const int i = (idx / (16 * kPackCols)) * 16 + idx % 16;
    _Float16 hi, lo;
    split2(tile[i][c], sexp[c], hi, lo);
    const size_t at = packed_index(kb, c0 + c, i, nbp);
    Bh[at] = hi;
    Bl[at] = lo;
  }
}

// Matrix kernel: a 64 x 256 tile per 4-wave workgroup, two workgroups per CU.
// Each wave owns 64 columns (2 x 2 MFMA tiles over all 64 rows).  Only the
// f tile is shared by the waves, so only it goes through LDS, double
// buffered: while the matrix cores work on block kb, the waves split block
// kb + 1 (already in registers) into the other buffer, and one barrier per
// block orders the buffers.  K fragments are read by each wave straight from
// the packed layout (1 MB for 512 x 512, resident in L2), one 16-g step
// ahead.  The two workgroups of a CU meet at their own barriers only, so one's
// rescale overlaps the other's matrix work.  Row and column scales are floats
// (2^-s), so the rescale is one multiply and one FMA per element.  Workgroups
// are numbered so that the column tiles of a row block run on one XCD (one
// L2) at about the same time, and f is fetched from HBM once.
//
// The block loop has no branches: the split of the next block always runs
// (after the last block it fills a buffer nobody reads), loads past the last
// step or block re-read the last one, and the f load of block kb + 2 is
// issued after the block's last K load.  Loads complete in order, so a K
// fragment issued after a load of f cannot be used before that f arrives
// from HBM; issued last, the f load has the rescale, the barrier and the next
// block's first step to arrive in.  The column scales of a block are loaded
// at its start and used at its end.
constexpr int kTM2 = 64;
constexpr int kThreads2 = 256;

// Rows of f given as terms instead of values (GEN): a power law with an
// exponential cutoff over the grid x_g,
//     f[d, g] = x_g^p_d e^(-x_g / xc_d) / N_d,
// N_d its trapezoid integral over the grid.  The kernels stage L_g = log2 x_g,
// x_g and the trapezoid weights w_g = (x_g+1 - x_g-1) / 2 (halved end
// intervals) in LDS, form t = p_d L_g - c_d x_g - t_max,d in FP64 (c_d =
// log2(e) / xc_d) and the value 2^t as 2^n exp2f(t - n), n = floor(t), within
// ~2 float ulps of the same expression in FP64, and accumulate N_d 2^-t_max,d
// = sum_g f_g w_g in FP64 from the same float values while generating them;
// the row's results are divided by it when stored.  t(x) = p log2 x - c x is
// convex in x for p < 0 (largest at an end of the grid) and concave for
// p >= 0 (largest at x* = p / (c ln 2) when x* lies inside the grid, else at
// an end), so t_max, the larger of the end values and t(x*), bounds every
// exponent: no value overflows and the largest is 1.
struct GenRows {
  const float* index;   // [nd] p_d
  const float* cutoff;  // [nd] xc_d
  const float* grid;    // [ng] x_g
};
constexpr int kGenMaxNg = 2048;  // three FP64 grid terms per g in LDS

__device__ __forceinline__ float gen_value(double p, double c, double o, double L, double x) {
  const double t = fma(p, L, fma(-c, x, o));
  // no branch: elements interleave; exponents below -1000 are clamped, and
  // ldexpf flushes anything below the float range to zero
  const double n = fmax(floor(t), -1000.0);
  return ldexpf(__builtin_amdgcn_exp2f(static_cast<float>(t - n)), static_cast<int>(n));
}

// Grid terms of generated rows into LDS: L_g, x_g, w_g (three arrays of ng).
__device__ __forceinline__ void stage_grid_terms(const float* __restrict__ grid, int ng, double* gterms) {
  for (int g = threadIdx.x; g < ng; g += blockDim.x) {
    const double x = grid[g];
    gterms[g] = log2(x);
    gterms[ng + g] = x;
    gterms[2 * ng + g] = 0.5 * (static_cast<double>(grid[min(g + 1, ng - 1)]) - static_cast<double>(grid[max(g - 1, 0)]));
  }
}

// f rows as the matrix kernels read them: thread (row ar, quarter aq) of a
// 256-thread workgroup holds 16 consecutive g of row d0 + ar of one g block.
//   LOAD 0  scalar loads, each element checked against nd and ng;
//   LOAD 1  16-byte loads (rows 16-byte aligned, ng % 4 == 0), each checked;
//   LOAD 2  16-byte loads of whole blocks (and ng % 64 == 0): rows past nd
//           read row nd - 1, whose results are computed but never stored;
//   GEN     generated (GenRows; grid terms staged in LDS by the kernel),
//           rows past nd generate row nd - 1; with LOAD 2 (ng % 64 == 0) no
//           element is checked against ng.
// load(kb, ra, count): block kb into ra; with count (GEN), the block's
// values also go into the row's norm (each block is counted once, though the
// kernels' branch-free loops load the last block more than once).
template <int LOAD, bool GEN>
struct RowBlocks {
  const float* row;      // LOAD 2: start of this thread's row and quarter
  const float* A;
  int nd, ng, d, g_off;  // g_off = 16 aq
  const double* gL;      // GEN: LDS grid terms L_g, x_g, w_g
  double gp, gc, go, nsum;

  __device__ RowBlocks(const float* A_, int nd_, int ng_, int d_, int aq, const GenRows& gen, const double* gterms)
      : row(nullptr), A(A_), nd(nd_), ng(ng_), d(d_), g_off(16 * aq), gL(gterms), gp(0.0), gc(0.0), go(0.0),
        nsum(0.0) {
    const int dc = min(d, nd - 1);
    if (LOAD == 2 && !GEN) row = A + static_cast<size_t>(dc) * ng + g_off;
    if (GEN) {
      gp = gen.index[dc];
      gc = 1.4426950408889634 / static_cast<double>(gen.cutoff[dc]);  // log2(e) / xc
      const double* gx = gL + ng;
      double tmax = fmax(fma(gp, gL[0], -gc * gx[0]), fma(gp, gL[ng - 1], -gc * gx[ng - 1]));
      const double xs = gp / (gc * 0.69314718055994530942);
      if (xs > gx[0] && xs < gx[ng - 1]) tmax = fmax(tmax, fma(gp, log2(xs), -gc * xs));
      go = -tmax;
    }
  }

  __device__ __forceinline__ void load(int kb, float (&ra)[16], bool count) {
    const int g0 = kb * kKB + g_off;
    if (GEN) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const bool in = LOAD == 2 || g0 + i < ng;
        ra[i] = in ? gen_value(gp, gc, go, gL[g0 + i], gL[ng + g0 + i]) : 0.0f;
        s = fma(static_cast<double>(ra[i]), in ? gL[2 * ng + g0 + i] : 0.0, s);
      }
      if (count) nsum += s;
    } else if (LOAD == 2) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float4 v = *reinterpret_cast<const float4*>(row + kb * kKB + 4 * k);
        ra[4 * k] = v.x;
        ra[4 * k + 1] = v.y;
        ra[4 * k + 2] = v.z;
        ra[4 * k + 3] = v.w;
      }
    } else if (LOAD == 1) {
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
  }

  // GEN: 1 / N_d of this thread's row (the row's four threads each hold a
  // quarter of the sum; every thread of the workgroup calls this)
  __device__ __forceinline__ float inv_norm() const {
    double t = nsum;
    t += __shfl_xor(t, 1);
    t += __shfl_xor(t, 2);
    return static_cast<float>(1.0 / t);
  }
};

template <int NPROD, int LOAD, bool GEN = false>
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
  // generated rows: the grid terms in LDS (read by every row)
  extern __shared__ __align__(16) double gterms[];  // GEN: [3][ng]
  if (GEN) {
    stage_grid_terms(gen.grid, ng, gterms);
    __syncthreads();
  }
  RowBlocks<LOAD, GEN> src(A, nd, ng, d0 + ar, aq, gen, gterms);
  float ra[16];
  auto split_a = [&](int buf) {
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) m = fmaxf(m, fabsf(ra[i]));
    m = fmaxf(m, __shfl_xor(m, 1));
    m = fmaxf(m, __shfl_xor(m, 2));
    const int sc = block_exp(m);
    rA[buf][ar] = ldexpf(1.0f, -sc);  // the row's four threads store the same value
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
  // g = 16 s + 8 (lane / 32) .. + 8 of step s; column tile j + 1 is 512 halves on
  const int kofs = (lane >> 5) * 8;
  const size_t b0 = packed_index(0, c0 + wc + (lane & 31), kofs, nbp);
  const int nsteps = nkb * (kKB / 16);
  auto ldb = [&](const _Float16* B, int s, int j) {
    return *reinterpret_cast<const half8*>(B + b0 + static_cast<size_t>(s) * nbp * 16 + j * 512);
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
  src.load(0, ra, true);
  split_a(0);
  src.load(min(1, nkb - 1), ra, nkb > 1);
  half8 bh[2], bl[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    bh[j] = ldb(Bh, 0, j);
    if (NPROD == 3) bl[j] = ldb(Bl, 0, j);
  }
  __syncthreads();
  for (int kb = 0; kb < nkb; ++kb) {
    const int cur = kb & 1;
    float sc[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) sc[j] = rB[static_cast<size_t>(kb) * nbp + c0 + wc + j * 32 + (lane & 31)];
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
      const int sn = min(kb * (kKB / 16) + ks + 1, nsteps - 1);
      half8 nbh[2], nbl[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        nbh[j] = ldb(Bh, sn, j);
        if (NPROD == 3) nbl[j] = ldb(Bl, sn, j);
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
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        bh[j] = nbh[j];
        if (NPROD == 3) bl[j] = nbl[j];
      }
      if (ks == 1) split_a(cur ^ 1);                                  // the next block's f
      if (ks == kKB / 16 - 1) src.load(min(kb + 2, nkb - 1), ra, kb + 2 < nkb);  // the one after
    }
    // acc += tmp 2^-(s_row + s_col); output element v of a lane sits at
    // row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32 of its 32x32 tile
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float4 sr = *reinterpret_cast<const float4*>(&rA[cur][i * 32 + 8 * r + 4 * (lane >> 5)]);
          acc[i][j][4 * r + 0] = fmaf(tmp[i][j][4 * r + 0] * sr.x, sc[j], acc[i][j][4 * r + 0]);
          acc[i][j][4 * r + 1] = fmaf(tmp[i][j][4 * r + 1] * sr.y, sc[j], acc[i][j][4 * r + 1]);
          acc[i][j][4 * r + 2] = fmaf(tmp[i][j][4 * r + 2] * sr.z, sc[j], acc[i][j][4 * r + 2]);
          acc[i][j][4 * r + 3] = fmaf(tmp[i][j][4 * r + 3] * sr.w, sc[j], acc[i][j][4 * r + 3]);
        }
      }
    }
    __syncthreads();  // buffer cur is free; buffer cur ^ 1 is written
  }
  // generated rows: divide by the row norms (via LDS, like the row scales)
  __shared__ float rnorm[GEN ? kTM2 : 1];
  if (GEN) {
    const float r = src.inv_norm();
    rnorm[ar] = r;  // the row's four threads store the same value
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = c0 + wc + j * 32 + (lane & 31);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rr = i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int dd = d0 + rr;
        const float val = GEN ? acc[i][j][v] * rnorm[rr] : acc[i][j][v];
        if (dd < nd && c < nb) __builtin_nontemporal_store(val, out + static_cast<size_t>(dd) * nb + c);
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

// K -> B{h,m,l} in the step-major packed layout.  One block per (16 columns,
// g block): the tile is formed in LDS (form_k_tile), split and written in
// packed order, consecutive threads to consecutive addresses.
__global__ void __launch_bounds__(256) pack_kernel_matrix_bf16(PlanDev P, const float* __restrict__ gbeta, int ng,
                                                               int dist_log, int nbp, __bf16* __restrict__ Bh,
                                                               __bf16* __restrict__ Bm, __bf16* __restrict__ Bl) {
  __shared__ float tile[kKB][kPackCols + 1];
  __shared__ Source src[kKB];
  const int c0 = blockIdx.x * kPackCols, kb = blockIdx.y;
  form_k_tile(P, gbeta, ng, dist_log, kb, c0, tile, src);
#pragma unroll
  for (int r = 0; r < kKB * kPackCols / 256; ++r) {
    const int idx = r * 256 + threadIdx.x;
    const int c = (idx / 16) % kPackCols, i = (idx / (16 * kPackCols)) * 16 + idx % 16;
    __bf16 h, m, l;
    split3(tile[i][c], h, m, l);
    const size_t at = packed_index(kb, c0 + c, i, nbp);
    Bh[at] = h;
    Bm[at] = m;
    Bl[at] = l;
  }
}

template <int LOAD, bool GEN>
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
  // generated rows: the grid terms in LDS (read by every row)
  extern __shared__ __align__(16) double gterms[];  // GEN: [3][ng]
  if (GEN) {
    stage_grid_terms(gen.grid, ng, gterms);
    __syncthreads();
  }
  RowBlocks<LOAD, GEN> src(A, nd, ng, d0 + ar, aq, gen, gterms);
  float ra[16];
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
  const size_t b0 = packed_index(0, c0 + wc + (lane & 31), kofs, nbp);
  const int nsteps = nkb * (kKB / 16);
  float16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
    }
  }
  src.load(0, ra, true);
  split_a(0);
  src.load(min(1, nkb - 1), ra, nkb > 1);
  // K fragments two 16-g steps ahead: b? holds step s, n1? step s + 1; steps
  // past the last re-read the last
  bf16x8 bh[2], bm[2], bl[2], n1h[2], n1m[2], n1l[2];
  auto load_b = [&](int step, bf16x8* h, bf16x8* m, bf16x8* l) {
    const size_t at = b0 + static_cast<size_t>(min(step, nsteps - 1)) * nbp * 16;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      h[j] = *reinterpret_cast<const bf16x8*>(Bh + at + j * 512);
      m[j] = *reinterpret_cast<const bf16x8*>(Bm + at + j * 512);
      l[j] = *reinterpret_cast<const bf16x8*>(Bl + at + j * 512);
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
      if (ks == 1) split_a(cur ^ 1);                                  // the next block's f
      if (ks == kKB / 16 - 1) src.load(min(kb + 2, nkb - 1), ra, kb + 2 < nkb);  // the one after
    }
    __syncthreads();  // buffer cur is free; buffer cur ^ 1 is written
  }
  // generated rows: divide by the row norms (via LDS, like the row scales)
  __shared__ float rnorm[GEN ? kTM2 : 1];
  if (GEN) {
    const float r = src.inv_norm();
    rnorm[ar] = r;  // the row's four threads store the same value
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = c0 + wc + j * 32 + (lane & 31);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rr = i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int dd = d0 + rr;
        const float val = GEN ? acc[i][j][v] * rnorm[rr] : acc[i][j][v];
        if (dd < nd && c < nb) __builtin_nontemporal_store(val, out + static_cast<size_t>(dd) * nb + c);
      }
    }
  }
}

}  // namespace

void contraction_f16(const SyncPlan& plan, const float* d_gbeta, int ng, int dist_log, const float* d_f, size_t nd,
                     int nprod, float* d_out, hipStream_t st) {
  RGN_REQUIRE(nprod == 1 || nprod == 3, RGN_ERR_INVALID, "split-FP16 products must be 1 or 3");
  RGN_REQUIRE(nd <= static_cast<size_t>(INT32_MAX) - kTM2, RGN_ERR_RANGE, "too many distributions");
  if (nd == 0) return;
  const int nb = plan.nbins;
  const int nbp = static_cast<int>(ceil_div(nb, kTN)) * kTN;
  const int nkb = static_cast<int>(ceil_div(ng, kKB));
  const size_t packed = static_cast<size_t>(nkb) * nbp * kKB;
  _Float16* Bh = static_cast<_Float16*>(scratch(2 * sizeof(_Float16) * packed + sizeof(float) * nkb * nbp, 16));
  _Float16* Bl = Bh + packed;
  float* rB = reinterpret_cast<float*>(Bl + packed);
  pack_kernel_matrix<<<dim3(nbp / kPackCols, nkb), 256, 0, st>>>(device_view(plan), d_gbeta, ng, dist_log, nbp, Bh,
                                                                  Bl, rB);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "pack_kernel_matrix");
  // f read with 16-byte loads when every row starts 16-byte aligned, of whole
  // blocks when ng is a multiple of 64
  const bool vec = ng % 4 == 0 && reinterpret_cast<uintptr_t>(d_f) % 16 == 0;
  const int nd_i = static_cast<int>(nd);
  const int ntn = nbp / kTN;
  const unsigned grid = static_cast<unsigned>(ceil_div(nd, kTM2) * ntn);
  const int load = vec ? (ng % kKB == 0 ? 2 : 1) : 0;
#define RGN_CONTRACT(P_, L_) \
  mfma_contract2<P_, L_><<<grid, kThreads2, 0, st>>>(d_f, nd_i, ng, Bh, Bl, rB, nb, nbp, nkb, ntn, d_out)
#define RGN_CONTRACT_P(P_)                   \
  do {                                       \
    if (load == 2) RGN_CONTRACT(P_, 2);      \
    else if (load == 1) RGN_CONTRACT(P_, 1); \
    else RGN_CONTRACT(P_, 0);                \
  } while (0)
  if (nprod == 3) RGN_CONTRACT_P(3);
  else RGN_CONTRACT_P(1);
#undef RGN_CONTRACT_P
#undef RGN_CONTRACT
  RGN_HIP(hipGetLastError());
  stage_mark(st, "mfma_contract");
}

void contraction_f16_plaw(const SyncPlan& plan, const float* d_gbeta, int ng, int dist_log, const float* d_index,
                          const float* d_cutoff, size_t nd, int nprod, float* d_out, hipStream_t st) {
  RGN_REQUIRE(nprod == 1 || nprod == 3, RGN_ERR_INVALID, "split-FP16 products must be 1 or 3");
  RGN_REQUIRE(ng >= 2, RGN_ERR_INVALID, "a generated distribution needs at least two grid points");
  RGN_REQUIRE(ng <= kGenMaxNg, RGN_ERR_RANGE, "generated distributions take at most 2048 grid points");
  RGN_REQUIRE(nd <= static_cast<size_t>(INT32_MAX) - kTM2, RGN_ERR_RANGE, "too many distributions");
  if (nd == 0) return;
  const int nb = plan.nbins;
  const int nbp = static_cast<int>(ceil_div(nb, kTN)) * kTN;
  const int nkb = static_cast<int>(ceil_div(ng, kKB));
  const size_t packed = static_cast<size_t>(nkb) * nbp * kKB;
  _Float16* Bh = static_cast<_Float16*>(scratch(2 * sizeof(_Float16) * packed + sizeof(float) * nkb * nbp, 16));
  _Float16* Bl = Bh + packed;
  float* rB = reinterpret_cast<float*>(Bl + packed);
  pack_kernel_matrix<<<dim3(nbp / kPackCols, nkb), 256, 0, st>>>(device_view(plan), d_gbeta, ng, dist_log, nbp, Bh,
                                                                  Bl, rB);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "pack_kernel_matrix");
  const GenRows gen{d_index, d_cutoff, d_gbeta};
  const size_t gsmem = 3 * sizeof(double) * static_cast<size_t>(ng);
  const int ntn = nbp / kTN;
  const unsigned grid = static_cast<unsigned>(ceil_div(nd, kTM2) * ntn);
  const int nd_i = static_cast<int>(nd);
#define RGN_CONTRACT_GEN(P_, L_)                                                                            \
  mfma_contract2<P_, L_, true><<<grid, kThreads2, gsmem, st>>>(nullptr, nd_i, ng, Bh, Bl, rB, nb, nbp, nkb, ntn, \
                                                              d_out, gen)
  const bool full = ng % kKB == 0;
  if (nprod == 3) {
    if (full) RGN_CONTRACT_GEN(3, 2); else RGN_CONTRACT_GEN(3, 0);
  } else {
    if (full) RGN_CONTRACT_GEN(1, 2); else RGN_CONTRACT_GEN(1, 0);
  }
#undef RGN_CONTRACT_GEN
  RGN_HIP(hipGetLastError());
  stage_mark(st, "mfma_contract");
}

// Split-BF16 form.  Rows come from d_f (nd x ng, row-major) or, when d_index is
// not null, are generated from (index, cutoff) over the grid d_gbeta.
void contraction_bf16(const SyncPlan& plan, const float* d_gbeta, int ng, int dist_log, const float* d_f,
                      const float* d_index, const float* d_cutoff, size_t nd, float* d_out, hipStream_t st) {
  RGN_REQUIRE(nd <= static_cast<size_t>(INT32_MAX) - kTM2, RGN_ERR_RANGE, "too many distributions");
  const bool generated = d_index != nullptr;
  RGN_REQUIRE(!generated || ng >= 2, RGN_ERR_INVALID, "a generated distribution needs at least two grid points");
  RGN_REQUIRE(!generated || ng <= kGenMaxNg, RGN_ERR_RANGE, "generated distributions take at most 2048 grid points");
  if (nd == 0) return;
  const int nb = plan.nbins;
  const int nbp = static_cast<int>(ceil_div(nb, kTN)) * kTN;
  const int nkb = static_cast<int>(ceil_div(ng, kKB));
  const size_t packed = static_cast<size_t>(nkb) * nbp * kKB;
  __bf16* Bh = static_cast<__bf16*>(scratch(3 * sizeof(__bf16) * packed, 19));
  __bf16* Bm = Bh + packed;
  __bf16* Bl = Bm + packed;
  pack_kernel_matrix_bf16<<<dim3(nbp / kPackCols, nkb), 256, 0, st>>>(device_view(plan), d_gbeta, ng, dist_log, nbp,
                                                                       Bh, Bm, Bl);
  RGN_HIP(hipGetLastError());
  stage_mark(st, "pack_kernel_matrix_bf16");
  const GenRows gen{d_index, d_cutoff, d_gbeta};
  const int ntn = nbp / kTN;
  const unsigned grid = static_cast<unsigned>(ceil_div(nd, kTM2) * ntn);
  const int nd_i = static_cast<int>(nd);
  const bool vec = ng % 4 == 0 && reinterpret_cast<uintptr_t>(d_f) % 16 == 0;
  const size_t gsmem = 3 * sizeof(double) * static_cast<size_t>(ng);
  if (generated && ng % kKB == 0) {
    mfma_contract_bf16<2, true><<<grid, kThreads2, gsmem, st>>>(nullptr, nd_i, ng, Bh, Bm, Bl, nb, nbp, nkb, ntn,
                                                                d_out, gen);
  } else if (generated) {
    mfma_contract_bf16<0, true><<<grid, kThreads2, gsmem, st>>>(nullptr, nd_i, ng, Bh, Bm, Bl, nb, nbp, nkb, ntn,
                                                                d_out, gen);
  } else if (vec && ng % kKB == 0) {
    mfma_contract_bf16<2, false><<<grid, kThreads2, 0, st>>>(d_f, nd_i, ng, Bh, Bm, Bl, nb, nbp, nkb, ntn, d_out, gen);
  } else if (vec) {
    mfma_contract_bf16<1, false><<<grid, kThreads2, 0, st>>>(d_f, nd_i, ng, Bh, Bm, Bl, nb, nbp, nkb, ntn, d_out, gen);
  } else {
    mfma_contract_bf16<0, false><<<grid, kThreads2, 0, st>>>(d_f, nd_i, ng, Bh, Bm, Bl, nb, nbp, nkb, ntn, d_out, gen);
  }
  RGN_HIP(hipGetLastError());
  stage_mark(st, "mfma_contract_bf16");
}

}  // namespace rgn
