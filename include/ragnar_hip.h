/*
 * ragnar_hip.h -- C ABI of the MI355X (gfx950) radiation hot path.
 *
 * Every entry point is plain C: pointers, sizes and scalars; no C++ types and
 * no exceptions cross this boundary.  Calls return RGN_OK (0) or an RGN_ERR_*
 * code; rgn_last_error() returns the message of the last failure on the
 * calling thread.
 *
 * Array arguments named d_* are device pointers (HBM); h_* are host pointers.
 * Particle data is structure-of-arrays: one float array per component.
 * Every compute call is enqueued on the given stream (0 = the null stream) and
 * returns without synchronising; callers that want the reference's blocking
 * semantics (every reference compute call ends in Kokkos::fence()) call
 * rgn_sync(stream).
 *
 * The library works on one device (rgn_init) and keeps one set of scratch
 * buffers on it.  Compute calls that use them are ordered by the library
 * whatever streams they are issued on: a call on another stream than the
 * previous such call first waits (hipStreamWaitEvent) for that call's
 * kernels, so calls on different streams run one after another on the
 * device, and host threads issue them one at a time.  Calls on one stream
 * are ordered by the stream itself.
 *
 * Reference interfaces replaced are cited as path:line in haykh/ragnar.
 */
#ifndef RAGNAR_HIP_H
#define RAGNAR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* rgn_stream_t; /* a hipStream_t */

enum {
  RGN_OK            = 0,
  RGN_ERR_INVALID   = 1, /* bad argument; the reference raises RuntimeError   */
  RGN_ERR_RANGE     = 2, /* bad range; the reference raises ValueError        */
  RGN_ERR_HIP       = 3, /* HIP runtime failure                               */
  RGN_ERR_NOMEM     = 4, /* device allocation failed                          */
  RGN_ERR_UNSUPPORTED = 5
};

/* F(x) table modes */
enum {
  RGN_FTABLE_REFERENCE = 0, /* the reference recipe and its quirks, synchrotron.cpp:28-65 */
  RGN_FTABLE_ACCURATE  = 1  /* F(x) = x int_x^inf K_{5/3}, FP64 integral representation   */
};

/* ---- runtime -------------------------------------------------------------
 * Replaces Kokkos::initialize / finalize behind ragnar.Initialize/Finalize,
 * pyinterface.cpp:30-49.  rgn_init selects the device and loads the kernels;
 * every other call initialises on first use if rgn_init was not called. */
int         rgn_init(int device);
int         rgn_finalize(void);
int         rgn_is_initialized(void);
const char* rgn_last_error(void);
int         rgn_sync(rgn_stream_t stream);
const char* rgn_version(void);

/* ---- F(x) = x int_x^inf K_{5/3}(s) ds table ------------------------------
 * Replaces sync::TabulateFfunc, synchrotron.cpp:48-65 (defaults npoints=200,
 * xmin=1e-6, xmax=100, synchrotron.hpp:26-29) and the TabulatedFunction<true>
 * it returns, tabulation.cpp:64-117.  The K_{5/3} evaluations (15.7k for the
 * reference recipe) run in FP64 on the GPU; the float grids and the float
 * trapezoid sum follow the reference arithmetic on the host.
 * h_x / h_y receive the node grid and the table (npoints floats each). */
typedef struct rgn_ftable rgn_ftable;
int rgn_ftable_build(int mode, int npoints, float xmin, float xmax, rgn_ftable** out);
/* A table from caller nodes (x must be the Logspace grid of n points between
 * min(x) and max(x): the kernels interpolate in log10 x with uniform node
 * spacing, as InterpolateTabulatedFunction<true> indexes, tabulation.hpp:33-35). */
int rgn_ftable_from_host(const float* h_x, const float* h_y, int n, rgn_ftable** out);
int rgn_ftable_nodes(const rgn_ftable* t, float* h_x, float* h_y);
int rgn_ftable_size(const rgn_ftable* t);
/* FP64 K_{5/3}(x_i) and dF/dx(x_i) = F/x - x K_{5/3} at the nodes (either may be NULL). */
int rgn_ftable_aux(const rgn_ftable* t, double* h_k53, double* h_dfdx);
int rgn_ftable_destroy(rgn_ftable* t);
/* sync::Ffunc_integrand(x), synchrotron.cpp:28-46 (bound as ragnar.Ffunc_integrand,
 * synchrotron.cpp:148): x * trapezoid of K_{5/3} on Logspace(x, 20, 100). */
int rgn_ffunc_integrand(float x, float* h_out);
/* FP64 K_nu(x) for n points on the GPU (used by the table builder). */
int rgn_bessel_k(double nu, const double* d_x, double* d_out, size_t n, rgn_stream_t stream);

/* ---- per-particle synchrotron emission -----------------------------------
 * A plan holds everything that depends on the photon bins and the table:
 * the per-bin interpolation coordinates and the padded device table.  It
 * replaces the setup inside SynchrotronSpectrum<D>, synchrotron.cpp:107-130,
 * and the Kernel<D> constructor, synchrotron.hpp:119-143 (the bins-unit check
 * stays with the caller).  h_bins: the photon-bin energies (nbins floats).
 * B0, g_syn, e_at: as SynchrotronSpectrum_3D(prtls, bins, B0, g_syn,
 * e_syn_at_g_syn), synchrotron.cpp:181-215. */
typedef struct rgn_sync_plan rgn_sync_plan;
int rgn_sync_plan_create(const rgn_ftable* table, const float* h_bins, int nbins, float B0,
                         float g_syn, float e_at, rgn_sync_plan** out);
int rgn_sync_plan_destroy(rgn_sync_plan* plan);

/* S[e] = e_e * sum_p w_p chi_p F(e_e / eps_p)  -- replaces the
 * parallel_for("SynchrotronSpectrum", MDRange(N, nbins), Kernel<D>) of
 * synchrotron.cpp:131-142 and Kernel<D>::operator(), synchrotron.hpp:145-232.
 * d_u/d_e/d_b: three component pointers each; any d_e[k] may be NULL (zero
 * field); d_w may be NULL (unit weights: identical to the reference).
 * d_out: nbins floats. */
int rgn_sync_spectrum(const rgn_sync_plan* plan, const float* const d_u[3],
                      const float* const d_e[3], const float* const d_b[3], const float* d_w,
                      size_t n, float* d_out, rgn_stream_t stream);

/* Adjoint of rgn_sync_spectrum: given d_gbar[e] = dL/dS[e], writes per-particle
 * dL/dw (d_gw), dL/du (d_gu[3]), dL/dB (d_gb[3]) and dL/dE (d_ge[3]); any output
 * pointer may be NULL.  gamma and the pitch-angle terms are recomputed, not
 * stored.  No reference counterpart (the reference is forward-only). */
int rgn_sync_adjoint(const rgn_sync_plan* plan, const float* const d_u[3],
                     const float* const d_e[3], const float* const d_b[3], const float* d_w,
                     size_t n, const float* d_gbar, float* d_gw, float* const d_gu[3],
                     float* const d_gb[3], float* const d_ge[3], rgn_stream_t stream);

/* Method selection for the two calls above.  RGN_METHOD_DIRECT evaluates every
 * particle x bin pair (VALU bound); RGN_METHOD_BINNED uses the exact
 * factorisation of the piecewise-linear interpolant: particles are partitioned
 * by the integer part of their table coordinate and reduced into per-sub-cell
 * moments (forward), or read one precomputed entry each (adjoint); cost is
 * ~independent of the bin count.  RGN_METHOD_AUTO picks the cheaper one from a
 * measured cost model (forward: binned from ~1.2e6 particles at 256 bins; env
 * RAGNAR_SYNC_METHOD=direct|binned overrides).  Binned needs fewer than 2^29
 * particles, at most 1024 distinct fractional bin offsets and a bin range of
 * at most 3072 table units (RGN_ERR_UNSUPPORTED otherwise; AUTO then runs
 * direct).  Same results to FP32/FP64 rounding (no reference counterpart). */
enum { RGN_METHOD_AUTO = 0, RGN_METHOD_DIRECT = 1, RGN_METHOD_BINNED = 2 };
int rgn_sync_spectrum_method(const rgn_sync_plan* plan, const float* const d_u[3],
                             const float* const d_e[3], const float* const d_b[3],
                             const float* d_w, size_t n, int method, float* d_out,
                             rgn_stream_t stream);
int rgn_sync_adjoint_method(const rgn_sync_plan* plan, const float* const d_u[3],
                            const float* const d_e[3], const float* const d_b[3],
                            const float* d_w, size_t n, const float* d_gbar, int method,
                            float* d_gw, float* const d_gu[3], float* const d_gb[3],
                            float* const d_ge[3], rgn_stream_t stream);

/* Spatially resolved spectra: cell c owns particles [d_offsets[c], d_offsets[c+1]);
 * d_out[c * nbins + e].  One persistent launch plus a fixed-order combine.
 * h_offsets: host copy of the ncells+1 offsets (used to balance the launch).
 * No reference counterpart: equals rgn_sync_spectrum applied per cell. */
int rgn_sync_cells(const rgn_sync_plan* plan, const float* const d_u[3],
                   const float* const d_e[3], const float* const d_b[3], const float* d_w,
                   const int64_t* h_offsets, int ncells, float* d_out, rgn_stream_t stream);

/* Diagnostics (no reference counterpart): per-particle (e_peak, chiR, W,
 * V_hi, V_lo) as the kernels see them (d_out: 5 * n floats), and the plan's
 * integer (Z, TP, NB, M, NK, binned method available) and double (K, Amin,
 * Amax) scalars. */
int rgn_sync_particle_terms(const rgn_sync_plan* plan, const float* const d_u[3],
                            const float* const d_e[3], const float* const d_b[3],
                            const float* d_w, size_t n, float* d_out, rgn_stream_t stream);
int rgn_sync_plan_info(const rgn_sync_plan* plan, int* ints6, double* dbl3);

/* Diagnostics (no reference counterpart): rgn_sync_spectrum_method /
 * rgn_sync_adjoint_method with a HIP event recorded on `stream` before the
 * first and after every kernel launch of the call.  On return (after waiting
 * for the last event) stage_ms[i] holds the elapsed milliseconds of launch i,
 * names[i] (if names is not NULL) a static string naming its kernel, and
 * *nstages the number of launches (at most 16; max_stages entries are
 * written).  The outputs are those of the untimed call. */
int rgn_sync_spectrum_stages(const rgn_sync_plan* plan, const float* const d_u[3],
                             const float* const d_e[3], const float* const d_b[3],
                             const float* d_w, size_t n, int method, float* d_out,
                             rgn_stream_t stream, float* stage_ms, const char** names,
                             int max_stages, int* nstages);
int rgn_sync_adjoint_stages(const rgn_sync_plan* plan, const float* const d_u[3],
                            const float* const d_e[3], const float* const d_b[3],
                            const float* d_w, size_t n, const float* d_gbar, int method,
                            float* d_gw, float* const d_gu[3], float* const d_gb[3],
                            float* const d_ge[3], rgn_stream_t stream, float* stage_ms,
                            const char** names, int max_stages, int* nstages);

/* ---- spectrum from a tabulated distribution ------------------------------
 * Replaces SynchrotronSpectrumFromDist, synchrotron.cpp:69-105, with
 * KernelFromDist::operator(), synchrotron.hpp:72-95.  d_gbeta/d_f: the
 * distribution's bins and values (ng floats); dist_log: the distribution's
 * bins are log spaced (multiplies by gbeta, synchrotron.hpp:89-93).
 * The plan's B0 is unused on this path. */
int rgn_sync_fromdist(const rgn_sync_plan* plan, const float* d_gbeta, const float* d_f,
                      int ng, int dist_log, float* d_out, rgn_stream_t stream);

/* Batched contraction S[d, e] = sum_g f[d, g] K[g, e] for nd distributions
 * sharing one gbeta grid (row-major d_f[nd * ng], d_out[nd * nbins]; both
 * may be NULL when nd == 0).  K[g, e] is KernelFromDist's term
 * (synchrotron.hpp:78-93) formed on the GPU.
 * precision: RGN_PREC_FP32 (the library GEMM: rocBLAS SGEMM, FP32
 * throughout), RGN_PREC_F16X3 (the hand-written split-FP16 MFMA kernel: each
 * operand split into two fp16 parts with power-of-two scales per block of 64
 * g, three products, FP32 accumulation; every element within 2^18 of its
 * block's largest keeps 22 bits, so the result holds ~1e-6 when each block of
 * 64 grid points spans about a decade of gbeta or less (512 log points over 6
 * decades: 4.5e-7 from FP64, better than the FP32 GEMM; 64 points over 4
 * decades: 0.19 in the highest bin), RGN_PREC_BF16X6 below has no such
 * condition),
 * RGN_PREC_F16X1 (one FP16 product; reported, not parity), RGN_PREC_BF16X6
 * (the hand-written split-BF16 MFMA kernel: each operand split exactly into
 * three bf16 parts, which need no scaling, six products, FP32 accumulation;
 * the dropped products are below ~2^-24 relative).  No reference
 * counterpart: equals rgn_sync_fromdist per row. */
enum { RGN_PREC_FP32 = 0, RGN_PREC_F16X3 = 1, RGN_PREC_F16X1 = 2, RGN_PREC_BF16X6 = 3 };
int rgn_sync_fromdist_batched(const rgn_sync_plan* plan, const float* d_gbeta, int ng,
                              int dist_log, const float* d_f, size_t nd, int precision,
                              float* d_out, rgn_stream_t stream);

/* rgn_sync_fromdist_batched with the rows given by two numbers each instead
 * of ng values: f[d, g] = x_g^p_d exp(-x_g / xc_d) / N_d, x_g = d_gbeta[g],
 * p_d = d_index[d], xc_d = d_cutoff[d], N_d the trapezoid integral of the
 * row over the grid (sum_g (f_g + f_g+1) (x_g+1 - x_g) / 2).  The rows and
 * their norms are generated inside the matrix-core kernels (FP64 exponent,
 * float values, FP64 norm), so the call reads 8 bytes per distribution
 * instead of 4 ng (d_index, d_cutoff and d_out may be NULL when nd == 0).  precision: RGN_PREC_F16X3, RGN_PREC_F16X1 or
 * RGN_PREC_BF16X6 (RGN_ERR_UNSUPPORTED otherwise); 2 <= ng <= 2048
 * (RGN_ERR_INVALID below, RGN_ERR_RANGE above: the grid's FP64 terms are
 * staged in LDS).
 * No reference counterpart (the reference evaluates a generator on the host
 * per distribution, distributions.cpp:37-66, and integrates each with
 * SynchrotronSpectrumFromDist, synchrotron.cpp:69-105). */
int rgn_sync_fromdist_batched_plaw(const rgn_sync_plan* plan, const float* d_gbeta, int ng,
                                   int dist_log, const float* d_index, const float* d_cutoff,
                                   size_t nd, int precision, float* d_out, rgn_stream_t stream);

/* ---- energy histogram -----------------------------------------------------
 * Replaces Particles<D>::energyDistribution, particles.cpp:189-260: energy =
 * |u| (fourvel) or sqrt(1+u^2) in double; log index clamped to the first /
 * last bin; log_bins adds (float)(1/E) per particle, otherwise 1.
 * A plan tabulates the index function at its discontinuities on the host
 * (bit-exact with the reference's double log10 indexing).
 * d_hist: nb floats; d_counts: nb uint64 particle counts (may be NULL). */
typedef struct rgn_hist_plan rgn_hist_plan;
int rgn_hist_plan_create(const float* h_bins, int nb, int log_bins, rgn_hist_plan** out);
int rgn_hist_plan_destroy(rgn_hist_plan* plan);
int rgn_energy_hist(const rgn_hist_plan* plan, const float* const d_u[3], size_t n,
                    int fourvel, float* d_hist, uint64_t* d_counts, rgn_stream_t stream);

/* rgn_energy_hist without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-block partials live in a fixed 8 MiB device array of the library (each
 * device holds its own copy until the process ends) instead of the library's
 * scratch; the call makes no hipMalloc/hipFree and no allocating library
 * call.  The grid is cut so that bins x blocks x 12 bytes fits the array
 * (the full grid of 2048 blocks up to 341 bins; fewer blocks beyond), so any
 * plan and particle count are accepted; counts equal rgn_energy_hist's bit
 * for bit and sums agree to FP64 rounding.  Every call on a device shares its
 * array, which keeps the last call's contents (each call writes every entry
 * it reads, so nothing needs clearing): calls that overlap on one device, from
 * different streams or threads, race, and the caller orders them.  The plan
 * is created beforehand (rgn_hist_plan_create allocates). */
int rgn_energy_hist_noalloc(const rgn_hist_plan* plan, const float* const d_u[3], size_t n,
                            int fourvel, float* d_hist, uint64_t* d_counts, rgn_stream_t stream);

/* ---- inverse-Compton spectrum ------------------------------------------------ */
/* ICSpectrum (src/physics/ic.cpp:15-46, per-triple kernel src/physics/ic.hpp:28-86):
 * E_ic^2 f(E_ic) at the neic energies d_eic (mc^2 units) of particles with
 * gamma bins d_g and values d_fg (ng; g_log: the bins are log spaced, which
 * divides by g instead of g^2) scattering soft photons of energies d_es (mc^2)
 * and values d_fes (nes).  Terms in the reference's FP32 arithmetic, summed per
 * output bin in FP64 in a fixed order.  The unit checks of the reference's
 * kernel constructor (ic.hpp:52-59) belong to the caller (the pybind layer
 * raises them as the reference does). */
int rgn_ic_spectrum(const float* d_g, const float* d_fg, int ng, int g_log, const float* d_es,
                    const float* d_fes, int nes, const float* d_eic, int neic, float* d_out,
                    rgn_stream_t stream);

/* ---- device buffers (used by the Python layer) ---------------------------- */
int rgn_malloc(void** d_ptr, size_t bytes);
int rgn_free(void* d_ptr);
int rgn_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes);
int rgn_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes);
int rgn_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes);
int rgn_memset(void* d_dst, int value, size_t bytes);
/* Pinned host staging and asynchronous copies (used by the HDF5 readers). */
int rgn_host_malloc(void** h_ptr, size_t bytes);
int rgn_host_free(void* h_ptr);
int rgn_memcpy_h2d_async(void* d_dst, const void* h_src, size_t bytes, rgn_stream_t stream);
int rgn_stream_create(rgn_stream_t* stream);
int rgn_stream_destroy(rgn_stream_t stream);
/* Events for overlapping host work with queued copies. */
int rgn_event_create(void** event);
int rgn_event_destroy(void* event);
int rgn_event_record(void* event, rgn_stream_t stream);
int rgn_event_sync(void* event);

#ifdef __cplusplus
}
#endif

#endif /* RAGNAR_HIP_H */
