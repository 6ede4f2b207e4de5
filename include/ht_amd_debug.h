/*
 * ht_amd_debug.h -- component entry points of the HT engine.
 *
 * Each function runs one kernel family of the decomposition on device buffers so it can
 * be checked on its own against the CPU oracle. They mirror the reference's L1/L2
 * helpers (listed per function) but are not part of the reference's driver surface;
 * ht_amd.h holds that. All pointers are device pointers unless stated otherwise; every
 * call is synchronous.
 */
#ifndef HT_AMD_DEBUG_H
#define HT_AMD_DEBUG_H

#include "ht_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Runs each MFMA shape the engine uses once on one wavefront (lane-layout check).
 * a, b: 32x16 row-major FP32 operands (see the kernel for the per-shape use);
 * outputs: 32x32 (bf16 32x32x16, f32 32x32x2), 16x16 (f32 16x16x4, f64 16x16x4,
 * bf16 16x16x32), row-major. */
ht_status ht_debug_mfma_probe(ht_handle_t h, const float* a, const float* b, float* o_bf16_32, float* o_f32_32,
                              float* o_f32_16, double* o_f64_16, float* o_bf16_16);

/* G = X_(t) X_(t)^T (FP64, rows_t x rows_t) for the contiguous mode range t = [lo, hi)
 * with rows_t <= 1024, read straight from the column-major tensor (no permuted copy).
 * Replaces mode2/mode2h + GemmEx: fourth_order_tensor_decomposition/single_GPU/CUDA/
 * Optimized/kernel.cu:36-63, htd4_tensorcore.cu:61-103. *sumsq (host) gets ||X||^2. */
ht_status ht_mode_gram(ht_handle_t h, const float* x, int32_t order, const int64_t* dims, int32_t lo, int32_t hi,
                       int32_t precision, double* G, double* sumsq);

/* The same Gram in BF16/FP16 from a 16-bit copy of X made first (one pass that also gives
 * *sumsq = ||X||^2), the path the randomised method takes for matricisations of 257-1024
 * rows (P == 1 with rows_t % 4 == 0, or P % 64 == 0); HT_ERR_UNSUPPORTED for other views.
 * x 16-byte aligned. Replaces the same call sites as ht_mode_gram. */
ht_status ht_mode_gram16(ht_handle_t h, const float* x, int32_t order, const int64_t* dims, int32_t lo,
                         int32_t hi, int32_t precision, double* G, double* sumsq);

/* Grams of the two leading modes of an n1 x n2 x S tensor from one pass (BF16/FP16). */
ht_status ht_slab_gram(ht_handle_t h, const float* x, int32_t n1, int32_t n2, int64_t S, int32_t precision,
                       double* G1, double* G2, double* sumsq);

/* z[:, :, s] = U1^T X[:, :, s] U2 (FP32; n1, n2 multiples of 16 and <= 128; r <= 32).
 * The two-mode ttm of third_order_tensor_decompsition/CUDA/Optimized/ttm.cu:2-36. */
ht_status ht_slab_project(ht_handle_t h, const float* x, int32_t n1, int32_t n2, int64_t S, const float* U1,
                          int32_t r1, const float* U2, int32_t r2, float* z);

/* Symmetric eigen decomposition (n <= 64, FP64 Jacobi), descending, canonical signs.
 * Replaces cusolverDnSsyevd: third_order_tensor_decompsition/CUDA/Optimized/kernel.cu:682-718. */
ht_status ht_sym_eig(ht_handle_t h, const double* A, int32_t n, double* w, double* V);

/* Top-k eigenpairs of a symmetric PSD n x n FP64 matrix (subspace iteration +
 * Rayleigh-Ritz for n > 64) to Ritz-residual tolerance tol; *iters (host) = iterations. */
ht_status ht_sym_topk(ht_handle_t h, const double* G, int32_t n, int32_t k, double tol, double* U, double* w,
                      int32_t* iters);

/* TSQR: W (m x b FP32, b <= 64) = Q R, Q explicit (FP32, m x b), R FP64 (b x b).
 * W is overwritten. Replaces tsqr_svd's QR tree: third_order_tensor_decompsition/CUDA/
 * TSQR/kernel.cu:500-677. */
ht_status ht_tsqr(ht_handle_t h, float* W, int64_t m, int32_t b, float* Q, double* R);

/* Orthonormal basis Q (m x b FP32) of the columns of Y (m x b FP32, b <= 64, m >= b) by
 * shifted CholeskyQR3 with FP64 Grams -- the randomised method's orthonormalisation of
 * its sketches -- or, when a later pass meets a non-positive pivot, by TSQR
 * (*fallback, host, optional: 1 then). Y is left unchanged. Replaces the QR step of
 * tsqr_svd for tall sketches: third_order_tensor_decompsition/CUDA/TSQR/kernel.cu:500-677. */
ht_status ht_tall_orth(ht_handle_t h, float* Y, int64_t m, int32_t b, float* Q, int32_t* fallback);

/* Skinny products with A (M x K column-major): Y = A^T B (transA=1, B: M x b) or
 * Y = A B (transA=0, B: K x b), b <= 64. The sketch/co-range GEMMs of rsvd:
 * fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:434-444. */
ht_status ht_skinny_gemm(ht_handle_t h, const float* A, int64_t M, int64_t K, int32_t transA, const float* B,
                         int32_t b, int32_t precision, float* Y);

/* The same products (BF16/FP16) reading a 16-bit copy of A made first: the path the
 * randomised method's range finders take once the run holds such a copy of X. A^T B needs
 * M % 8 == 0 and 16-byte aligned A and B, A B needs M % 4 == 0; HT_ERR_UNSUPPORTED
 * otherwise. Replaces the same call sites as ht_skinny_gemm. */
ht_status ht_skinny_gemm16(ht_handle_t h, const float* A, int64_t M, int64_t K, int32_t transA, const float* B,
                           int32_t b, int32_t precision, float* Y);

/* n Philox4x32-10 standard normals of (seed, stream) on the device, and on the host
 * (out is host memory) -- the shared generator of every synthetic input. */
ht_status ht_philox_normals(ht_handle_t h, uint64_t seed, uint32_t stream, int64_t n, double* out_dev);
void ht_philox_normals_host(uint64_t seed, uint32_t stream, int64_t n, double* out);

/* dst[0..n) = src[0..n) by a plain streaming kernel (copy_f32_kernel): the known-traffic
 * reference (8 n bytes) against which rocprofv3 FETCH_SIZE / WRITE_SIZE counts of the
 * engine's kernels are scaled. No counterpart in the reference. */
ht_status ht_debug_copy(ht_handle_t h, const float* src, float* dst, int64_t n);
/* Diagnostics: process-wide count of small-solver QR calls that fell back from Cholesky
 * QR to Householder (the first round after few powers uses Householder directly). */
ht_status ht_debug_qr_fallbacks(ht_handle_t h, int64_t* count);

/* Per-kernel device timing. When enabled, every decomposition records HIP events on
 * the handle's stream around its main kernels and accumulates their durations. Kernel
 * ids: 0 slab Gram, 1 slab projection, 2 mode Gram, 3 residual, 4 skinny GEMM, 5 TSQR,
 * 6 small eigen solves, 7 small contractions, 8 generators. ht_profile_read copies the
 * accumulated milliseconds and launch counts of the first n ids (and clears them when
 * reset != 0). */
ht_status ht_profile_enable(ht_handle_t h, int32_t on);
ht_status ht_profile_read(ht_handle_t h, double* ms, int64_t* launches, int32_t n, int32_t reset);
/* Restrict the timing to the kernel ids whose bit (1 << id) is set in mask (default: all).
 * Each timed launch adds two event records to the stream; timing only the passes over X
 * keeps that cost off a throughput measurement. */
ht_status ht_profile_select(ht_handle_t h, int32_t mask);

#ifdef __cplusplus
}
#endif

#endif /* HT_AMD_DEBUG_H */
