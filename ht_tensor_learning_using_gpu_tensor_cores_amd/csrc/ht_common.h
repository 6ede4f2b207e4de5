// Shared device helpers and kernel launcher declarations of the HT engine (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HT_HIP_CHECK(expr)                                  \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return ht::hip_fail(_e, #expr);   \
  } while (0)

namespace ht {

int hip_fail(hipError_t e, const char* what);

// ---------------------------------------------------------------- MFMA operand types
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

enum Prec { PREC_F32 = 0, PREC_BF16 = 1, PREC_F16 = 2, PREC_FP8 = 3 };

// FP8 (OCP E4M3, gfx950's v_cvt_pk_fp8_f32: round to nearest even, largest finite 448,
// NaN past it) of four FP32 values, packed low byte first.
__device__ __forceinline__ uint32_t fp8x4(float a, float b, float c, float d) {
  int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
  return (uint32_t)w;
}
__device__ __forceinline__ uint8_t fp8_bits(float a) {
  return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(a, 0.f, 0, false) & 0xff);
}
// Power-of-two scale that brings a tile whose largest magnitude is m to at most 224 (half
// the FP8 range, so rounding never reaches NaN); 1 for an all-zero tile.
__device__ __forceinline__ float fp8_scale(float m) {
  if (!(m > 0.f) || !isfinite(m)) return 1.f;
  const float e = fminf(fmaxf(floorf(log2f(224.f / m)), -120.f), 120.f);
  return ldexpf(1.f, (int)e);
}

// FP32 -> BF16 bits, round to nearest even (inputs are finite: the engine checks them).
__device__ __forceinline__ uint16_t f2bf_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ uint16_t f2h_bits(float f) {
  _Float16 h = (_Float16)f;
  return __builtin_bit_cast(uint16_t, h);
}
template <int PREC>
__device__ __forceinline__ uint16_t to16(float f) {
  if constexpr (PREC == PREC_F16) return f2h_bits(f);
  else return f2bf_bits(f);
}

// Two FP32 values to one packed pair of 16-bit values (low half = a), round to nearest
// even: one v_cvt_pk_bf16_f32 on gfx950 for BF16.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <int PREC>
__device__ __forceinline__ uint32_t cvt2(float a, float b) {
  const f32x2 v = {a, b};
  if constexpr (PREC == PREC_F16) return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, f16x2));
  else return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// Streaming 16-byte load that the caches need not keep (X is read once per pass).
__device__ __forceinline__ float4 load_stream(const float* p) {
  const f32x4 t = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p));
  return make_float4(t[0], t[1], t[2], t[3]);
}

// 32x32x16 MFMA on 16-bit operands held as raw bits.
template <int PREC>
__device__ __forceinline__ f32x16 mfma32x32x16(u16x8 a, u16x8 b, f32x16 c) {
  if constexpr (PREC == PREC_F16)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Output position of element e (0..15) of a 32x32 MFMA accumulator in lane l:
//   row = 8*(e/4) + 4*(l/32) + (e%4), col = l%32.
__device__ __forceinline__ int acc32_row(int e, int lane) { return 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3); }
// 16x16 f32 accumulators (f32x4): row = 4*(l/16) + e, col = l%16.
__device__ __forceinline__ int acc16_row(int e, int lane) { return 4 * (lane >> 4) + e; }
// 16x16x4 f64 accumulators (f64x4) interleave the other way: row = 4*e + l/16, col = l%16.
__device__ __forceinline__ int acc16_row_f64(int e, int lane) { return 4 * e + (lane >> 4); }

__device__ __forceinline__ float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Philox4x32-10 (Salmon et al., SC'11). Key = (seed, 0x5EED0000 ^ stream),
// counter = (block lo, block hi, 0, 0); a block yields 4 normals by Box-Muller.
struct Philox4 {
  uint32_t v[4];
};
__host__ __device__ inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                 uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  Philox4 o;
  o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
  return o;
}
// The i-th standard normal of (seed, stream), in double precision.
__host__ __device__ inline double philox_normal(uint64_t seed, uint32_t stream, uint64_t i) {
  uint64_t blk = i >> 2;
  Philox4 w = philox4x32_10((uint32_t)blk, (uint32_t)(blk >> 32), 0u, 0u, (uint32_t)seed,
                            0x5EED0000u ^ stream);
  int j = (int)(i & 3);
  double ua = ((double)w.v[(j >> 1) * 2] + 0.5) * 2.3283064365386963e-10;
  double ub = ((double)w.v[(j >> 1) * 2 + 1] + 0.5) * 2.3283064365386963e-10;
  double r = sqrt(-2.0 * log(ua));
  double t = 6.283185307179586 * ub;
  return (j & 1) ? r * sin(t) : r * cos(t);
}
__host__ __device__ inline double philox_uniform(uint64_t seed, uint32_t stream, uint64_t i) {
  uint64_t blk = i >> 2;
  Philox4 w = philox4x32_10((uint32_t)blk, (uint32_t)(blk >> 32), 0u, 0u, (uint32_t)seed,
                            0x5EED0000u ^ stream);
  return ((double)w.v[i & 3] + 0.5) * 2.3283064365386963e-10;
}

// ------------------------------------------------------------------ tensor views
// (P, M, S) view of a column-major tensor: element (m, c) of the matricisation with
// c = p + P*s lives at x[p + P*m + P*M*s].
struct PMS {
  long P, M, S;
};

// ------------------------------------------------------------------ launchers
// All launchers enqueue on `stream` and return 0 or an ht_status code.

// Leaf Gram partials of an M-row matricisation (M <= 1024) in tile x tile pieces,
// tile = gram_pms_tile(x, v, prec) (128, or 256 for 16-bit products of more than 256
// rows in contiguous layouts). part: [gram_pms_pairs][n_split][tile*tile] FP32; norms:
// [pairs*n_split] FP64 sums of squares whose total is ||X||^2.
int gram_pms_launch(hipStream_t st, const float* x, PMS v, int prec, int n_split, float* part,
                    double* norms);
int gram_pms_tile(const float* x, PMS v, int prec);
int gram_pms_pairs(const float* x, PMS v, int prec);
int gram_pms_nsplit(const float* x, PMS v, int prec);
// Sum partials into a full symmetric FP64 Gram (ld = M).
int gram_reduce_launch(hipStream_t st, const float* part, int M, int n_split, double* G, int tile = 128);
// Two Gram reductions (n1, n2 partials each; and the sum of the per-block norms when
// norms != nullptr) in one launch.
int gram_reduce2_launch(hipStream_t st, const float* p1, int M1, int n1, double* G1, const float* p2, int M2, int n2,
                        double* G2, const double* norms, int n_norms, double* norm_out, int tile1 = 128,
                        int tile2 = 128);
int gram_reduce_norms_launch(hipStream_t st, const double* norms, int n, double* out);
// 16-bit copy y of the n floats of x (16-byte aligned; prec BF16 or FP16, round to nearest
// even) and *sumsq = ||x||^2 of the FP32 values; part holds cast16_part_size() doubles.
int cast16_part_size();
int cast16_sumsq_launch(hipStream_t st, const float* x, long n, int prec, uint16_t* y, double* part,
                        double* sumsq);
// Gram partials from such a copy (16-byte aligned), in the layout of gram_pms_launch's
// 256-row tiles (part: [gram16_pairs][n_split][256*256] FP32), for views where gram16_ok
// holds (257 to 1024 rows; P == 1 with M % 4 == 0, or P % 64 == 0).
bool gram16_ok(PMS v);
int gram16_pairs(PMS v);
int gram16_nsplit(PMS v);
int gram16_launch(hipStream_t st, const uint16_t* x, PMS v, int prec, int n_split, float* part);

// Fused slab kernels on the two leading modes (x as n1 x n2 x S, n1,n2 <= 128).
// GRAM: part1/part2 [n_blocks][128*128] partials of X_(1)X_(1)^T and X_(2)X_(2)^T.
int slab_gram_launch(hipStream_t st, const float* x, int n1, int n2, long S, int prec, int n_blocks,
                     float* part1, float* part2, double* norms);
// PROJ: z[a + r1*b + r1*r2*s] = sum_ij U1[i,a] X[i,j,s] U2[j,b]  (r1, r2 <= 64, FP32).
// max_blocks: the grid's cap (3 blocks per CU fit; 768 fills the chip)
int slab_proj_launch(hipStream_t st, const float* x, int n1, int n2, long S, const float* U1, int r1,
                     const float* U2, int r2, float* z, int max_blocks = 768);

// Generic mode projection y = x x_t U^T on a (P,M,S) view: y[p + P*j + P*r*s] =
// sum_m U[m + M*j] x[p + P*m + P*M*s]. T = float or double.
int proj_pms_f32_launch(hipStream_t st, const float* x, PMS v, const float* U, int r, float* y);
int proj_pms_f64_launch(hipStream_t st, const double* x, PMS v, const double* U, int r, double* y);
// The same with U given transposed: Ut is r x M column-major (U[m, j] = Ut[j + r m]).
int proj_pms_f64t_launch(hipStream_t st, const double* x, PMS v, const double* Ut, int r, double* y);
// The same with an FP32 result (FP64 products, rounded once at the store).
int proj_pms_f64t_f32out_launch(hipStream_t st, const double* x, PMS v, const double* Ut, int r, float* y);
// Two independent proj_pms_f64t products in one launch.
int proj_pms_f64t_pair_launch(hipStream_t st, const double* x0, PMS v0, const double* Ut0, int r0, double* y0,
                              const double* x1, PMS v1, const double* Ut1, int r1, double* y1);
int proj_pms_f32in_f64_launch(hipStream_t st, const float* x, PMS v, const double* U, int r, double* y);

// Skinny FP32 GEMMs over a big column-major A (M x K, lda = M), b <= 64:
//   Y(K x b) = A^T B(M x b)       (gemm_atb)
//   Y(M x b) = A B(K x b)         (gemm_ab)
// With prec = BF16/F16 the A and B tiles are rounded to 16 bits (FP32 accumulation).
int gemm_atb_launch(hipStream_t st, const float* A, long M, long K, const float* B, int b, int prec,
                    float* Y, float* ws, size_t ws_floats);
int gemm_ab_launch(hipStream_t st, const float* A, long M, long K, const float* B, int b, int prec,
                   float* Y, float* ws, size_t ws_floats);
size_t gemm_ws_floats(long M, long K, int b);
// The same products with A a 16-bit copy in the precision's format (prec BF16/F16; the
// values gemm_*_launch would round A to). gemm16_ok: gemm_atb16 (op_atb = 1) takes
// M % 8 == 0 with A and B 16-byte aligned, gemm_ab16 M % 4 == 0 with A 8-byte aligned.
bool gemm16_ok(const uint16_t* A, long M, const float* B, int op_atb);
int gemm_atb16_launch(hipStream_t st, const uint16_t* A, long M, long K, const float* B, int b, int prec,
                      float* Y, float* ws, size_t ws_floats);
int gemm_ab16_launch(hipStream_t st, const uint16_t* A, long M, long K, const float* B, int b, int prec,
                     float* Y, float* ws, size_t ws_floats);

// Low-rank tiles C = L R^T (L: M x k, R: N x k, both column-major, k <= 256).
//   mode 0: store C into out (ld M);  mode 1: accumulate sum (X - C)^2 and sum X^2 (FP64)
//   into acc[2*block] per block; the host sums the blocks.
int lowrank_launch(hipStream_t st, int mode, const float* L, const float* R, long M, long N, int k,
                   const float* X, float* out, double* acc, int* n_blocks_out);

// Small dense FP64 kernels (single work-group unless noted).
int jacobi_eig_launch(hipStream_t st, const double* A, int n, int lda, double* w, double* V, int ldv,
                      int* info, int max_sweeps, double tol);
int jacobi_eig_batched_launch(hipStream_t st, const double* A, int n, long strideA, double* w, long stridew,
                              double* V, long strideV, int batch, int* info, int max_sweeps, double tol);
// C(m x n) = op(A) op(B) with op = transpose flags; multi-block; all FP64, column-major.
int dgemm_small_launch(hipStream_t st, int ta, int tb, int m, int n, int k, double alpha, const double* A,
                       int lda, const double* B, int ldb, double beta, double* C, int ldc);
// The same product (no transposes) with A in FP32, widened to FP64 as it is loaded.
int dgemm_small_f32a_launch(hipStream_t st, int m, int n, int k, double alpha, const float* A, int lda,
                            const double* B, int ldb, double beta, double* C, int ldc);
// The same with the k range split over work-groups (small C, long k): ws holds
// dgemm_splitk_slices(m, n, k) * m * n doubles; the slices are summed in a fixed order.
int dgemm_splitk_slices(int m, int n, int k);
int dgemm_splitk_launch(hipStream_t st, int ta, int tb, int m, int n, int k, double alpha, const double* A, int lda,
                        const double* B, int ldb, double beta, double* C, int ldc, double* ws);
// Cholesky QR with column replacement (Z: m x b FP64, in place); R out (b x b).
int cholqr_launch(hipStream_t st, double* Z, int m, int b, double* W, double* R, uint64_t seed,
                  uint32_t stream, int* info);
// One pass of shifted CholeskyQR from W = Z^T Z (first pass: column equilibration and
// shift): Rinv (b x b) with Z Rinv orthonormal; Z is m x b.
int chol_rinv_launch(hipStream_t st, const double* W, int b, long m, int first, double* Rinv, int* info);
int refresh_zero_cols_launch(hipStream_t st, double* Z, long m, int b, double* W, uint64_t seed, uint32_t stream);
int ritz_resid_launch(hipStream_t st, const double* GU, const double* U, const double* w, int n, int k,
                      double* out);
// G32 (optional): an FP32 copy of G written alongside.
int gram_pms_f64_launch(hipStream_t st, const double* x, PMS v, double* G, float* G32 = nullptr);
int sum_strided_launch(hipStream_t st, const double* acc, int n, double* out);
int sqrt_clamp_launch(hipStream_t st, const double* w, int k, double* s);
// One FP64 -> FP32 copy of cast_multi_launch (sqrt_clamp: sqrt(max(x, 0)) before rounding).
struct CastItem {
  const double* src;
  float* dst;
  long n;
  int sqrt_clamp;
};
constexpr int CAST_MAX = 48;
int cast_multi_launch(hipStream_t st, const CastItem* items, int n);
// U[:, j] = A[:, j] / sqrt(max(w_j, 0)) with canonical signs, A[r, j] at in[r rs + j cs]
// (in may alias U), m <= 4096; ok (optional) receives the rank test w[0] > 0 and
// w[k-1] > 1e-24 w[0].
int cols_unit_sign_launch(hipStream_t st, const double* in, long rs, long cs, double* U, int m, int k, int ldu,
                          const double* w, int* ok = nullptr);
int scale_cols_inv_launch(hipStream_t st, double* A, int m, int n, const double* s);
// Batched small symmetric eigen solvers (k_small.hip).
int jacobi_full_launch(hipStream_t st, const double* A, long sA, int n, double* w, long sw, double* V, long sV,
                       int batch, int* info, int max_sweeps, double tol);
// W == nullptr: the start block is the Philox normals (gseed, gstream) of element
// i = problem * sWY + row * b + column, the values fill_normals_f64 writes.
int power_launch(hipStream_t st, const double* G, long sG, const double* W, double* Y, long sWY, int n, int b, int q,
                 int mode, const double* cut, double* T, int batch, const float* G32 = nullptr, uint64_t gseed = 0,
                 uint32_t gstream = 0);
int qr_launch(hipStream_t st, const double* Y, double* Q, long sYQ, int n, int b, int try_chol, int batch);
// Running count of qr_kernel calls that took the Householder fallback (process-wide).
int qr_fallback_count(unsigned long long* out);
int rr_launch(hipStream_t st, const double* Q, const double* Z, long sQ, const double* T, int n, int b, int k,
              double* Uf, double* U, long sU, double* w, long sw, double* wall, double* cut, double tol,
              double tol_rel, double floor, int* converged, double* rres, int batch);
// Casts and sign canonicalisation.
int cast_f64_f32_launch(hipStream_t st, const double* a, float* b, long n);
int cast_f32_f64_launch(hipStream_t st, const float* a, double* b, long n);
// ws: canon_signs_ws_doubles(m, k) doubles (none below the split size).
size_t canon_signs_ws_doubles(int m, int k);
int canon_signs_launch(hipStream_t st, double* U, int m, int k, int ldu, double* ws);
int fill_normals_f64_launch(hipStream_t st, double* a, long n, uint64_t seed, uint32_t stream);
int fill_normals_f32_launch(hipStream_t st, float* a, long n, uint64_t seed, uint32_t stream);
int copy_cols_f64_launch(hipStream_t st, const double* a, int lda, double* b, int ldb, int m, int n);
int sumsq_f64_launch(hipStream_t st, const double* a, long n, double* out);
int sumsq_f32_launch(hipStream_t st, const float* a, long n, double* out);

// TSQR of a tall-skinny FP32 W (m x b, ld m, b <= 64): explicit Q (m x b, FP32) and R
// (b x b, FP64). ws must hold tsqr_ws_bytes(m, b) bytes.
size_t tsqr_ws_bytes(long m, int b);
int tsqr_launch(hipStream_t st, float* W, long m, int b, float* Q, double* R, void* ws);

// Generators.
int gen_noise_sumsq_launch(hipStream_t st, long n, uint64_t seed, uint32_t stream, double* acc, int* nb);
int gen_add_noise_launch(hipStream_t st, float* x, long n, uint64_t seed, uint32_t stream, double scale);
int copy_f32_launch(hipStream_t st, const float* src, float* dst, long n);
int touch_stream(hipStream_t st);
// Batched HT (k_batch.hip): one descriptor per tensor, prepared on the host.
struct BatchItem {
  long x_off;     // element offset of the tensor in x
  long out_off;   // float offset of its packed output block
  int n[3];       // mode sizes
  int k[5];       // heap ranks (k[0] = 1)
  int u_off[5];   // offsets within the block (floats) of U_t (leaves 1, 3, 4), else -1
  int b_off[5];   // offsets of B_t (interior 0, 2), else -1
  int s_off[5];   // offsets of the singular values of nodes 1..4
};
size_t batch_scratch_bytes(int blocks);
int batch_ht_launch(hipStream_t st, const float* x, const void* items, const int* order, int count, float* out,
                    double* rel_err, double* scratch, int blocks, int kc);
// Bump tensor from its factors (n_bumps x order x n) and amplitudes. With sumsq != null
// also sumsq[0] = ||x||^2 of the FP32 values, part holding gen_bumps_planes(order, n)
// doubles (order >= 2, n <= 1024, n_bumps <= 8, x 16-byte aligned; else an error).
long gen_bumps_planes(int order, long n);
int gen_bumps_launch(hipStream_t st, float* x, int order, long n, const float* factors, const float* amp,
                     int n_bumps, double* sumsq, double* part);

// Learned per-mode sketch (k_sketch.hip). S_m is w x n_m row-major at off[m] of the
// flat sketch array; a node's test matrix is the Khatri-Rao product over its column
// modes m0..m1-1 (cols = prod n, column-major over those modes).
#define HT_KR_MAX 7
int kr_omega_launch(hipStream_t st, const float* S, const int64_t* n, const int64_t* off, int m0, int m1, int w,
                    float* om);
int kr_grad_launch(hipStream_t st, const float* S, const float* dom, const int64_t* n, const int64_t* off, int m0,
                   int m1, int w, float scale, float* grad);
// out = A^T B (wa x wb, FP64) of two tall FP32 matrices (rows x wa, rows x wb); part
// holds tall_gram_blocks(rows) * wa * wb doubles.
int tall_gram_blocks(long rows);
int tall_gram_launch(hipStream_t st, const float* A, int wa, const float* B, int wb, long rows, double* part,
                     double* out);
int sketch_solve_launch(hipStream_t st, const double* W, const double* E, const double* xsq, int w, double ridge,
                        double* C1, double* C2s, double* coef, double* loss_acc, int* info);
int tall_small_launch(hipStream_t st, const float* A, long rows, int w, const double* C, const float* B,
                      const double* beta, float* out);
int adam_launch(hipStream_t st, float* S, float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                float scale, float bc1, float bc2, const double* step_loss, double min_loss_sum);
int fold_loss_launch(hipStream_t st, double* step, double* acc);
int fill_f32_launch(hipStream_t st, float* a, long n, float val);
int fill_f64_launch(hipStream_t st, double* a, long n, double val);
// out[0] = sum_i ||P_i - Q QtP_i||^2 (Q: rows x b, P: rows x np, QtP: b x np FP64);
// part holds probe_resid_blocks(rows) doubles.
int probe_resid_blocks(long rows);
int probe_resid_launch(hipStream_t st, const float* Q, const float* P, const double* QtP, long rows, int b, int np,
                       double* part, double* out);

// MFMA lane-layout probe (diagnostic): runs each MFMA shape once on given inputs.
int mfma_probe_launch(hipStream_t st, const float* a, const float* b, float* out_bf16_32, float* out_f32_32,
                      float* out_f32_16, double* out_f64_16, float* out_bf16_16);

}  // namespace ht
