// GEMM-shaped kernels of the HT engine (gfx950).
//
// gemm_atb / gemm_ab: skinny products with the big matricisation A = X_(t) (M x K,
// lda = M) -- the range-finder sketches X_(t) W, the co-range products X_(t)^T Q and the
// subspace-iteration products of the exact solver. They replace the cuBLAS SGEMM calls of
// rsvd (fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:434-444,
// .../Baseline/kernel.cu:380-390). MFMA with FP32 accumulation; 16-bit operands when the
// precision asks for it.
//
// proj_pms: y = x x_t U^T for a (P,M,S) view -- the mode products of the ttm contraction
// (third_order_tensor_decompsition/CUDA/Optimized/ttm.cu:2-36) on small tensors.
//
// lowrank: tiles of C = L R^T with two epilogues: store C (generator, explicit
// reconstruction) or accumulate ||X - C||^2 and ||X||^2 in one pass (the residual that
// the reference forms with SGEMM + Saxpy + two Snrm2:
// third_order_tensor_decompsition/CUDA/Baseline/htd_unopt.cu:202-266).
#include "ht_common.h"

namespace ht {

constexpr int SK_B = 64;  // padded skinny width
constexpr int SK_K = 64;  // reduction chunk

// LDS element type and operand fetch of the skinny products: 16-bit tiles for the BF16 /
// FP16 precisions (converted once, when the chunk is stored), FP32 tiles for FP32.
template <int PREC>
struct SkT {
  typedef uint16_t T;
  static constexpr int LD = SK_K + 8;  // 144-byte rows: 16-byte operand reads without conflicts
};
template <>
struct SkT<PREC_F32> {
  typedef float T;
  static constexpr int LD = SK_K + 4;
};
template <>
struct SkT<PREC_FP8> {
  typedef uint8_t T;
  static constexpr int LD = SK_K + 8;  // 72-byte rows, 8-byte operand reads
};

// Four consecutive values of a tile row into LDS (converted for 16-bit tiles).
template <int PREC>
__device__ __forceinline__ void sk_put4(typename SkT<PREC>::T* dst, float a, float b, float c, float d) {
  if constexpr (PREC == PREC_F32) {
    *reinterpret_cast<float4*>(dst) = make_float4(a, b, c, d);
  } else if constexpr (PREC == PREC_FP8) {
    *reinterpret_cast<uint32_t*>(dst) = fp8x4(a, b, c, d);
  } else {
    uint2 w;
    w.x = cvt2<PREC>(a, b);
    w.y = cvt2<PREC>(c, d);
    *reinterpret_cast<uint2*>(dst) = w;
  }
}

template <int PREC>
__device__ __forceinline__ typename SkT<PREC>::T sk_one(float v) {
  if constexpr (PREC == PREC_F32) return v;
  else if constexpr (PREC == PREC_FP8) return fp8_bits(v);
  else return to16<PREC>(v);
}

// acc[bb] += (rows 32 bb.. of tile R) x (rows 32 w.. of tile C)^T over one 64-deep chunk:
// D[r][c] = sum_k R[r][k] C[c][k], R and C row-major [row][k] in LDS. D's lane layout puts
// consecutive c in consecutive lanes, which the output stores rely on.
template <int PREC, int NB>
__device__ __forceinline__ void sk_chunk(const typename SkT<PREC>::T* __restrict__ R,
                                         const typename SkT<PREC>::T* __restrict__ C, int w, int lane,
                                         f32x16 (&acc)[NB]) {
  constexpr int LD = SkT<PREC>::LD;
  if constexpr (PREC == PREC_F32) {
#pragma unroll 8
    for (int kk = 0; kk < SK_K / 2; ++kk) {
      const float cf = C[(32 * w + (lane & 31)) * LD + 2 * kk + (lane >> 5)];
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        const float rf = R[(32 * bb + (lane & 31)) * LD + 2 * kk + (lane >> 5)];
        acc[bb] = __builtin_amdgcn_mfma_f32_32x32x2f32(rf, cf, acc[bb], 0, 0, 0);
      }
    }
  } else if constexpr (PREC == PREC_FP8) {
#pragma unroll
    for (int kk = 0; kk < SK_K / 16; ++kk) {
      const long cf = *reinterpret_cast<const long*>(&C[(32 * w + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        const long rf = *reinterpret_cast<const long*>(&R[(32 * bb + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)]);
        acc[bb] = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(rf, cf, acc[bb], 0, 0, 0);
      }
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < SK_K / 16; ++kk) {
      const u16x8 cf = *reinterpret_cast<const u16x8*>(&C[(32 * w + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        const u16x8 rf =
            *reinterpret_cast<const u16x8*>(&R[(32 * bb + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)]);
        acc[bb] = mfma32x32x16<PREC>(rf, cf, acc[bb]);
      }
    }
  }
}

// FP8 tiles carry a power-of-two scale each (sk_scales): the wave maxima of the two
// register tiles go to red before the barrier that precedes the LDS stores, and every
// thread derives the block's scales from them. The chunk's MFMA partial sums are then
// scaled back (exactly: powers of two) as they join the running accumulators.
template <int NA, int NBF>
__device__ __forceinline__ void sk_maxima(const float (&ra)[NA], const float (&rb)[NBF], int lane, int w,
                                          float* red) {
  float ma = 0.f, mb = 0.f;
#pragma unroll
  for (int q = 0; q < NA; ++q) ma = fmaxf(ma, fabsf(ra[q]));
#pragma unroll
  for (int q = 0; q < NBF; ++q) mb = fmaxf(mb, fabsf(rb[q]));
  for (int o = 32; o > 0; o >>= 1) {
    ma = fmaxf(ma, __shfl_xor(ma, o, 64));
    mb = fmaxf(mb, __shfl_xor(mb, o, 64));
  }
  if (lane == 0) {
    red[w] = ma;
    red[16 + w] = mb;
  }
}
__device__ __forceinline__ float sk_scale(const float* red, int nw) {
  float m = 0.f;
  for (int i = 0; i < nw; ++i) m = fmaxf(m, red[i]);
  return fp8_scale(m);
}
template <int PREC, int NB>
__device__ __forceinline__ void sk_chunk_acc(const typename SkT<PREC>::T* __restrict__ R,
                                             const typename SkT<PREC>::T* __restrict__ C, int w, int lane,
                                             f32x16 (&acc)[NB], float sr, float sc) {
  if constexpr (PREC == PREC_FP8) {
    f32x16 part[NB];
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) part[bb] = f32x16{};
    sk_chunk<PREC, NB>(R, C, w, lane, part);
    const float ir = 1.f / sr, ic = 1.f / sc;
#pragma unroll
    for (int bb = 0; bb < NB; ++bb)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[bb][e] += (part[bb][e] * ir) * ic;
  } else {
    sk_chunk<PREC, NB>(R, C, w, lane, acc);
  }
}

// Output of a block: acc[bb] holds D[j][c] for j = 32 bb + acc32_row(e, lane), c = c0 +
// 32 w + lane % 32. With one split the values go straight to Y (column-major, rows x b:
// lanes write consecutive rows); otherwise to part[split][j][row] for skinny_sum_kernel.
template <int NB>
__device__ __forceinline__ void sk_write(const f32x16 (&acc)[NB], long c0, int w, int lane, long rows, int b,
                                         int n_split, int split, float* __restrict__ Y, float* __restrict__ part) {
  const long c = c0 + 32 * w + (lane & 31);
  if (c >= rows) return;
  float* dst = n_split == 1 ? Y : part + (long)split * SK_B * rows;
#pragma unroll
  for (int bb = 0; bb < NB; ++bb)
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int j = 32 * bb + acc32_row(e, lane);
      if (j < b) dst[c + rows * j] = acc[bb][e];
    }
}

// ------------------------------------------------------------------ gemm_atb
// Y(K x b) = A^T B, A: M x K (lda M), B: M x b (ldb M). A block owns 128 columns c of A
// (output rows) and all b outputs: 256 threads = 4 waves, wave w the columns 32 w.. (NB
// 32-row accumulators over j). The reduction runs over m in chunks of 64: chunk ch + 1 is
// loaded into registers (every load unconditional, its index clamped into the matrix:
// a load behind a divergent branch is waited for at the join) while chunk ch is
// multiplied; stored values with m >= M are zero, columns c >= K and j >= b are copies
// whose outputs are never written.
// VEC (M % 4 == 0, A and B 16-byte aligned): float4 loads along m.
// TA = uint16_t: A is a 16-bit copy in the precision's format (M % 8 == 0, VEC): 16-byte
// loads of 8 values along m, stored to LDS as they are.
template <int PREC, int NB, bool VEC, typename TA = float>
__global__ __launch_bounds__(256) void gemm_atb_kernel(const TA* __restrict__ A, long M, long K,
                                                       const float* __restrict__ B, int b, int n_split,
                                                       float* __restrict__ Y, float* __restrict__ part) {
  typedef typename SkT<PREC>::T T;
  constexpr int LD = SkT<PREC>::LD;
  constexpr int NBJ = 32 * NB;                    // j rows of the B tile
  constexpr int RA = 32, RB = NBJ * SK_K / 256;   // values per thread per chunk
  constexpr bool A16 = sizeof(TA) == 2;
  __shared__ __attribute__((aligned(16))) T As[128 * LD];
  __shared__ __attribute__((aligned(16))) T Bs[NBJ * LD];
  __shared__ float smax[32];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long c0 = (long)blockIdx.x * 128;
  const int split = blockIdx.y;
  const long n_chunks = (M + SK_K - 1) / SK_K;
  const long k_begin = (long)split * n_chunks / n_split, k_end = (long)(split + 1) * n_chunks / n_split;
  float ra[A16 ? 1 : RA], rb[RB];
  uint32_t rh[A16 ? RA / 2 : 1];
  // VEC: thread element q of A is (col f / 16, m 4 (f % 16) + 0..3), f = tid + 256 (q / 4);
  // scalar: (col e / 64, m e % 64), e = tid + 256 q. B likewise with j for col. 16-bit A:
  // group it is (col f / 8, m 8 (f % 8) + 0..7), f = tid + 256 it.
  auto load = [&](long ch) {
    const long m0 = ch * SK_K;
    if constexpr (A16) {
#pragma unroll
      for (int it = 0; it < RA / 8; ++it) {
        const int f = tid + 256 * it;
        const long c = min(c0 + (f >> 3), K - 1), m = min(m0 + 8 * (f & 7), M - 8);
        const uint4 v = *reinterpret_cast<const uint4*>(A + m + M * c);
        rh[4 * it] = v.x; rh[4 * it + 1] = v.y; rh[4 * it + 2] = v.z; rh[4 * it + 3] = v.w;
      }
    } else if constexpr (VEC) {
#pragma unroll
      for (int it = 0; it < RA / 4; ++it) {
        const int f = tid + 256 * it;
        const long c = min(c0 + (f >> 4), K - 1), m = min(m0 + 4 * (f & 15), M - 4);
        const float4 v = *reinterpret_cast<const float4*>(A + m + M * c);
        ra[4 * it] = v.x; ra[4 * it + 1] = v.y; ra[4 * it + 2] = v.z; ra[4 * it + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < RA; ++q) {
        const int e = tid + 256 * q;
        const long c = min(c0 + (e >> 6), K - 1), m = min(m0 + (e & 63), M - 1);
        ra[q] = A[m + M * c];
      }
    }
    if constexpr (VEC) {
#pragma unroll
      for (int it = 0; it < RB / 4; ++it) {
        const int f = tid + 256 * it;
        const long j = min(f >> 4, b - 1), m = min(m0 + 4 * (f & 15), M - 4);
        const float4 v = *reinterpret_cast<const float4*>(B + m + M * j);
        rb[4 * it] = v.x; rb[4 * it + 1] = v.y; rb[4 * it + 2] = v.z; rb[4 * it + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        const int e = tid + 256 * q;
        const long j = min(e >> 6, b - 1), m = min(m0 + (e & 63), M - 1);
        rb[q] = B[m + M * j];
      }
    }
  };
  auto store = [&](long ch, float sa, float sb) {
    const long m0 = ch * SK_K;
    if constexpr (A16) {
#pragma unroll
      for (int it = 0; it < RA / 8; ++it) {
        const int f = tid + 256 * it;
        const uint32_t keep = m0 + 8 * (f & 7) < M ? 0xffffffffu : 0u;
        *reinterpret_cast<uint4*>(&As[(f >> 3) * LD + 8 * (f & 7)]) =
            make_uint4(rh[4 * it] & keep, rh[4 * it + 1] & keep, rh[4 * it + 2] & keep, rh[4 * it + 3] & keep);
      }
    } else if constexpr (VEC) {
#pragma unroll
      for (int it = 0; it < RA / 4; ++it) {
        const int f = tid + 256 * it;
        const bool ok = m0 + 4 * (f & 15) < M;
        sk_put4<PREC>(&As[(f >> 4) * LD + 4 * (f & 15)], ok ? ra[4 * it] * sa : 0.f, ok ? ra[4 * it + 1] * sa : 0.f,
                      ok ? ra[4 * it + 2] * sa : 0.f, ok ? ra[4 * it + 3] * sa : 0.f);
      }
    } else {
      const bool ok = m0 + (tid & 63) < M;  // e % 64 == tid % 64 for every q
#pragma unroll
      for (int q = 0; q < RA; ++q) {
        const int e = tid + 256 * q;
        As[(e >> 6) * LD + (e & 63)] = sk_one<PREC>(ok ? ra[q] * sa : 0.f);
      }
    }
    if constexpr (VEC) {
#pragma unroll
      for (int it = 0; it < RB / 4; ++it) {
        const int f = tid + 256 * it;
        const bool ok = m0 + 4 * (f & 15) < M;
        sk_put4<PREC>(&Bs[(f >> 4) * LD + 4 * (f & 15)], ok ? rb[4 * it] * sb : 0.f, ok ? rb[4 * it + 1] * sb : 0.f,
                      ok ? rb[4 * it + 2] * sb : 0.f, ok ? rb[4 * it + 3] * sb : 0.f);
      }
    } else {
      const bool ok = m0 + (tid & 63) < M;
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        const int e = tid + 256 * q;
        Bs[(e >> 6) * LD + (e & 63)] = sk_one<PREC>(ok ? rb[q] * sb : 0.f);
      }
    }
  };
  f32x16 acc[NB];
#pragma unroll
  for (int bb = 0; bb < NB; ++bb) acc[bb] = f32x16{};
  if (k_begin < k_end) load(k_begin);
  for (long ch = k_begin; ch < k_end; ++ch) {
    if constexpr (PREC == PREC_FP8) sk_maxima(ra, rb, lane, w, smax);
    __syncthreads();
    const float sa = PREC == PREC_FP8 ? sk_scale(smax, 4) : 1.f, sb = PREC == PREC_FP8 ? sk_scale(smax + 16, 4) : 1.f;
    store(ch, sa, sb);
    __syncthreads();
    load(ch + 1 < k_end ? ch + 1 : ch);  // the last chunk again at the end (never stored)
    __builtin_amdgcn_sched_barrier(0);
    sk_chunk_acc<PREC, NB>(Bs, As, w, lane, acc, sb, sa);
    __builtin_amdgcn_sched_barrier(0);
  }
  sk_write<NB>(acc, c0, w, lane, K, b, n_split, split, Y, part);
}

// ------------------------------------------------------------------ gemm_ab
// Y(M x b) = A B, A: M x K (lda M), B: K x b (ldb K). A block owns 256 rows m (output
// rows) and all b outputs: 512 threads = 8 waves, wave w the rows 32 w... The reduction
// runs over columns k of A in chunks of 64, pipelined as in gemm_atb. The A tile is
// transposed on its way to LDS ([m][k]): thread t reads row m0 + t % 256 of the 32
// columns 32 (t / 256).. (a wave's load reads 64 consecutive rows of a column) and
// stores them as one row piece. Rows m >= M are copies whose outputs are never
// written; columns k >= K are stored as zero.
// VEC (K % 4 == 0, B 16-byte aligned): float4 loads of B along k.
// TA = uint16_t: A is a 16-bit copy (M % 4 == 0, 8-byte aligned): thread t reads rows
// 4 (t % 64) .. + 3 of the 8 columns 8 (t / 64).. (8-byte loads; a wave reads 512
// consecutive bytes of a column) and transposes them in registers.
template <int PREC, int NB, bool VEC, typename TA = float>
__global__ __launch_bounds__(512) void gemm_ab_kernel(const TA* __restrict__ A, long M, long K,
                                                      const float* __restrict__ B, int b, int n_split,
                                                      float* __restrict__ Y, float* __restrict__ part) {
  typedef typename SkT<PREC>::T T;
  constexpr int LD = SkT<PREC>::LD;
  constexpr int NBJ = 32 * NB;
  constexpr int RB = NBJ * SK_K / 512;
  constexpr bool A16 = sizeof(TA) == 2;
  __shared__ __attribute__((aligned(16))) T As[256 * LD];
  __shared__ __attribute__((aligned(16))) T Bs[NBJ * LD];
  __shared__ float smax[32];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long r0 = (long)blockIdx.x * 256;
  const int split = blockIdx.y;
  const long n_chunks = (K + SK_K - 1) / SK_K;
  const long k_begin = (long)split * n_chunks / n_split, k_end = (long)(split + 1) * n_chunks / n_split;
  const long arow = A16 ? min(r0 + 4 * (tid & 63), M - 4) : min(r0 + (tid & 255), M - 1);
  const int kh = A16 ? 8 * (tid >> 6) : 32 * (tid >> 8);
  float ra[A16 ? 1 : 32], rb[RB];
  uint32_t rh[A16 ? 16 : 1];
  auto load = [&](long ch) {
    const long k0 = ch * SK_K;
    const TA* a = A + arow;
    if constexpr (A16) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint2 v = *reinterpret_cast<const uint2*>(a + M * min(k0 + kh + q, K - 1));
        rh[2 * q] = v.x; rh[2 * q + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 32; ++q) ra[q] = a[M * min(k0 + kh + q, K - 1)];
    }
    if constexpr (VEC) {
#pragma unroll
      for (int it = 0; it < RB / 4; ++it) {
        const int f = tid + 512 * it;
        const long j = min(f >> 4, b - 1), k = min(k0 + 4 * (f & 15), K - 4);
        const float4 v = *reinterpret_cast<const float4*>(B + k + K * j);
        rb[4 * it] = v.x; rb[4 * it + 1] = v.y; rb[4 * it + 2] = v.z; rb[4 * it + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        const int e = tid + 512 * q;
        const long j = min(e >> 6, b - 1), k = min(k0 + (e & 63), K - 1);
        rb[q] = B[k + K * j];
      }
    }
  };
  auto store = [&](long ch, float sa, float sb) {
    const long k0 = ch * SK_K;
    const long nk = K - (k0 + kh);  // valid columns among this thread's
    if constexpr (A16) {
      // rh[2 q] = rows 0, 1 of column q (low half row 0), rh[2 q + 1] = rows 2, 3
      uint32_t c[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) c[i] = (i >> 1) < nk ? rh[i] : 0u;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int h = r >> 1;
        const bool hi = (r & 1) != 0;
        uint32_t wv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t x0 = c[2 * (2 * j) + h], x1 = c[2 * (2 * j + 1) + h];
          wv[j] = hi ? ((x0 >> 16) | (x1 & 0xffff0000u)) : ((x0 & 0xffffu) | (x1 << 16));
        }
        *reinterpret_cast<uint4*>(&As[(4 * (tid & 63) + r) * LD + kh]) = make_uint4(wv[0], wv[1], wv[2], wv[3]);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 32; q += 4)
        sk_put4<PREC>(&As[(tid & 255) * LD + kh + q], q < nk ? ra[q] * sa : 0.f, q + 1 < nk ? ra[q + 1] * sa : 0.f,
                      q + 2 < nk ? ra[q + 2] * sa : 0.f, q + 3 < nk ? ra[q + 3] * sa : 0.f);
    }
    if constexpr (VEC) {
#pragma unroll
      for (int it = 0; it < RB / 4; ++it) {
        const int f = tid + 512 * it;
        const bool ok = k0 + 4 * (f & 15) < K;
        sk_put4<PREC>(&Bs[(f >> 4) * LD + 4 * (f & 15)], ok ? rb[4 * it] * sb : 0.f, ok ? rb[4 * it + 1] * sb : 0.f,
                      ok ? rb[4 * it + 2] * sb : 0.f, ok ? rb[4 * it + 3] * sb : 0.f);
      }
    } else {
      const bool ok = k0 + (tid & 63) < K;
#pragma unroll
      for (int q = 0; q < RB; ++q) {
        const int e = tid + 512 * q;
        Bs[(e >> 6) * LD + (e & 63)] = sk_one<PREC>(ok ? rb[q] * sb : 0.f);
      }
    }
  };
  f32x16 acc[NB];
#pragma unroll
  for (int bb = 0; bb < NB; ++bb) acc[bb] = f32x16{};
  if (k_begin < k_end) load(k_begin);
  for (long ch = k_begin; ch < k_end; ++ch) {
    if constexpr (PREC == PREC_FP8) sk_maxima(ra, rb, lane, w, smax);
    __syncthreads();
    const float sa = PREC == PREC_FP8 ? sk_scale(smax, 8) : 1.f, sb = PREC == PREC_FP8 ? sk_scale(smax + 16, 8) : 1.f;
    store(ch, sa, sb);
    __syncthreads();
    load(ch + 1 < k_end ? ch + 1 : ch);
    __builtin_amdgcn_sched_barrier(0);
    sk_chunk_acc<PREC, NB>(Bs, As, w, lane, acc, sb, sa);
    __builtin_amdgcn_sched_barrier(0);
  }
  sk_write<NB>(acc, r0, w, lane, M, b, n_split, split, Y, part);
}

// Y[r + rows j] = sum over splits of part[s][j][r] (FP64 sum, fixed order), j < b.
__global__ void skinny_sum_kernel(const float* __restrict__ part, long rows, int b, int n_split,
                                  float* __restrict__ Y) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * b) return;
  double s = 0.0;
  for (int k = 0; k < n_split; ++k) s += (double)part[(long)k * rows * SK_B + e];
  Y[e] = (float)s;
}

// Splits of the reduction: enough blocks to fill the chip (target blocks), at most 256
// and at most one per chunk. The splits' FP32 partials are summed in FP64, so short
// reductions split finely also keep their FP32 rounding small.
static int pick_split(long tiles, long chunks, long target) {
  long s = target / (tiles > 0 ? tiles : 1);
  if (s > chunks) s = chunks;
  if (s > 256) s = 256;
  if (s < 1) s = 1;
  return (int)s;
}
static constexpr long ATB_TARGET = 1024, AB_TARGET = 512;

size_t gemm_ws_floats(long M, long K, int b) {
  (void)b;
  const long s1 = pick_split((K + 127) / 128, (M + SK_K - 1) / SK_K, ATB_TARGET);
  const long s2 = pick_split((M + 255) / 256, (K + SK_K - 1) / SK_K, AB_TARGET);
  const long a = s1 > 1 ? s1 * K * SK_B : 0, c = s2 > 1 ? s2 * M * SK_B : 0;
  return (size_t)std::max(std::max(a, c), 1L);
}

template <int PREC, int NB, typename TA>
static void atb_dispatch(hipStream_t st, dim3 grid, bool vec, const TA* A, long M, long K, const float* B, int b,
                         int ns, float* Y, float* ws) {
  if (vec)
    hipLaunchKernelGGL((gemm_atb_kernel<PREC, NB, true, TA>), grid, dim3(256), 0, st, A, M, K, B, b, ns, Y, ws);
  else if constexpr (sizeof(TA) == 4)
    hipLaunchKernelGGL((gemm_atb_kernel<PREC, NB, false, TA>), grid, dim3(256), 0, st, A, M, K, B, b, ns, Y, ws);
}
template <int PREC, int NB, typename TA>
static void ab_dispatch(hipStream_t st, dim3 grid, bool vec, const TA* A, long M, long K, const float* B, int b,
                        int ns, float* Y, float* ws) {
  if (vec)
    hipLaunchKernelGGL((gemm_ab_kernel<PREC, NB, true, TA>), grid, dim3(512), 0, st, A, M, K, B, b, ns, Y, ws);
  else
    hipLaunchKernelGGL((gemm_ab_kernel<PREC, NB, false, TA>), grid, dim3(512), 0, st, A, M, K, B, b, ns, Y, ws);
}

template <typename TA>
static int atb_launch(hipStream_t st, const TA* A, long M, long K, const float* B, int b, int prec, float* Y,
                      float* ws, size_t ws_floats, bool vec) {
  const long tiles = (K + 127) / 128, chunks = (M + SK_K - 1) / SK_K;
  const int ns = pick_split(tiles, chunks, ATB_TARGET);
  if (ns > 1 && (size_t)ns * K * SK_B > ws_floats) return 4;
  dim3 grid((unsigned)tiles, ns);
  const bool nb1 = b <= 32;
  if (prec == PREC_BF16) {
    if (nb1) atb_dispatch<PREC_BF16, 1>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
    else atb_dispatch<PREC_BF16, 2>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
  } else if (prec == PREC_F16) {
    if (nb1) atb_dispatch<PREC_F16, 1>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
    else atb_dispatch<PREC_F16, 2>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
  } else if constexpr (sizeof(TA) == 4) {
    if (prec == PREC_FP8) {
      if (nb1) atb_dispatch<PREC_FP8, 1>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
      else atb_dispatch<PREC_FP8, 2>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
    } else {
      if (nb1) atb_dispatch<PREC_F32, 1>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
      else atb_dispatch<PREC_F32, 2>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
    }
  }
  if (ns > 1) {
    const long tot = K * b;
    hipLaunchKernelGGL(skinny_sum_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, ws, K, b, ns, Y);
  }
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

template <typename TA>
static int ab_launch(hipStream_t st, const TA* A, long M, long K, const float* B, int b, int prec, float* Y,
                     float* ws, size_t ws_floats) {
  const long tiles = (M + 255) / 256, chunks = (K + SK_K - 1) / SK_K;
  const int ns = pick_split(tiles, chunks, AB_TARGET);
  if (ns > 1 && (size_t)ns * M * SK_B > ws_floats) return 4;
  const bool vec = K % 4 == 0 && ((uintptr_t)B & 15) == 0;
  dim3 grid((unsigned)tiles, ns);
  const bool nb1 = b <= 32;
  if (prec == PREC_BF16) {
    if (nb1) ab_dispatch<PREC_BF16, 1>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
    else ab_dispatch<PREC_BF16, 2>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
  } else if (prec == PREC_F16) {
    if (nb1) ab_dispatch<PREC_F16, 1>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
    else ab_dispatch<PREC_F16, 2>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
  } else if constexpr (sizeof(TA) == 4) {
    if (prec == PREC_FP8) {
      if (nb1) ab_dispatch<PREC_FP8, 1>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
      else ab_dispatch<PREC_FP8, 2>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
    } else {
      if (nb1) ab_dispatch<PREC_F32, 1>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
      else ab_dispatch<PREC_F32, 2>(st, grid, vec, A, M, K, B, b, ns, Y, ws);
    }
  }
  if (ns > 1) {
    const long tot = M * b;
    hipLaunchKernelGGL(skinny_sum_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, ws, M, b, ns, Y);
  }
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int gemm_atb_launch(hipStream_t st, const float* A, long M, long K, const float* B, int b, int prec, float* Y,
                    float* ws, size_t ws_floats) {
  if (b < 1 || b > SK_B || M < 1 || K < 1) return 8;
  const bool vec = M % 4 == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0;
  return atb_launch(st, A, M, K, B, b, prec, Y, ws, ws_floats, vec);
}

int gemm_ab_launch(hipStream_t st, const float* A, long M, long K, const float* B, int b, int prec, float* Y,
                   float* ws, size_t ws_floats) {
  if (b < 1 || b > SK_B || M < 1 || K < 1) return 8;
  return ab_launch(st, A, M, K, B, b, prec, Y, ws, ws_floats);
}

bool gemm16_ok(const uint16_t* A, long M, const float* B, int op_atb) {
  if (op_atb) return M % 8 == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0;
  return M % 4 == 0 && ((uintptr_t)A & 7) == 0;
}

int gemm_atb16_launch(hipStream_t st, const uint16_t* A, long M, long K, const float* B, int b, int prec, float* Y,
                      float* ws, size_t ws_floats) {
  if (b < 1 || b > SK_B || M < 1 || K < 1 || (prec != PREC_BF16 && prec != PREC_F16) || !gemm16_ok(A, M, B, 1))
    return 8;
  return atb_launch(st, A, M, K, B, b, prec, Y, ws, ws_floats, true);
}

int gemm_ab16_launch(hipStream_t st, const uint16_t* A, long M, long K, const float* B, int b, int prec, float* Y,
                     float* ws, size_t ws_floats) {
  if (b < 1 || b > SK_B || M < 1 || K < 1 || (prec != PREC_BF16 && prec != PREC_F16) || !gemm16_ok(A, M, B, 0))
    return 8;
  return ab_launch(st, A, M, K, B, b, prec, Y, ws, ws_floats);
}

// ------------------------------------------------------------------ proj_pms
// One thread per output element (p, j, s); consecutive threads take consecutive p.
// U[m, j] is read at U[m * um + j * uj] (um = 1, uj = M: column-major M x r).
template <typename TI, typename TU, typename TO>
__global__ void proj_pms_kernel(const TI* __restrict__ x, PMS v, const TU* __restrict__ U, long um, long uj, int r,
                                TO* __restrict__ y) {
  const long total = v.P * (long)r * v.S;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long p = e % v.P;
    const long rest = e / v.P;
    const int j = (int)(rest % r);
    const long s = rest / r;
    const TI* xp = x + p + v.P * v.M * s;
    const TU* up = U + uj * (long)j;
    double acc = 0.0;
    for (long m = 0; m < v.M; ++m) acc += (double)up[um * m] * (double)xp[v.P * m];
    y[e] = (TO)acc;
  }
}

// The projection as FP64 MFMA tiles: y[j, n] = sum_m U[m, j] X[m, n] with n = p + P s and
// X[m, n] = x[p + P m + P M s]. One wave per 16 (j) x 16 (n) output tile, 16x16x4 steps
// over m in batches of eight whose loads all issue before the products (indices past M,
// r or N are clamped and their values zeroed). A lane's n (hence p, s and the base of
// its X column) is fixed for the whole tile.
template <typename TI, typename TU, typename TO>
__device__ __forceinline__ void proj_mfma_tile(const TI* __restrict__ x, PMS v, const TU* __restrict__ U, long um,
                                               long uj, int r, TO* __restrict__ y, long ntn, long wave) {
  const int lane = threadIdx.x & 63;
  const long njt = (r + 15) / 16;
  if (wave >= njt * ntn) return;
  const long NT = wave / njt, JT = wave % njt;
  const int i16 = lane & 15, kk = lane >> 4;
  const long M = v.M, N = v.P * v.S;
  const long n = NT * 16 + i16;
  const bool nok = n < N;
  const long nc = nok ? n : N - 1;
  const long ss = nc / v.P, pp = nc - ss * v.P;
  const TI* xb = x + pp + v.P * M * ss;
  const long j = JT * 16 + i16;
  const bool jok = j < r;
  const TU* ub = U + uj * (jok ? j : r - 1);
  constexpr int UB = 8;
  f64x4 acc = {};
  for (long m0 = 0; m0 < M; m0 += 4 * UB) {
    double a[UB], b[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      const long m = m0 + 4 * u + kk;
      const long mc = m < M ? m : M - 1;
      a[u] = (double)ub[um * mc];
      b[u] = (double)xb[v.P * mc];
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      const bool ok = m0 + 4 * u + kk < M;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ok && jok ? a[u] : 0.0, ok && nok ? b[u] : 0.0, acc, 0, 0, 0);
    }
  }
  if (nok) {
    TO* yb = y + pp + v.P * (long)r * ss;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const long jj = JT * 16 + acc16_row_f64(e, lane);
      if (jj < r) yb[v.P * jj] = (TO)acc[e];
    }
  }
}
template <typename TI, typename TU, typename TO>
__global__ __launch_bounds__(256) void proj_mfma_kernel(const TI* __restrict__ x, PMS v, const TU* __restrict__ U,
                                                        long um, long uj, int r, TO* __restrict__ y, long ntn) {
  proj_mfma_tile(x, v, U, um, uj, r, y, ntn, ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
}
// Two independent FP64 projections in one launch: blocks below j0.blocks take job 0.
struct ProjJob {
  const double* x;
  PMS v;
  const double* U;
  long um, uj;
  int r;
  double* y;
  long ntn, blocks;
};
__global__ __launch_bounds__(256) void proj_mfma2_kernel(ProjJob j0, ProjJob j1) {
  const bool second = blockIdx.x >= j0.blocks;
  const ProjJob& j = second ? j1 : j0;
  const long blk = second ? blockIdx.x - j0.blocks : blockIdx.x;
  proj_mfma_tile(j.x, j.v, j.U, j.um, j.uj, j.r, j.y, j.ntn, (blk * blockDim.x + threadIdx.x) >> 6);
}

template <typename TI, typename TU, typename TO>
static int proj_launch(hipStream_t st, const TI* x, PMS v, const TU* U, long um, long uj, int r, TO* y) {
  const long total = v.P * (long)r * v.S;
  if (total <= 0) return 0;
  {
    const long ntn = (v.P * v.S + 15) / 16, waves = ((r + 15) / 16) * ntn;
    const long blocks = (waves + 3) / 4;
    if (blocks <= (1L << 30)) {
      hipLaunchKernelGGL((proj_mfma_kernel<TI, TU, TO>), dim3((unsigned)blocks), dim3(256), 0, st, x, v, U, um, uj, r,
                         y, ntn);
      return hipGetLastError() == hipSuccess ? 0 : 5;
    }
  }
  long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((proj_pms_kernel<TI, TU, TO>), dim3((unsigned)blocks), dim3(256), 0, st, x, v, U, um, uj, r, y);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
int proj_pms_f32_launch(hipStream_t st, const float* x, PMS v, const float* U, int r, float* y) {
  return proj_launch(st, x, v, U, 1, v.M, r, y);
}
int proj_pms_f64_launch(hipStream_t st, const double* x, PMS v, const double* U, int r, double* y) {
  return proj_launch(st, x, v, U, 1, v.M, r, y);
}
int proj_pms_f64t_launch(hipStream_t st, const double* x, PMS v, const double* Ut, int r, double* y) {
  return proj_launch(st, x, v, Ut, r, 1, r, y);
}
int proj_pms_f64t_f32out_launch(hipStream_t st, const double* x, PMS v, const double* Ut, int r, float* y) {
  return proj_launch(st, x, v, Ut, r, 1, r, y);
}
int proj_pms_f64t_pair_launch(hipStream_t st, const double* x0, PMS v0, const double* Ut0, int r0, double* y0,
                              const double* x1, PMS v1, const double* Ut1, int r1, double* y1) {
  ProjJob j[2] = {{x0, v0, Ut0, r0, 1, r0, y0, 0, 0}, {x1, v1, Ut1, r1, 1, r1, y1, 0, 0}};
  for (auto& q : j) {
    q.ntn = (q.v.P * q.v.S + 15) / 16;
    q.blocks = (((q.r + 15) / 16) * q.ntn + 3) / 4;
    if (q.v.P * (long)q.r * q.v.S <= 0 || q.blocks > (1L << 29)) return 8;
  }
  hipLaunchKernelGGL(proj_mfma2_kernel, dim3((unsigned)(j[0].blocks + j[1].blocks)), dim3(256), 0, st, j[0], j[1]);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
int proj_pms_f32in_f64_launch(hipStream_t st, const float* x, PMS v, const double* U, int r, double* y) {
  return proj_launch(st, x, v, U, 1, v.M, r, y);
}

// ------------------------------------------------------------------ lowrank
// C^T tile (128 n x 128 m) = R L^T so that the accumulator's column index (lane%32) runs
// along m, the contiguous dimension of X and of the output. LDS tiles are [k][row].
constexpr int LR_T = 128, LR_K = 16;
template <int MODE>
__global__ __launch_bounds__(256) void lowrank_kernel(const float* __restrict__ L, const float* __restrict__ R,
                                                      long M, long N, int k, const float* __restrict__ X,
                                                      float* __restrict__ out, double* __restrict__ acc_out) {
  __shared__ float Ls[LR_K * LR_T];
  __shared__ float Rs[LR_K * LR_T];
  __shared__ double red[8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1;  // wave quadrant: n rows 64wr.., m cols 64wc..
  const long m0 = (long)blockIdx.x * LR_T, n0 = (long)blockIdx.y * LR_T;
  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x16{};
  for (int k0 = 0; k0 < k; k0 += LR_K) {
    __syncthreads();
    for (int e = tid; e < LR_K * LR_T; e += 256) {
      const int kk = e / LR_T, row = e % LR_T;
      const long m = m0 + row, n = n0 + row;
      Ls[e] = (k0 + kk < k && m < M) ? L[m + M * (long)(k0 + kk)] : 0.f;
      Rs[e] = (k0 + kk < k && n < N) ? R[n + N * (long)(k0 + kk)] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < LR_K / 2; ++kk) {
      float af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = Rs[(2 * kk + (lane >> 5)) * LR_T + 64 * wr + 32 * a + (lane & 31)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Ls[(2 * kk + (lane >> 5)) * LR_T + 64 * wc + 32 * b + (lane & 31)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
  }
  double se = 0.0, sx = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const long n = n0 + 64 * wr + 32 * a + acc32_row(e, lane);
        const long m = m0 + 64 * wc + 32 * b + (lane & 31);
        if (n < N && m < M) {
          if constexpr (MODE == 0) {
            out[m + M * n] = acc[a][b][e];
          } else {
            const float xv = X[m + M * n];
            const float d = xv - acc[a][b][e];
            se += (double)d * d;
            sx += (double)xv * xv;
          }
        }
      }
  if constexpr (MODE == 1) {
    se = wave_sum(se);
    sx = wave_sum(sx);
    if (lane == 0) { red[w] = se; red[4 + w] = sx; }
    __syncthreads();
    if (tid == 0) {
      const long bid = (long)blockIdx.y * gridDim.x + blockIdx.x;
      acc_out[2 * bid] = red[0] + red[1] + red[2] + red[3];
      acc_out[2 * bid + 1] = red[4] + red[5] + red[6] + red[7];
    }
  }
}

template <int KP, bool FULL>
__global__ void residual_kernel(const float* __restrict__ L, const float* __restrict__ R, long M, long N, int k,
                                const float* __restrict__ X, int nts, double* __restrict__ acc_out);
template <int KP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KP <= 16 ? 4 : 2, KP <= 16 ? 8 : 2))) void
residual_stream_kernel(const float* __restrict__ L, const float* __restrict__ R, long M, long N, int k,
                       const float* __restrict__ X, int nts, double* __restrict__ acc_out);

int lowrank_launch(hipStream_t st, int mode, const float* L, const float* R, long M, long N, int k,
                   const float* X, float* out, double* acc, int* n_blocks_out) {
  if (M < 1 || N < 1 || k < 1 || k > 4096) return 8;
  if (mode == 1 && M > (1L << 24)) return 8;  // 32-bit column offsets of the residual pass
  dim3 grid((unsigned)((M + LR_T - 1) / LR_T), (unsigned)((N + LR_T - 1) / LR_T));
  if (mode == 0) {
    if (n_blocks_out) *n_blocks_out = (int)(grid.x * grid.y);
    hipLaunchKernelGGL(lowrank_kernel<0>, grid, dim3(256), 0, st, L, R, M, N, k, X, out, acc);
    return hipGetLastError() == hipSuccess ? 0 : 5;
  }
  if (k > 64) return 8;
  const bool full = M % LR_T == 0 && N % LR_T == 0;
  // Residual blocks: 128 rows x 32 nts columns (nts <= 32). Waves that stream more
  // subtiles keep their prefetch busy for longer; nts shrinks while the grid would fall
  // below 2048 blocks (8 per CU).
  int nts = 4;
  if (full) {
    nts = 32;
    while (nts > 4 && (N % (32L * nts) != 0 || grid.x * (N / (32L * nts)) < 2048)) nts /= 2;
    grid.y = (unsigned)(N / (32L * nts));
  }
  if (n_blocks_out) *n_blocks_out = (int)(grid.x * grid.y);
  if (full) {
    if (k <= 16)
      hipLaunchKernelGGL(residual_stream_kernel<16>, grid, dim3(256), 0, st, L, R, M, N, k, X, nts, acc);
    else if (k <= 32)
      hipLaunchKernelGGL(residual_stream_kernel<32>, grid, dim3(256), 0, st, L, R, M, N, k, X, nts, acc);
    else
      hipLaunchKernelGGL(residual_stream_kernel<64>, grid, dim3(256), 0, st, L, R, M, N, k, X, nts, acc);
  } else if (k <= 16)
    hipLaunchKernelGGL((residual_kernel<16, false>), grid, dim3(256), 0, st, L, R, M, N, k, X, nts, acc);
  else if (k <= 32)
    hipLaunchKernelGGL((residual_kernel<32, false>), grid, dim3(256), 0, st, L, R, M, N, k, X, nts, acc);
  else
    hipLaunchKernelGGL((residual_kernel<64, false>), grid, dim3(256), 0, st, L, R, M, N, k, X, nts, acc);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// Residual pass: acc[2*block] = sum (X - L R^T)^2, acc[2*block+1] = sum X^2 over one
// 128 (m) x 32 nts (n) region of X (ld M). No LDS: wave w takes the m strip
// [m0 + 32 w, +32) and nts 32-column subtiles; the MFMA operands (R rows as A, L rows
// as B, k <= KP) come straight from global memory (L2-resident), and X is read in the
// accumulator's own layout (element e of lane l is (m0 + l % 32, n0 + acc32_row(e, l)):
// each half-wave reads one 128-byte column segment), so the comparison needs no
// transposition. FULL (M and N multiples of 128): no bounds checks. Loads are never
// behind a per-lane branch (indices past k are clamped and the values zeroed at their
// use): a branch join would make the compiler wait for them on the spot.
template <int KP, bool FULL>
__global__ __launch_bounds__(256) void residual_kernel(const float* __restrict__ L, const float* __restrict__ R, long M,
                                                       long N, int k, const float* __restrict__ X, int nts,
                                                       double* __restrict__ acc_out) {
  __shared__ double red[8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 31, kh = lane >> 5;
  const long mw = (long)blockIdx.x * 128 + 32 * w;  // wave's first row
  const long m = mw + li;
  const long nbase = (long)blockIdx.y * 32 * nts;
  const bool mok = FULL || m < M;
  const long mc = mok ? m : M - 1;
  float bl[KP / 2];
#pragma unroll
  for (int j = 0; j < KP / 2; ++j) {
    const int kk = 2 * j + kh;
    const float v = L[mc + M * (long)min(kk, k - 1)];
    bl[j] = (kk < k && mok) ? v : 0.f;
  }
  // X element e of this lane in subtile nb: (m, nb + 8 (e / 4) + 4 kh + e % 4); the column
  // part of the offset is a 32-bit multiple of M from the wave-uniform base X + mw + M nb.
  const unsigned mu = (unsigned)M;
  auto load_x = [&](float* xv, int t) {
    const long nb = nbase + 32 * t;
    const float* base = X + mw + M * nb;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int row = 8 * (e >> 2) + 4 * kh + (e & 3);
      if (FULL) {
        xv[e] = base[(unsigned)li + mu * (unsigned)row];
      } else {
        const bool ok = mok && nb + row < N;
        const float v = X[ok ? mw + li + M * (nb + row) : 0];
        xv[e] = ok ? v : 0.f;
      }
    }
  };
  double se = 0.0, sx = 0.0;
  auto tile = [&](const float* xv, int t) {
    const long nb = nbase + 32 * t;
    const long n = nb + li;
    const long nc = (FULL || n < N) ? n : N - 1;
    float ar[KP / 2];
#pragma unroll
    for (int j = 0; j < KP / 2; ++j) {
      const int kk = 2 * j + kh;
      const float v = R[nc + N * (long)min(kk, k - 1)];
      ar[j] = (kk < k && (FULL || n < N)) ? v : 0.f;
    }
    f32x16 acc = f32x16{};
#pragma unroll
    for (int j = 0; j < KP / 2; ++j) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[j], bl[j], acc, 0, 0, 0);
    float e2 = 0.f, x2 = 0.f;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const float d = xv[e] - acc[e];  // out-of-range elements: X = 0 and acc = 0
      e2 += d * d;
      x2 += xv[e] * xv[e];
    }
    se += (double)e2;
    sx += (double)x2;
  };
#pragma unroll 1
  for (int t = 0; t < nts; ++t) {
    float xv[16];
    load_x(xv, t);
    tile(xv, t);
  }
  se = wave_sum(se);
  sx = wave_sum(sx);
  if (lane == 0) { red[w] = se; red[4 + w] = sx; }
  __syncthreads();
  if (tid == 0) {
    const long bid = (long)blockIdx.y * gridDim.x + blockIdx.x;
    acc_out[2 * bid] = red[0] + red[1] + red[2] + red[3];
    acc_out[2 * bid + 1] = red[4] + red[5] + red[6] + red[7];
  }
}

// The residual pass for k <= KP (16, 32 or 64) with M a multiple of 128 and N of 32 nts:
// the same block, wave and lane layout as residual_kernel, no bounds checks, and the
// next subtile's X and R rows are loaded while the current one is multiplied. X is read
// once, with streaming (nontemporal) loads; each half-wave load covers whole 128-byte
// lines. The subtile after the last is the last again (loaded, never used), so every
// path issues the same loads and each wait covers only the subtile being multiplied; a
// scheduling barrier keeps the compiler from sinking the loads below the products.
// Wider KP holds more operands in flight: two waves per SIMD leave it 256 registers.
template <int KP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KP <= 16 ? 4 : 2, KP <= 16 ? 8 : 2))) void
residual_stream_kernel(
    const float* __restrict__ L, const float* __restrict__ R, long M, long N, int k, const float* __restrict__ X,
    int nts, double* __restrict__ acc_out) {
  __shared__ double red[8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 31, kh = lane >> 5;
  const long mw = (long)blockIdx.x * 128 + 32 * w;
  const long m = mw + li;
  const long nbase = (long)blockIdx.y * 32 * nts;
  float bl[KP / 2];
#pragma unroll
  for (int j = 0; j < KP / 2; ++j) {
    const int kk = 2 * j + kh;
    const float v = L[m + M * (long)min(kk, k - 1)];
    bl[j] = kk < k ? v : 0.f;
  }
  const unsigned mu = (unsigned)M;
  auto load_x = [&](float* xv, int t) {
    const float* base = X + mw + M * (nbase + 32 * t);
#pragma unroll
    for (int e = 0; e < 16; ++e)
      xv[e] = __builtin_nontemporal_load(base + ((unsigned)li + mu * (unsigned)acc32_row(e, lane)));
  };
  // R rows of subtile t as the MFMA's A operand (columns past k are copies of column
  // k - 1, zeroed at their use)
  auto load_r = [&](float* rv, int t) {
    const long n = nbase + 32 * t + li;
#pragma unroll
    for (int j = 0; j < KP / 2; ++j) rv[j] = R[n + N * (long)min(2 * j + kh, k - 1)];
  };
  double se = 0.0, sx = 0.0;
  auto tile = [&](const float* xv, const float* rv) {
    f32x16 acc = f32x16{};
#pragma unroll
    for (int j = 0; j < KP / 2; ++j)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(2 * j + kh < k ? rv[j] : 0.f, bl[j], acc, 0, 0, 0);
    float e2 = 0.f, x2 = 0.f;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const float d = xv[e] - acc[e];
      e2 += d * d;
      x2 += xv[e] * xv[e];
    }
    se += (double)e2;
    sx += (double)x2;
  };
  float xa[16], xb[16], ra[KP / 2], rb[KP / 2];
  load_x(xa, 0);
  load_r(ra, 0);
  for (int t = 0; t < nts; t += 2) {
    load_x(xb, min(t + 1, nts - 1));
    load_r(rb, min(t + 1, nts - 1));
    __builtin_amdgcn_sched_barrier(0);
    tile(xa, ra);
    load_x(xa, min(t + 2, nts - 1));
    load_r(ra, min(t + 2, nts - 1));
    __builtin_amdgcn_sched_barrier(0);
    if (t + 1 < nts) tile(xb, rb);
  }
  se = wave_sum(se);
  sx = wave_sum(sx);
  if (lane == 0) { red[w] = se; red[4 + w] = sx; }
  __syncthreads();
  if (tid == 0) {
    const long bid = (long)blockIdx.y * gridDim.x + blockIdx.x;
    acc_out[2 * bid] = red[0] + red[1] + red[2] + red[3];
    acc_out[2 * bid + 1] = red[4] + red[5] + red[6] + red[7];
  }
}

}  // namespace ht
