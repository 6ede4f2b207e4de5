// Leaf Gram kernels of the HT engine (gfx950, MFMA).
//
// gram_pms: G = X_(t) X_(t)^T for the (P,M,S) view of a mode range t, read straight
// from the column-major tensor into LDS tiles with coalesced strided loads -- no
// permuted copy of X is ever written (the reference materialises mode-2/3 unfoldings
// first: third_order_tensor_decompsition/CUDA/Baseline/kernel.cu:45-69,
// fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:36-63, then
// calls GemmEx: .../Optimized/htd4_tensorcore.cu:61-103).
//
// slab_gram: the Grams of the two leading modes from one read of each n1 x n2 slab.
// slab_proj: z_s = U1^T S_s U2 per slab (the leaf-core projection of the two leading
// modes), one wavefront per slab, FP32 MFMA, U1 and U2 staged in LDS.
#include <algorithm>

#include "ht_common.h"

namespace ht {

// ------------------------------------------------------------------ MFMA layout probe
// Runs each MFMA shape used by the engine once, on a 64-lane wave, so a test can check
// the assumed operand/accumulator lane layouts against a host matmul.
//   a, b: 32x16 (row-major) operands for the 16-bit shapes; the first 32x2 / 16x4 part
//   (row-major, k fastest) is used for the f32/f64 shapes.
__global__ void mfma_probe_kernel(const float* a, const float* b, float* o_bf16_32, float* o_f32_32,
                                  float* o_f32_16, double* o_f64_16, float* o_bf16_16) {
  const int l = threadIdx.x;
  // 32x32x16 bf16: A[row=l%32][k=8*(l/32)+i], B[k=8*(l/32)+i][col=l%32]; b is stored as
  // B^T (col-major B) row-major 32x16, i.e. b[col*16 + k].
  {
    u16x8 av, bv;
    for (int i = 0; i < 8; ++i) {
      av[i] = f2bf_bits(a[(l & 31) * 16 + 8 * (l >> 5) + i]);
      bv[i] = f2bf_bits(b[(l & 31) * 16 + 8 * (l >> 5) + i]);
    }
    f32x16 acc = {};
    acc = mfma32x32x16<PREC_BF16>(av, bv, acc);
    for (int e = 0; e < 16; ++e) o_bf16_32[acc32_row(e, l) * 32 + (l & 31)] = acc[e];
  }
  // 32x32x2 f32: A[row=l%32][k=l/32], B[k=l/32][col=l%32].
  {
    float av = a[(l & 31) * 16 + (l >> 5)];
    float bv = b[(l & 31) * 16 + (l >> 5)];
    f32x16 acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
    for (int e = 0; e < 16; ++e) o_f32_32[acc32_row(e, l) * 32 + (l & 31)] = acc[e];
  }
  // 16x16x4 f32: A[row=l%16][k=l/16], B[k=l/16][col=l%16].
  {
    float av = a[(l & 15) * 16 + (l >> 4)];
    float bv = b[(l & 15) * 16 + (l >> 4)];
    f32x4 acc = {};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) o_f32_16[acc16_row(e, l) * 16 + (l & 15)] = acc[e];
  }
  // 16x16x4 f64: same operand layout, accumulator rows 4 e + l / 16.
  {
    double av = a[(l & 15) * 16 + (l >> 4)];
    double bv = b[(l & 15) * 16 + (l >> 4)];
    f64x4 acc = {};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) o_f64_16[acc16_row_f64(e, l) * 16 + (l & 15)] = acc[e];
  }
  // 16x16x32 bf16: A[row=l%16][k=8*(l/16)+i], B[k=8*(l/16)+i][col=l%16]; a, b as 16x32.
  {
    u16x8 av, bv;
    for (int i = 0; i < 8; ++i) {
      av[i] = f2bf_bits(a[(l & 15) * 32 + 8 * (l >> 4) + i]);
      bv[i] = f2bf_bits(b[(l & 15) * 32 + 8 * (l >> 4) + i]);
    }
    f32x4 acc = {};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av),
                                                  __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) o_bf16_16[acc16_row(e, l) * 16 + (l & 15)] = acc[e];
  }
}

int mfma_probe_launch(hipStream_t st, const float* a, const float* b, float* o1, float* o2, float* o3,
                      double* o4, float* o5) {
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, st, a, b, o1, o2, o3, o4, o5);
  return (int)hipGetLastError() == 0 ? 0 : 5;
}

// ------------------------------------------------------------------ gram_pms
// Tile: 128 rows of the matricisation x 64 columns per chunk; 256 threads = 4 waves,
// wave w owns the 64x64 quadrant (w>>1, w&1) of the 128x128 output tile pair.
constexpr int GT_ROWS = 128;
constexpr int GT_K = 64;

enum Loader { LD_ROWS_CONTIG = 0, LD_COLS_CONTIG = 1, LD_GENERIC = 2 };

template <int PREC>
struct GramLds;
template <>
struct GramLds<PREC_F32> {
  static constexpr int LD = GT_K + 4;  // floats per row
  typedef float T;
};
template <>
struct GramLds<PREC_BF16> {
  static constexpr int LD = GT_K + 8;  // 16-bit elements per row (row = 144 bytes)
  typedef uint16_t T;
};
template <>
struct GramLds<PREC_F16> {
  static constexpr int LD = GT_K + 8;
  typedef uint16_t T;
};

// Global -> registers for one 128 x 64 tile (rows m0.., columns col0..), 32 floats.
template <int LOADER>
__device__ __forceinline__ void gram_load(const float* __restrict__ x, PMS v, long K, int m0, long col0,
                                          int tid, float (&r)[32]) {
  if constexpr (LOADER == LD_COLS_CONTIG) {
    // P % 64 == 0: the 64 columns share one s; each row is 64 contiguous floats.
    const long s = col0 / v.P;
    const long p0 = col0 - s * v.P;
    const float* base = x + p0 + v.P * v.M * s;
    const int kq = tid & 15;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int row = (tid >> 4) + 16 * i;
      const long m = m0 + row;
      float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
      if (m < v.M) q = *reinterpret_cast<const float4*>(base + v.P * m + 4 * kq);
      r[4 * i + 0] = q.x; r[4 * i + 1] = q.y; r[4 * i + 2] = q.z; r[4 * i + 3] = q.w;
    }
  } else if constexpr (LOADER == LD_ROWS_CONTIG) {
    // P == 1, M % 4 == 0: column c is M contiguous floats; 4 rows per float4.
    const int rq = tid & 31;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = (tid >> 5) + 8 * i;
      const long c = col0 + k;
      const long m = m0 + 4 * rq;
      float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
      if (m < v.M && c < K) q = *reinterpret_cast<const float4*>(x + m + v.M * c);
      r[4 * i + 0] = q.x; r[4 * i + 1] = q.y; r[4 * i + 2] = q.z; r[4 * i + 3] = q.w;
    }
  } else {
#pragma unroll 4
    for (int i = 0; i < 32; ++i) {
      const int e = tid + 256 * i;
      const int row = e >> 6, k = e & 63;
      const long m = m0 + row, c = col0 + k;
      float val = 0.f;
      if (m < v.M && c < K) {
        const long s = c / v.P, p = c - s * v.P;
        val = x[p + v.P * m + v.P * v.M * s];
      }
      r[i] = val;
    }
  }
}

// Registers -> LDS tile [row][k].
template <int PREC, int LOADER>
__device__ __forceinline__ void gram_store(typename GramLds<PREC>::T* __restrict__ T, int tid,
                                           const float (&r)[32]) {
  constexpr int LD = GramLds<PREC>::LD;
  if constexpr (LOADER == LD_COLS_CONTIG) {
    const int kq = tid & 15;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int row = (tid >> 4) + 16 * i;
      if constexpr (PREC == PREC_F32) {
        *reinterpret_cast<float4*>(&T[row * LD + 4 * kq]) =
            make_float4(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]);
      } else {
        uint2 w;
        w.x = (uint32_t)to16<PREC>(r[4 * i]) | ((uint32_t)to16<PREC>(r[4 * i + 1]) << 16);
        w.y = (uint32_t)to16<PREC>(r[4 * i + 2]) | ((uint32_t)to16<PREC>(r[4 * i + 3]) << 16);
        *reinterpret_cast<uint2*>(&T[row * LD + 4 * kq]) = w;
      }
    }
  } else if constexpr (LOADER == LD_ROWS_CONTIG) {
    const int rq = tid & 31;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = (tid >> 5) + 8 * i;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if constexpr (PREC == PREC_F32) T[(4 * rq + j) * LD + k] = r[4 * i + j];
        else T[(4 * rq + j) * LD + k] = to16<PREC>(r[4 * i + j]);
      }
    }
  } else {
#pragma unroll 4
    for (int i = 0; i < 32; ++i) {
      const int e = tid + 256 * i;
      const int row = e >> 6, k = e & 63;
      if constexpr (PREC == PREC_F32) T[row * LD + k] = r[i];
      else T[row * LD + k] = to16<PREC>(r[i]);
    }
  }
}

template <int PREC>
__device__ __forceinline__ void gram_compute(const typename GramLds<PREC>::T* __restrict__ Ai,
                                             const typename GramLds<PREC>::T* __restrict__ Aj, int wr,
                                             int wc, int lane, f32x16 (&acc)[2][2]) {
  constexpr int LD = GramLds<PREC>::LD;
  if constexpr (PREC == PREC_F32) {
#pragma unroll 4
    for (int kk = 0; kk < GT_K / 2; ++kk) {
      float af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = Ai[(64 * wr + 32 * a + (lane & 31)) * LD + 2 * kk + (lane >> 5)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Aj[(64 * wc + 32 * b + (lane & 31)) * LD + 2 * kk + (lane >> 5)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < GT_K / 16; ++kk) {
      u16x8 af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
        af[a] = *reinterpret_cast<const u16x8*>(&Ai[(64 * wr + 32 * a + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int b = 0; b < 2; ++b)
        bf[b] = *reinterpret_cast<const u16x8*>(&Aj[(64 * wc + 32 * b + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mfma32x32x16<PREC>(af[a], bf[b], acc[a][b]);
    }
  }
}

template <int PREC, int LOADER>
__global__ __launch_bounds__(256) void gram_pms_kernel(const float* __restrict__ x, PMS v, long K,
                                                       long n_chunks, int n_split, int n_tiles,
                                                       float* __restrict__ part, double* __restrict__ norms) {
  typedef typename GramLds<PREC>::T T;
  constexpr int LD = GramLds<PREC>::LD;
  __shared__ __attribute__((aligned(16))) T Ai[GT_ROWS * LD];
  __shared__ __attribute__((aligned(16))) T Aj[GT_ROWS * LD];
  __shared__ double red[4];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1;
  const int split = blockIdx.x, pair = blockIdx.y;
  int bi = 0, bj = 0;
  {
    int p = pair;
    for (bi = 0; bi < n_tiles; ++bi) {
      const int cnt = n_tiles - bi;
      if (p < cnt) { bj = bi + p; break; }
      p -= cnt;
    }
  }
  const bool diag = (bi == bj);
  const long c_begin = (long)split * n_chunks / n_split;
  const long c_end = (long)(split + 1) * n_chunks / n_split;

  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x16{};

  float ri[32], rj[32];
  double sumsq = 0.0;
  if (c_begin < c_end) {
    gram_load<LOADER>(x, v, K, GT_ROWS * bi, c_begin * GT_K, tid, ri);
    if (!diag) gram_load<LOADER>(x, v, K, GT_ROWS * bj, c_begin * GT_K, tid, rj);
  }
  for (long c = c_begin; c < c_end; ++c) {
    __syncthreads();
    gram_store<PREC, LOADER>(Ai, tid, ri);
    if (!diag) gram_store<PREC, LOADER>(Aj, tid, rj);
    if (diag) {
      float s = 0.f;
#pragma unroll
      for (int i = 0; i < 32; ++i) s += ri[i] * ri[i];
      sumsq += (double)s;
    }
    __syncthreads();
    if (c + 1 < c_end) {
      gram_load<LOADER>(x, v, K, GT_ROWS * bi, (c + 1) * GT_K, tid, ri);
      if (!diag) gram_load<LOADER>(x, v, K, GT_ROWS * bj, (c + 1) * GT_K, tid, rj);
    }
    gram_compute<PREC>(Ai, diag ? Ai : Aj, wr, wc, lane, acc);
  }

  float* out = part + ((long)pair * n_split + split) * (GT_ROWS * GT_ROWS);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = 64 * wr + 32 * a + acc32_row(e, lane);
        const int col = 64 * wc + 32 * b + (lane & 31);
        out[row * GT_ROWS + col] = acc[a][b][e];
      }
  sumsq = wave_sum(sumsq);
  if (lane == 0) red[w] = sumsq;
  __syncthreads();
  if (tid == 0) norms[(long)pair * n_split + split] = red[0] + red[1] + red[2] + red[3];
}


// ------------------------------------------------------------------ gram_big
// Matricisations of more than 256 rows with 16-bit products. A block owns a 256 x 256
// tile pair (i, j) of G over one range of columns: 512 threads = 8 waves in a 4 x 2 grid,
// wave (wr, wc) multiplying rows 64 wr.. of tile i against rows 128 wc.. of tile j
// (2 x 4 MFMA 32x32x16 accumulators). Columns advance in chunks of 64 through two LDS
// buffers: chunk c + 1 is loaded into registers while chunk c is multiplied, then stored
// to the other buffer, so one barrier per chunk orders both. Against 128 x 128 tiles each
// column chunk is read half as many times (4 tile pairs per row tile of a 1024-row
// matricisation instead of 8); the bytes each CU pulls from L2 per MFMA are what bound
// the smaller tiles at these sizes. Diagonal pairs (i == j) load one tile and run their
// own copy of the loop (no branch on i == j inside it).
// Tile b starts at row gb_origin(b) = min(256 b, M - 256): the last tile overlaps its
// neighbour instead of running past row M, so every row a block reads exists (entries
// of the overlap are produced by two tile pairs, with the same value) and the loads need
// no row bounds. Their addresses are a block- or wave-uniform base plus a fixed 32-bit
// lane offset.
constexpr int GB_ROWS = 256, GB_K = 64, GB_LD = GB_K + 8;  // LDS rows of 144 B (9 x 16 B)
constexpr int GB_TILE = GB_ROWS * GB_LD;                    // one tile, 16-bit elements

__device__ __forceinline__ long gb_origin(int b, long M) { return min((long)GB_ROWS * b, M - GB_ROWS); }

// One 256 x 64 tile (rows m0.., columns col0..) into registers, 32 floats per thread.
template <int LOADER>
__device__ __forceinline__ void gb_load(const float* __restrict__ x, PMS v, long K, long m0, long col0, int tid,
                                        float (&r)[32]) {
  if constexpr (LOADER == LD_COLS_CONTIG) {
    // P % 64 == 0: the chunk's 64 columns share one s and are contiguous in every row (K
    // is a multiple of 64, so chunks are whole); 16 lanes read a row's 256 bytes (a
    // float4 each), rows tid / 16 + 32 i.
    const long s = col0 / v.P, p0 = col0 - s * v.P;
    const float* base = x + p0 + v.P * (v.M * s + m0);
    const unsigned off = (unsigned)v.P * (unsigned)(tid >> 4) + 4u * (unsigned)(tid & 15);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const float4 q = *reinterpret_cast<const float4*>(base + v.P * (32 * i) + off);
      r[4 * i + 0] = q.x; r[4 * i + 1] = q.y; r[4 * i + 2] = q.z; r[4 * i + 3] = q.w;
    }
  } else {
    // P == 1. Thread owns row m0 + tid % 256 and columns col0 + 32 (tid / 256) + i, i < 32
    // (the column clamped to K - 1; gb_store zeroes the copies): a wave's load i reads
    // 256 contiguous bytes of one column.
    const long cb = col0 + 32 * __builtin_amdgcn_readfirstlane(tid >> 8);
    const unsigned off = (unsigned)(tid & 255);
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      const float* col = x + (v.M * min(cb + i, K - 1) + m0);
      r[i] = col[off];
    }
  }
}

// Registers -> LDS tile [row][k] as 16-bit values (row stride GB_LD). Columns past K are
// zeroed in r as well (the diagonal pairs' norm sums r after the store).
template <int PREC, int LOADER>
__device__ __forceinline__ void gb_store(uint16_t* __restrict__ T, int tid, float (&r)[32], long K, long col0) {
  if constexpr (LOADER == LD_COLS_CONTIG) {
    const int kq = tid & 15;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int row = (tid >> 4) + 32 * i;
      uint2 w;
      w.x = cvt2<PREC>(r[4 * i], r[4 * i + 1]);
      w.y = cvt2<PREC>(r[4 * i + 2], r[4 * i + 3]);
      *reinterpret_cast<uint2*>(&T[row * GB_LD + 4 * kq]) = w;
    }
  } else {
    const int row = tid & 255, k0 = 32 * (tid >> 8);
    const long nv = K - (col0 + k0);  // columns of this thread inside the matrix
#pragma unroll
    for (int i = 0; i < 32; ++i)
      if (i >= nv) r[i] = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint4 w;
      w.x = cvt2<PREC>(r[8 * q + 0], r[8 * q + 1]);
      w.y = cvt2<PREC>(r[8 * q + 2], r[8 * q + 3]);
      w.z = cvt2<PREC>(r[8 * q + 4], r[8 * q + 5]);
      w.w = cvt2<PREC>(r[8 * q + 6], r[8 * q + 7]);
      *reinterpret_cast<uint4*>(&T[row * GB_LD + k0 + 8 * q]) = w;
    }
  }
}

// Sum of squares of the rows of r that tile b owns (rows from 256 b on: the overlapped
// rows of the last tile belong to its neighbour), so the diagonal pairs' norms add up to
// ||X||^2.
template <int LOADER>
__device__ __forceinline__ float gb_sumsq(const float (&r)[32], int tid, int own_lo) {
  float s = 0.f;
  if constexpr (LOADER == LD_COLS_CONTIG) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const float t = r[4 * i] * r[4 * i] + r[4 * i + 1] * r[4 * i + 1] + r[4 * i + 2] * r[4 * i + 2] +
                      r[4 * i + 3] * r[4 * i + 3];
      s += (tid >> 4) + 32 * i >= own_lo ? t : 0.f;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 32; ++i) s += r[i] * r[i];
    if ((tid & 255) < own_lo) s = 0.f;
  }
  return s;
}

template <int PREC, int LOADER, bool DIAG>
__device__ __forceinline__ void gb_body(const float* __restrict__ x, PMS v, long K, long c_begin, long c_end,
                                        long mi, long mj, int own_lo, uint16_t* __restrict__ lds,
                                        f32x16 (&acc)[2][4], double& sumsq) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1;
  constexpr int BUF = DIAG ? GB_TILE : 2 * GB_TILE;
  float ri[32], rj[32];
  gb_load<LOADER>(x, v, K, mi, c_begin * GB_K, tid, ri);
  if constexpr (!DIAG) gb_load<LOADER>(x, v, K, mj, c_begin * GB_K, tid, rj);
  gb_store<PREC, LOADER>(lds, tid, ri, K, c_begin * GB_K);
  if constexpr (!DIAG) gb_store<PREC, LOADER>(lds + GB_TILE, tid, rj, K, c_begin * GB_K);
  __syncthreads();
  int cur = 0;
  for (long c = c_begin; c < c_end; ++c) {
    if constexpr (DIAG) sumsq += (double)gb_sumsq<LOADER>(ri, tid, own_lo);  // ri = chunk c
    // the next chunk (the last one again at the end: loaded and stored, never used)
    const long cn = c + 1 < c_end ? c + 1 : c;
    gb_load<LOADER>(x, v, K, mi, cn * GB_K, tid, ri);
    if constexpr (!DIAG) gb_load<LOADER>(x, v, K, mj, cn * GB_K, tid, rj);
    // keep the loads ahead of the multiply (the scheduler would otherwise sink them to
    // their stores and wait on them inside the MFMA sequence)
    __builtin_amdgcn_sched_barrier(0);
    const uint16_t* Ai = lds + cur * BUF;
    const uint16_t* Aj = DIAG ? Ai : Ai + GB_TILE;
#pragma unroll
    for (int kk = 0; kk < GB_K / 16; ++kk) {
      u16x8 af[2], bf[4];
#pragma unroll
      for (int a = 0; a < 2; ++a)
        af[a] = *reinterpret_cast<const u16x8*>(&Ai[(64 * wr + 32 * a + (lane & 31)) * GB_LD + 16 * kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int b = 0; b < 4; ++b)
        bf[b] = *reinterpret_cast<const u16x8*>(&Aj[(128 * wc + 32 * b + (lane & 31)) * GB_LD + 16 * kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = mfma32x32x16<PREC>(af[a], bf[b], acc[a][b]);
    }
    __builtin_amdgcn_sched_barrier(0);
    cur ^= 1;
    gb_store<PREC, LOADER>(lds + cur * BUF, tid, ri, K, cn * GB_K);
    if constexpr (!DIAG) gb_store<PREC, LOADER>(lds + cur * BUF + GB_TILE, tid, rj, K, cn * GB_K);
    __syncthreads();
  }
}

// 16-bit copies of X (the values the FP32 loaders above convert to when they store to
// LDS): a chunk is half the bytes, so the body keeps two chunks in flight in the same
// registers. One 256 x 64 tile, 16 uint32 per thread.
template <int LOADER>
__device__ __forceinline__ void gb_load16(const uint16_t* __restrict__ x, PMS v, long K, long m0, long col0,
                                          int tid, uint32_t (&r)[16]) {
  if constexpr (LOADER == LD_COLS_CONTIG) {
    // 8 lanes read a row's 128 bytes (16 each), rows tid / 8 + 64 i
    const long s = col0 / v.P, p0 = col0 - s * v.P;
    const uint16_t* base = x + p0 + v.P * (v.M * s + m0);
    const unsigned off = (unsigned)v.P * (unsigned)(tid >> 3) + 8u * (unsigned)(tid & 7);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint4 q = *reinterpret_cast<const uint4*>(base + v.P * (64 * i) + off);
      r[4 * i + 0] = q.x; r[4 * i + 1] = q.y; r[4 * i + 2] = q.z; r[4 * i + 3] = q.w;
    }
  } else {
    // P == 1, M % 4 == 0. Thread owns rows m0 + 4 (tid % 64) .. + 3 and columns
    // col0 + 8 (tid / 64) + i, i < 8 (clamped to K - 1; gb_store16 zeroes the copies): a
    // wave's load i reads the 512 contiguous bytes of one column.
    const long cb = col0 + 8 * __builtin_amdgcn_readfirstlane(tid >> 6);
    const unsigned off = 4u * (unsigned)(tid & 63);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint2 q = *reinterpret_cast<const uint2*>(x + (v.M * min(cb + i, K - 1) + m0) + off);
      r[2 * i] = q.x; r[2 * i + 1] = q.y;
    }
  }
}

// Registers -> LDS tile [row][k] (row stride GB_LD); columns past K, and every column
// when keep == 0 (a chunk past the block's range), stored as zeros.
template <int LOADER>
__device__ __forceinline__ void gb_store16(uint16_t* __restrict__ T, int tid, const uint32_t (&r)[16], long K,
                                           long col0, uint32_t keep) {
  if constexpr (LOADER == LD_COLS_CONTIG) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      *reinterpret_cast<uint4*>(&T[((tid >> 3) + 64 * i) * GB_LD + 8 * (tid & 7)]) =
          make_uint4(r[4 * i] & keep, r[4 * i + 1] & keep, r[4 * i + 2] & keep, r[4 * i + 3] & keep);
  } else {
    // r[2 i] = rows 0, 1 of column i (low half row 0), r[2 i + 1] = rows 2, 3: transpose
    // to 8 consecutive columns of each of the 4 rows
    const int row = 4 * (tid & 63), k0 = 8 * (tid >> 6);
    const long nv = K - (col0 + k0);  // columns of this thread inside the matrix
    uint32_t c[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) c[i] = (i >> 1) < nv ? (r[i] & keep) : 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int h = q >> 1;            // word of the column: rows 0-1 or 2-3
      const bool hi = (q & 1) != 0;    // upper half of that word
      uint32_t w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t a = c[2 * (2 * j) + h], b = c[2 * (2 * j + 1) + h];
        w[j] = hi ? ((a >> 16) | (b & 0xffff0000u)) : ((a & 0xffffu) | (b << 16));
      }
      *reinterpret_cast<uint4*>(&T[(row + q) * GB_LD + k0]) = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
}

template <int PREC, bool DIAG>
__device__ __forceinline__ void gb_mma(const uint16_t* __restrict__ Ai, const uint16_t* __restrict__ Aj, int wr,
                                       int wc, int lane, f32x16 (&acc)[2][4]) {
#pragma unroll
  for (int kk = 0; kk < GB_K / 16; ++kk) {
    u16x8 af[2], bf[4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
      af[a] = *reinterpret_cast<const u16x8*>(&Ai[(64 * wr + 32 * a + (lane & 31)) * GB_LD + 16 * kk + 8 * (lane >> 5)]);
#pragma unroll
    for (int b = 0; b < 4; ++b)
      bf[b] = *reinterpret_cast<const u16x8*>(&Aj[(128 * wc + 32 * b + (lane & 31)) * GB_LD + 16 * kk + 8 * (lane >> 5)]);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = mfma32x32x16<PREC>(af[a], bf[b], acc[a][b]);
  }
}

// Two register stages: while chunk c is multiplied from one LDS buffer, chunk c + 1
// (loaded one step earlier) is stored to the other and chunk c + 2 is being loaded. The
// loop runs over pairs of chunks with no exit in between (an exit between the halves lets
// the compiler move the first half's loads into the second and spill); an odd count ends
// with one chunk of zeros. Loads past the range reread the last chunk.
template <int PREC, int LOADER, bool DIAG>
__device__ __forceinline__ void gb_body16(const uint16_t* __restrict__ x, PMS v, long K, long c_begin, long c_end,
                                          long mi, long mj, uint16_t* __restrict__ lds, f32x16 (&acc)[2][4]) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1;
  constexpr int BUF = DIAG ? GB_TILE : 2 * GB_TILE;
  constexpr int OFF_J = DIAG ? 0 : GB_TILE;
  uint32_t ri0[16], rj0[16], ri1[16], rj1[16];
  const long last = c_end - 1;
  const long c_stop = c_begin + ((c_end - c_begin + 1) & ~1L);
  auto keep = [&](long c) { return c < c_end ? 0xffffffffu : 0u; };
  gb_load16<LOADER>(x, v, K, mi, c_begin * GB_K, tid, ri0);
  if constexpr (!DIAG) gb_load16<LOADER>(x, v, K, mj, c_begin * GB_K, tid, rj0);
  gb_load16<LOADER>(x, v, K, mi, min(c_begin + 1, last) * GB_K, tid, ri1);
  if constexpr (!DIAG) gb_load16<LOADER>(x, v, K, mj, min(c_begin + 1, last) * GB_K, tid, rj1);
  gb_store16<LOADER>(lds, tid, ri0, K, c_begin * GB_K, ~0u);
  if constexpr (!DIAG) gb_store16<LOADER>(lds + GB_TILE, tid, rj0, K, c_begin * GB_K, ~0u);
  __syncthreads();
  for (long c = c_begin; c < c_stop; c += 2) {
    // buffer 0 holds chunk c, stage 1 chunk c + 1
    gb_load16<LOADER>(x, v, K, mi, min(c + 2, last) * GB_K, tid, ri0);
    if constexpr (!DIAG) gb_load16<LOADER>(x, v, K, mj, min(c + 2, last) * GB_K, tid, rj0);
    __builtin_amdgcn_sched_barrier(0);
    gb_mma<PREC, DIAG>(lds, lds + OFF_J, wr, wc, lane, acc);
    __builtin_amdgcn_sched_barrier(0);
    gb_store16<LOADER>(lds + BUF, tid, ri1, K, min(c + 1, last) * GB_K, keep(c + 1));
    if constexpr (!DIAG) gb_store16<LOADER>(lds + BUF + GB_TILE, tid, rj1, K, min(c + 1, last) * GB_K, keep(c + 1));
    __syncthreads();
    // buffer 1 holds chunk c + 1, stage 0 chunk c + 2
    gb_load16<LOADER>(x, v, K, mi, min(c + 3, last) * GB_K, tid, ri1);
    if constexpr (!DIAG) gb_load16<LOADER>(x, v, K, mj, min(c + 3, last) * GB_K, tid, rj1);
    __builtin_amdgcn_sched_barrier(0);
    gb_mma<PREC, DIAG>(lds + BUF, lds + BUF + OFF_J, wr, wc, lane, acc);
    __builtin_amdgcn_sched_barrier(0);
    gb_store16<LOADER>(lds, tid, ri0, K, min(c + 2, last) * GB_K, keep(c + 2));
    if constexpr (!DIAG) gb_store16<LOADER>(lds + GB_TILE, tid, rj0, K, min(c + 2, last) * GB_K, keep(c + 2));
    __syncthreads();
  }
}

// T = float: FP32 X converted as it is stored to LDS, norms = the diagonal pairs' sums of
// squares. T = uint16_t: a 16-bit copy of X, no norms (the copy's pass takes them).
template <int PREC, int LOADER, typename T>
__global__ __launch_bounds__(512) void gram_big_kernel(const T* __restrict__ x, PMS v, long K, long n_chunks,
                                                       int n_split, int n_tiles, float* __restrict__ part,
                                                       double* __restrict__ norms) {
  __shared__ __attribute__((aligned(16))) uint16_t lds[4 * GB_TILE];
  __shared__ double red[8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1;
  // Work-groups are dealt to the 8 XCDs in turn (linear id mod 8). With n_split a
  // multiple of 8, every tile pair of a column split runs on the same XCD, so a row
  // panel's chunk, read by the n_tiles pairs that contain its tile, comes from HBM once
  // into that XCD's L2 rather than once per pair.
  const int n_pairs = n_tiles * (n_tiles + 1) / 2;
  int split, pair;
  if (n_split % 8 == 0) {
    const int q = blockIdx.x >> 3;
    split = (blockIdx.x & 7) + 8 * (q / n_pairs);
    pair = q % n_pairs;
  } else {
    split = blockIdx.x % n_split;
    pair = blockIdx.x / n_split;
  }
  int bi = 0, bj = 0;
  {
    int p = pair;
    for (bi = 0; bi < n_tiles; ++bi) {
      const int cnt = n_tiles - bi;
      if (p < cnt) { bj = bi + p; break; }
      p -= cnt;
    }
  }
  const long c_begin = (long)split * n_chunks / n_split;
  const long c_end = (long)(split + 1) * n_chunks / n_split;
  const long mi = gb_origin(bi, v.M), mj = gb_origin(bj, v.M);
  f32x16 acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x16{};
  double sumsq = 0.0;
  if (c_begin < c_end) {
    if constexpr (sizeof(T) == 2) {
      if (bi == bj) gb_body16<PREC, LOADER, true>(x, v, K, c_begin, c_end, mi, mi, lds, acc);
      else gb_body16<PREC, LOADER, false>(x, v, K, c_begin, c_end, mi, mj, lds, acc);
    } else {
      if (bi == bj)
        gb_body<PREC, LOADER, true>(x, v, K, c_begin, c_end, mi, mi, (int)((long)GB_ROWS * bi - mi), lds, acc,
                                    sumsq);
      else
        gb_body<PREC, LOADER, false>(x, v, K, c_begin, c_end, mi, mj, 0, lds, acc, sumsq);
    }
  }
  float* out = part + ((long)pair * n_split + split) * (GB_ROWS * GB_ROWS);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = 64 * wr + 32 * a + acc32_row(e, lane);
        const int col = 128 * wc + 32 * b + (lane & 31);
        out[row * GB_ROWS + col] = acc[a][b][e];
      }
  if constexpr (sizeof(T) == 4) {
    sumsq = wave_sum(sumsq);
    if (lane == 0) red[w] = sumsq;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int q = 0; q < 8; ++q) t += red[q];
      norms[(long)pair * n_split + split] = t;
    }
  }
}

// Tile edge of the Gram partials: 256 for 16-bit products of more than 256 rows whose
// columns are single floats of a row's stretch (P == 1) or aligned runs of 64
// (P % 64 == 0, 16-byte aligned data, lane offsets P * 15 + 60 within 32 bits).
int gram_pms_tile(const float* x, PMS v, int prec) {
  const bool rows = v.P == 1;
  const bool cols = v.P % 64 == 0 && ((uintptr_t)x & 15) == 0 && v.P <= (1L << 24);
  return (prec != PREC_F32 && v.M > GB_ROWS && (rows || cols)) ? GB_ROWS : GT_ROWS;
}

int gram_pms_pairs(const float* x, PMS v, int prec) {
  const int tile = gram_pms_tile(x, v, prec);
  const int nt = (int)((v.M + tile - 1) / tile);
  return nt * (nt + 1) / 2;
}

// Column splits: 256-row tiles take at most one block per CU (256 blocks, LDS-limited to
// one per CU), a multiple of 8 splits where there are that many (the XCD placement of
// gram_big_kernel); 128-row tiles two blocks per CU.
int gram_pms_nsplit(const float* x, PMS v, int prec) {
  const int tile = gram_pms_tile(x, v, prec);
  const int kc = tile == GB_ROWS ? GB_K : GT_K;
  const long n_chunks = (v.P * v.S + kc - 1) / kc;
  long ns = (tile == GB_ROWS ? 256 : 512) / gram_pms_pairs(x, v, prec);
  if (tile == GB_ROWS && ns >= 8) ns -= ns % 8;
  if (ns < 1) ns = 1;
  if (ns > n_chunks) ns = n_chunks;
  return (int)ns;
}

template <int PREC>
static void gram_dispatch(hipStream_t st, const float* x, PMS v, int n_split, float* part, double* norms) {
  const long K = v.P * v.S;
  const bool aligned = ((uintptr_t)x & 15) == 0;
  if constexpr (PREC != PREC_F32) {
    if (gram_pms_tile(x, v, PREC) == GB_ROWS) {
      const long n_chunks = (K + GB_K - 1) / GB_K;
      const int nt = (int)((v.M + GB_ROWS - 1) / GB_ROWS);
      dim3 grid(n_split * (nt * (nt + 1) / 2));
      if (v.P == 1)
        hipLaunchKernelGGL((gram_big_kernel<PREC, LD_ROWS_CONTIG, float>), grid, dim3(512), 0, st, x, v, K, n_chunks,
                           n_split, nt, part, norms);
      else
        hipLaunchKernelGGL((gram_big_kernel<PREC, LD_COLS_CONTIG, float>), grid, dim3(512), 0, st, x, v, K, n_chunks,
                           n_split, nt, part, norms);
      return;
    }
  }
  const long n_chunks = (K + GT_K - 1) / GT_K;
  const int nt = (int)((v.M + GT_ROWS - 1) / GT_ROWS);
  const int pairs = nt * (nt + 1) / 2;
  dim3 grid(n_split, pairs);
  if (aligned && v.P % 64 == 0)
    hipLaunchKernelGGL((gram_pms_kernel<PREC, LD_COLS_CONTIG>), grid, dim3(256), 0, st, x, v, K, n_chunks,
                       n_split, nt, part, norms);
  else if (aligned && v.P == 1 && v.M % 4 == 0)
    hipLaunchKernelGGL((gram_pms_kernel<PREC, LD_ROWS_CONTIG>), grid, dim3(256), 0, st, x, v, K, n_chunks,
                       n_split, nt, part, norms);
  else
    hipLaunchKernelGGL((gram_pms_kernel<PREC, LD_GENERIC>), grid, dim3(256), 0, st, x, v, K, n_chunks, n_split,
                       nt, part, norms);
}

int gram_pms_launch(hipStream_t st, const float* x, PMS v, int prec, int n_split, float* part, double* norms) {
  if (v.M <= 0 || v.M > 1024 || v.P <= 0 || v.S <= 0 || n_split < 1) return 8;
  if (prec == PREC_F32) gram_dispatch<PREC_F32>(st, x, v, n_split, part, norms);
  else if (prec == PREC_F16) gram_dispatch<PREC_F16>(st, x, v, n_split, part, norms);
  else gram_dispatch<PREC_BF16>(st, x, v, n_split, part, norms);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// Sum the split partials of every tile pair into the full symmetric FP64 Gram.
// Sums the n_split FP32 partial tiles of each tile pair into the FP64 Gram. A block takes
// 64 entries of one 128 x 128 tile (one per lane, consecutive in memory) and its 16 waves
// split the partials (wave w: w, w + 16, ..., eight loads in flight per batch); the
// waves' sums are added in a fixed order through LDS. Diagonal tiles produce every entry
// themselves, off-diagonal ones also write the mirrored entry.
// Up to two reductions in one launch (blockIdx.z picks the job); block (0, 0, 0) also sums
// the n_norms per-block norms into *norm_out when norms != nullptr.
struct GramReduceJob {
  const float* part;
  int M, n_split;
  double* G;
  int tile;  // 128 or 256
};
__global__ __launch_bounds__(1024) void gram_reduce_kernel(GramReduceJob j0, GramReduceJob j1,
                                                           const double* __restrict__ norms, int n_norms,
                                                           double* __restrict__ norm_out) {
  __shared__ double red[16][64];
  const GramReduceJob job = blockIdx.z ? j1 : j0;
  const float* __restrict__ part = job.part;
  const int M = job.M, tile = job.tile, nt = (M + tile - 1) / tile, n_split = job.n_split;
  double* __restrict__ G = job.G;
  const int pair = blockIdx.y;
  if (pair >= nt * (nt + 1) / 2) return;  // the smaller job of a launch with two
  if (norms && blockIdx.x == 0 && pair == 0 && blockIdx.z == 0 && threadIdx.x < 64) {
    double t = 0.0;
    for (int i = threadIdx.x; i < n_norms; i += 64) t += norms[i];
    t = wave_sum(t);
    if (threadIdx.x == 0) *norm_out = t;
  }
  int bi = 0, bj = 0;
  {
    int p = pair;
    for (bi = 0; bi < nt; ++bi) {
      const int cnt = nt - bi;
      if (p < cnt) { bj = bi + p; break; }
      p -= cnt;
    }
  }
  const int T2 = tile * tile;
  if (blockIdx.x * 64 >= T2) return;  // the smaller tile of a launch with two
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + lane;
  const float* p = part + (long)pair * n_split * T2 + e;
  double s = 0.0;
  constexpr int UB = 8;
  for (int k0 = w; k0 < n_split; k0 += 16 * UB) {
    float v[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      const int k = k0 + 16 * u;
      v[u] = p[(long)(k < n_split ? k : n_split - 1) * T2];
    }
#pragma unroll
    for (int u = 0; u < UB; ++u)
      if (k0 + 16 * u < n_split) s += (double)v[u];
  }
  red[w][lane] = s;
  __syncthreads();
  if (w == 0) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) t += red[q][lane];
    // 256-row tiles: the last one starts at M - 256 (see gram_big_kernel)
    const int r = e / tile, c = e % tile;
    const int oi = tile == GT_ROWS ? tile * bi : std::min(tile * bi, M - tile);
    const int oj = tile == GT_ROWS ? tile * bj : std::min(tile * bj, M - tile);
    const int i = oi + r, j = oj + c;
    if (i < M && j < M) {
      G[i + (long)M * j] = t;
      if (bi != bj) G[j + (long)M * i] = t;
    }
  }
}

int gram_reduce_launch(hipStream_t st, const float* part, int M, int n_split, double* G, int tile) {
  const int nt = (M + tile - 1) / tile;
  const int pairs = nt * (nt + 1) / 2;
  dim3 grid(tile * tile / 64, pairs);
  const GramReduceJob j{part, M, n_split, G, tile};
  hipLaunchKernelGGL(gram_reduce_kernel, grid, dim3(1024), 0, st, j, j, nullptr, 0, nullptr);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
int gram_reduce2_launch(hipStream_t st, const float* p1, int M1, int n1, double* G1, const float* p2, int M2, int n2,
                        double* G2, const double* norms, int n_norms, double* norm_out, int tile1, int tile2) {
  const int nt = std::max((M1 + tile1 - 1) / tile1, (M2 + tile2 - 1) / tile2);
  const int tile = std::max(tile1, tile2);
  dim3 grid(tile * tile / 64, nt * (nt + 1) / 2, 2);
  hipLaunchKernelGGL(gram_reduce_kernel, grid, dim3(1024), 0, st, GramReduceJob{p1, M1, n1, G1, tile1},
                     GramReduceJob{p2, M2, n2, G2, tile2}, norms, n_norms, norm_out);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

__global__ void sum_f64_kernel(const double* __restrict__ a, int n, double* __restrict__ out) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += a[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

int gram_reduce_norms_launch(hipStream_t st, const double* norms, int n, double* out) {
  hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, st, norms, n, out);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// ------------------------------------------------------------------ 16-bit copy of X
// Copy of X in 16-bit values (round to nearest even: the values the Gram and skinny GEMM
// kernels form when they read FP32 X in 16-bit precision) with ||X||^2 of the FP32 values
// (per block one FP64 partial, added by sum_f64_kernel). Each thread converts groups of 8
// floats (two 16-byte loads, one 16-byte store), four groups in flight.
constexpr int CAST16_BLOCKS = 2048;

template <int PREC>
__global__ __launch_bounds__(256) void cast16_sumsq_kernel(const float* __restrict__ x, long n,
                                                           uint16_t* __restrict__ y, double* __restrict__ part) {
  __shared__ double red[4];
  const long n8 = n / 8, stride = (long)gridDim.x * 256;
  double s = 0.0;
  auto one = [&](long g, float4 a, float4 b) {
    uint4 w;
    w.x = cvt2<PREC>(a.x, a.y);
    w.y = cvt2<PREC>(a.z, a.w);
    w.z = cvt2<PREC>(b.x, b.y);
    w.w = cvt2<PREC>(b.z, b.w);
    *reinterpret_cast<uint4*>(y + 8 * g) = w;
    s += (double)(a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w + b.x * b.x + b.y * b.y + b.z * b.z + b.w * b.w);
  };
  long g = blockIdx.x * 256L + threadIdx.x;
  for (; g + 3 * stride < n8; g += 4 * stride) {
    float4 a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = load_stream(x + 8 * (g + u * stride));
      b[u] = load_stream(x + 8 * (g + u * stride) + 4);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) one(g + u * stride, a[u], b[u]);
  }
  for (; g < n8; g += stride) one(g, load_stream(x + 8 * g), load_stream(x + 8 * g + 4));
  if (blockIdx.x == 0 && threadIdx.x < n - 8 * n8) {
    const long i = 8 * n8 + threadIdx.x;
    const float f = x[i];
    y[i] = to16<PREC>(f);
    s += (double)(f * f);
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

int cast16_part_size() { return CAST16_BLOCKS; }

int cast16_sumsq_launch(hipStream_t st, const float* x, long n, int prec, uint16_t* y, double* part,
                        double* sumsq) {
  if (n <= 0 || ((uintptr_t)x & 15) || ((uintptr_t)y & 15)) return 8;
  if (prec == PREC_F16)
    hipLaunchKernelGGL(cast16_sumsq_kernel<PREC_F16>, dim3(CAST16_BLOCKS), dim3(256), 0, st, x, n, y, part);
  else
    hipLaunchKernelGGL(cast16_sumsq_kernel<PREC_BF16>, dim3(CAST16_BLOCKS), dim3(256), 0, st, x, n, y, part);
  hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, st, part, CAST16_BLOCKS, sumsq);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// Grams from the 16-bit copy (16-byte aligned): gram_big_kernel's tiles over
// matricisations of 257 to 1024 rows whose columns are single values of a row's stretch
// (P == 1, M % 4 == 0) or runs of 64 (P % 64 == 0). part: [pairs][n_split][256 * 256] FP32 (gram_pms_pairs and
// gram_pms_nsplit of the same view in 16-bit precision).
bool gram16_ok(PMS v) {
  const bool rows = v.P == 1 && v.M % 4 == 0;
  const bool cols = v.P % 64 == 0 && v.P <= (1L << 24);
  return v.M > GB_ROWS && v.M <= 1024 && v.S > 0 && (rows || cols);
}

int gram16_pairs(PMS v) {
  const int nt = (int)((v.M + GB_ROWS - 1) / GB_ROWS);
  return nt * (nt + 1) / 2;
}

// as gram_pms_nsplit for 256-row tiles
int gram16_nsplit(PMS v) {
  const long n_chunks = (v.P * v.S + GB_K - 1) / GB_K;
  long ns = 256 / gram16_pairs(v);
  if (ns >= 8) ns -= ns % 8;
  if (ns < 1) ns = 1;
  if (ns > n_chunks) ns = n_chunks;
  return (int)ns;
}

int gram16_launch(hipStream_t st, const uint16_t* x, PMS v, int prec, int n_split, float* part) {
  if (!gram16_ok(v) || ((uintptr_t)x & 15) || n_split < 1 || (prec != PREC_BF16 && prec != PREC_F16)) return 8;
  const long K = v.P * v.S;
  const long n_chunks = (K + GB_K - 1) / GB_K;
  const int nt = (int)((v.M + GB_ROWS - 1) / GB_ROWS);
  dim3 grid(n_split * (nt * (nt + 1) / 2));
#define HT_G16(P, L) \
  hipLaunchKernelGGL((gram_big_kernel<P, L, uint16_t>), grid, dim3(512), 0, st, x, v, K, n_chunks, n_split, nt, part, nullptr)
  if (prec == PREC_F16) {
    if (v.P == 1) HT_G16(PREC_F16, LD_ROWS_CONTIG);
    else HT_G16(PREC_F16, LD_COLS_CONTIG);
  } else {
    if (v.P == 1) HT_G16(PREC_BF16, LD_ROWS_CONTIG);
    else HT_G16(PREC_BF16, LD_COLS_CONTIG);
  }
#undef HT_G16
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// ------------------------------------------------------------------ slab_gram
// 512 threads = 8 waves. Waves 0-3: quadrants of G1 = sum_s S S^T (A operand rows i,
// k = j, from the row-major copy); waves 4-7: quadrants of G2 = sum_s S^T S (rows j,
// k = i, from the column-major copy). Slabs are padded to 128 x 128 with zeros.
// The next slab's loads are issued unconditionally (clamped to valid addresses, the
// padding applied when the values are stored to LDS): a load behind a divergent branch
// makes the compiler wait for it at the branch join, which would serialise every slab's
// HBM latency with the multiply.
template <int PREC>
__global__ __launch_bounds__(512) void slab_gram_kernel(const float* __restrict__ x, int n1, int n2, long S,
                                                        int n_blocks, float* __restrict__ part1,
                                                        float* __restrict__ part2, double* __restrict__ norms) {
  // Scm[j][i] = S[i][j] (the slab's own column-major order, padded rows); Srm holds the
  // transpose, Srm[i][j ^ swz(i)] with swz(i) = 8 (((i / 4) ^ 4 i) mod 16): the XOR keeps
  // each aligned group of 8 columns contiguous (so MFMA operands stay single 16-byte
  // reads); it is a bijection both over 16 consecutive rows (the operand reads, lane =
  // row) and over rows 4m + e, m = 0..15 (the 4-element row stores), so neither access
  // has bank conflicts.
  // Two copies of (Srm, Scm): the slab being multiplied and the next one being stored.
  constexpr int LDC = 128 + 8;
  constexpr int BUF = 128 * 128 + 128 * LDC;
  __shared__ __attribute__((aligned(16))) uint16_t lds[2 * BUF];
  __shared__ double red[8];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int g = w >> 2, q = w & 3, wr = q >> 1, wc = q & 1;
  for (int i = tid; i < 2 * BUF; i += 512) lds[i] = 0;

  const long s_begin = (long)blockIdx.x * S / n_blocks;
  const long s_end = (long)(blockIdx.x + 1) * S / n_blocks;
  const long slab = (long)n1 * n2;

  // this thread's 4 x 4 blocks: b = tid + 512 u, rows i4 = 4 (b % 32), columns j4 = 4 (b / 32)
  float4 r[2][4];
  double sumsq = 0.0;
  auto load = [&](long s) {
    const float* base = x + s * slab;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int bb = tid + 512 * u, i4 = 4 * (bb & 31), j4 = 4 * (bb >> 5);
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const bool ok = i4 < n1 && j4 + d < n2;
        r[u][d] = load_stream(base + (ok ? (long)(j4 + d) * n1 + i4 : 0));
      }
    }
  };
  auto swz = [](int i) { return 8 * (((i >> 2) ^ (i << 2)) & 15); };
  auto store = [&](uint16_t* Srm, uint16_t* Scm) {
    float ss = 0.f;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int bb = tid + 512 * u, i4 = 4 * (bb & 31), j4 = 4 * (bb >> 5);
      float f[4][4];  // [d][e]: column j4 + d, row i4 + e (zero outside the slab)
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const bool ok = i4 < n1 && j4 + d < n2;
        f[d][0] = ok ? r[u][d].x : 0.f;
        f[d][1] = ok ? r[u][d].y : 0.f;
        f[d][2] = ok ? r[u][d].z : 0.f;
        f[d][3] = ok ? r[u][d].w : 0.f;
        ss += f[d][0] * f[d][0] + f[d][1] * f[d][1] + f[d][2] * f[d][2] + f[d][3] * f[d][3];
        uint2 wv;
        wv.x = cvt2<PREC>(f[d][0], f[d][1]);
        wv.y = cvt2<PREC>(f[d][2], f[d][3]);
        *reinterpret_cast<uint2*>(&Scm[(j4 + d) * LDC + i4]) = wv;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = i4 + e;
        uint2 wv;
        wv.x = cvt2<PREC>(f[0][e], f[1][e]);
        wv.y = cvt2<PREC>(f[2][e], f[3][e]);
        *reinterpret_cast<uint2*>(&Srm[i * 128 + (j4 ^ swz(i))]) = wv;
      }
    }
    sumsq += (double)ss;
  };

  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x16{};

  if (s_begin < s_end) {
    load(s_begin);
    __syncthreads();  // zero fill done
    store(lds, lds + 128 * 128);
    load(min(s_begin + 1, s_end - 1));
    __syncthreads();
  }
  for (long s = s_begin; s < s_end; ++s) {
    const uint16_t* Srm = lds + ((s - s_begin) & 1) * BUF;
    const uint16_t* Scm = Srm + 128 * 128;
    const int kmax = g == 0 ? n2 : n1;
    for (int kk = 0; kk < (kmax + 15) / 16; ++kk) {
      const int kc = 16 * kk + 8 * (lane >> 5);
      u16x8 af[2], bf[2];
      if (g == 0) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const int row = 64 * wr + 32 * a + (lane & 31);
          af[a] = *reinterpret_cast<const u16x8*>(&Srm[row * 128 + (kc ^ swz(row))]);
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const int row = 64 * wc + 32 * b + (lane & 31);
          bf[b] = *reinterpret_cast<const u16x8*>(&Srm[row * 128 + (kc ^ swz(row))]);
        }
      } else {
#pragma unroll
        for (int a = 0; a < 2; ++a)
          af[a] = *reinterpret_cast<const u16x8*>(&Scm[(64 * wr + 32 * a + (lane & 31)) * LDC + kc]);
#pragma unroll
        for (int b = 0; b < 2; ++b)
          bf[b] = *reinterpret_cast<const u16x8*>(&Scm[(64 * wc + 32 * b + (lane & 31)) * LDC + kc]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mfma32x32x16<PREC>(af[a], bf[b], acc[a][b]);
    }
    if (s + 1 < s_end) {
      uint16_t* nx = lds + ((s + 1 - s_begin) & 1) * BUF;
      store(nx, nx + 128 * 128);
      load(min(s + 2, s_end - 1));
    }
    __syncthreads();
  }
  float* out = (g == 0 ? part1 : part2) + (long)blockIdx.x * (128 * 128);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = 64 * wr + 32 * a + acc32_row(e, lane);
        const int col = 64 * wc + 32 * b + (lane & 31);
        out[row * 128 + col] = acc[a][b][e];
      }
  sumsq = wave_sum(sumsq);
  if (lane == 0) red[w] = sumsq;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int i = 0; i < 8; ++i) t += red[i];
    norms[blockIdx.x] = t;
  }
}

int slab_gram_launch(hipStream_t st, const float* x, int n1, int n2, long S, int prec, int n_blocks,
                     float* part1, float* part2, double* norms) {
  if (n1 < 1 || n2 < 1 || n1 > 128 || n2 > 128 || n1 % 4 || n2 % 4 || S < 1 || n_blocks < 1) return 8;
  if (((uintptr_t)x & 15) != 0) return 8;
  if (prec == PREC_F16)
    hipLaunchKernelGGL(slab_gram_kernel<PREC_F16>, dim3(n_blocks), dim3(512), 0, st, x, n1, n2, S, n_blocks,
                       part1, part2, norms);
  else
    hipLaunchKernelGGL(slab_gram_kernel<PREC_BF16>, dim3(n_blocks), dim3(512), 0, st, x, n1, n2, S, n_blocks,
                       part1, part2, norms);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// Orders the LDS accesses of one wave (lanes read what other lanes of the same wave wrote
// before): LDS operations of a wave complete in order; the fences keep the compiler from
// moving loads or stores across this point.
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------ slab_proj
// One wavefront per slab S (n1 x n2, column-major, n1 % 16 == 0, n2 % 16 == 0,
// r1, r2 <= 32). FP32 MFMA 16x16x4 with the reduction index permuted so that every
// lane reads float4 runs of a slab column:
//   T^T (n2 x r1) = S^T U1   -- A operand S^T: lane (j = l%16, g = l/16) at step s uses
//                               i = i0 + 4g + s, i.e. S[i0+4g .. i0+4g+3][j] (one float4);
//   Z^T (r2 x r1) = U2^T T^T -- the T^T accumulators are already in B-operand layout
//                               (their element e is row 4g + e = the slot of step e).
// U1 and U2 sit in LDS in the B/A-operand order (one float4 = the four steps of a lane).
// The slab's 16-column blocks (16 n1 contiguous floats) come from HBM as whole 1 KB wave
// loads (lane l takes floats 4 l + 256 i), streaming (nontemporal), into registers one
// block ahead; each block is then written to a per-wave LDS tile (column stride 132
// floats: conflict-free float4 reads) and read back in the operand layout. A wave's
// loads and LDS use are its own, so only wave-level ordering is needed. The loads are
// unconditional (past the last block the current one is loaded again), so the waits stay
// one block behind; a scheduling barrier keeps them above the MFMAs.
template <int RB1, int RB2>
__global__ __launch_bounds__(256) void slab_proj_kernel(const float* __restrict__ x, int n1, int n2, long S,
                                                        const float* __restrict__ U1, int r1,
                                                        const float* __restrict__ U2, int r2,
                                                        float* __restrict__ z) {
  // lu1[((a * 8 + ig) * 4 + g) * 16 + col] = U1[16 ig + 4 g + 0..3][16 a + col], lu2 alike
  __shared__ float4 lu1[RB1 * 512], lu2[RB2 * 512];
  constexpr int SLD = 132;
  __shared__ __attribute__((aligned(16))) float stage[4][16 * SLD];
  const int n1g = n1 / 16;  // row groups of 16 (<= 8)
  const int nb2 = n2 / 16;  // column blocks of 16 (<= 8)
  for (int e = threadIdx.x; e < RB1 * 512; e += blockDim.x) {
    const int cl = e & 15, gg = (e >> 4) & 3, ig = (e >> 6) & 7, a = e >> 9;
    const int c = 16 * a + cl, i = 16 * ig + 4 * gg;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ig < n1g && c < r1) {
      const float* p = U1 + i + (long)n1 * c;
      v = make_float4(p[0], p[1], p[2], p[3]);
    }
    lu1[e] = v;
  }
  for (int e = threadIdx.x; e < RB2 * 512; e += blockDim.x) {
    const int cl = e & 15, gg = (e >> 4) & 3, ig = (e >> 6) & 7, b = e >> 9;
    const int c = 16 * b + cl, i = 16 * ig + 4 * gg;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ig < nb2 && c < r2) {
      const float* p = U2 + i + (long)n2 * c;
      v = make_float4(p[0], p[1], p[2], p[3]);
    }
    lu2[e] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, g = lane >> 4;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const long n_waves = ((long)gridDim.x * blockDim.x) >> 6;
  if (wave >= S) return;
  const long slab = (long)n1 * n2;
  float* stg = stage[wv];

  auto load = [&](f32x4* R, long s, int bb) {
    const float* base = x + s * slab + (long)(16 * bb) * n1;
#pragma unroll
    for (int i = 0; i < 8; ++i)
      R[i] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(base + 4 * lane + 256 * min(i, n1g - 1)));
  };
  auto stage_store = [&](const f32x4* R) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (i < n1g) {
        const int off = 4 * lane + 256 * i, c = off / n1, r = off - c * n1;
        *reinterpret_cast<f32x4*>(&stg[c * SLD + r]) = R[i];
      }
    }
  };
  f32x4 zt[RB2][RB1];
#pragma unroll
  for (int b = 0; b < RB2; ++b)
#pragma unroll
    for (int a = 0; a < RB1; ++a) zt[b][a] = f32x4{};
  auto compute = [&](long s, int bb) {
    f32x4 sv[8];
#pragma unroll
    for (int ig = 0; ig < 8; ++ig)
      sv[ig] = ig < n1g ? *reinterpret_cast<const f32x4*>(&stg[col * SLD + 16 * ig + 4 * g]) : f32x4{};
    f32x4 tt[RB1];
#pragma unroll
    for (int a = 0; a < RB1; ++a) tt[a] = f32x4{};
#pragma unroll
    for (int ig = 0; ig < 8; ++ig) {
#pragma unroll
      for (int a = 0; a < RB1; ++a) {
        const float4 u = lu1[((a * 8 + ig) * 4 + g) * 16 + col];
        const float us[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int st = 0; st < 4; ++st) tt[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(sv[ig][st], us[st], tt[a], 0, 0, 0);
      }
    }
    // tt[a] element e: T^T[row 16bb + 4g + e][col 16a + col].
#pragma unroll
    for (int b = 0; b < RB2; ++b) {
      const float4 u = lu2[((b * 8 + bb) * 4 + g) * 16 + col];
      const float us[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int a = 0; a < RB1; ++a)
          zt[b][a] = __builtin_amdgcn_mfma_f32_16x16x4f32(us[st], tt[a][st], zt[b][a], 0, 0, 0);
    }
    if (bb == nb2 - 1) {
      // zt[b][a] element e: Z^T[row 16b + 4g + e][col 16a + col] = Z[a'][b'].
      float* zs = z + s * (long)r1 * r2;
#pragma unroll
      for (int b = 0; b < RB2; ++b)
#pragma unroll
        for (int a = 0; a < RB1; ++a) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int bp = 16 * b + 4 * g + e, ap = 16 * a + col;
            if (ap < r1 && bp < r2) zs[ap + (long)r1 * bp] = zt[b][a][e];
          }
          zt[b][a] = f32x4{};
        }
    }
  };
  f32x4 R[8];
  long s = wave;
  int bb = 0;
  load(R, s, bb);
  while (true) {
    long s2 = s;
    int bb2 = bb + 1;
    if (bb2 == nb2) { bb2 = 0; s2 += n_waves; }
    const bool more = s2 < S;
    wave_sync_lds();  // the previous block's operand reads are done
    stage_store(R);
    wave_sync_lds();
    load(R, more ? s2 : s, more ? bb2 : bb);
    __builtin_amdgcn_sched_barrier(0);
    compute(s, bb);
    if (!more) break;
    s = s2;
    bb = bb2;
  }
}

int slab_proj_launch(hipStream_t st, const float* x, int n1, int n2, long S, const float* U1, int r1,
                     const float* U2, int r2, float* z, int max_blocks) {
  if (n1 < 16 || n2 < 16 || n1 % 16 || n2 % 16 || n1 > 128 || n2 > 128 || r1 < 1 || r2 < 1 || r1 > 32 || r2 > 32)
    return 8;
  if (((uintptr_t)x & 15) != 0) return 8;
  // three blocks per CU fit the LDS (U tiles and four 8.4 KB stages each)
  long waves = S;
  long blocks = (waves + 3) / 4;
  if (blocks > max_blocks) blocks = max_blocks;
  if (blocks < 1) blocks = 1;
  dim3 grid((unsigned)blocks);
  if (r1 <= 16 && r2 <= 16)
    hipLaunchKernelGGL((slab_proj_kernel<1, 1>), grid, dim3(256), 0, st, x, n1, n2, S, U1, r1, U2, r2, z);
  else if (r1 <= 16)
    hipLaunchKernelGGL((slab_proj_kernel<1, 2>), grid, dim3(256), 0, st, x, n1, n2, S, U1, r1, U2, r2, z);
  else if (r2 <= 16)
    hipLaunchKernelGGL((slab_proj_kernel<2, 1>), grid, dim3(256), 0, st, x, n1, n2, S, U1, r1, U2, r2, z);
  else
    hipLaunchKernelGGL((slab_proj_kernel<2, 2>), grid, dim3(256), 0, st, x, n1, n2, S, U1, r1, U2, r2, z);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

}  // namespace ht
