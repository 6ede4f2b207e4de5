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

constexpr int SK_B = 64;     // padded skinny width
constexpr int SK_TILE = 128; // rows of the output tile
constexpr int SK_K = 64;     // reduction chunk

// ------------------------------------------------------------------ gemm_atb
// Y(K x b) = A^T B. Block: 128 columns of A (128 output rows) x 64; 4 waves, wave w owns
// output rows 32w..32w+31 and all 64 columns (two 32x32 accumulators). LDS tiles are
// [column][m] for A and [j][m] for B: both are column-major in memory along m.
// NB = number of 32-column accumulators (1 when b <= 32: half the MFMA work).
template <int PREC, int NB>
__global__ __launch_bounds__(256) void gemm_atb_kernel(const float* __restrict__ A, long M, long K,
                                                       const float* __restrict__ B, int b, int n_split,
                                                       float* __restrict__ part) {
  constexpr int LD = SK_K + 4;
  __shared__ __attribute__((aligned(16))) float As[SK_TILE * LD];
  __shared__ __attribute__((aligned(16))) float Bs[SK_B * LD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long c0 = (long)blockIdx.x * SK_TILE;
  const int split = blockIdx.y;
  const long n_chunks = (M + SK_K - 1) / SK_K;
  const long k_begin = (long)split * n_chunks / n_split, k_end = (long)(split + 1) * n_chunks / n_split;
  const bool vecA = (M % 4 == 0) && (((uintptr_t)A & 15) == 0);
  const bool vecB = (M % 4 == 0) && (((uintptr_t)B & 15) == 0);

  f32x16 acc[NB];
#pragma unroll
  for (int bb = 0; bb < NB; ++bb) acc[bb] = f32x16{};
  for (long ch = k_begin; ch < k_end; ++ch) {
    const long m0 = ch * SK_K;
    __syncthreads();
    // A tile: 128 columns x 64 m (float4 along m): 2048 float4 / 256 threads = 8 each.
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int f = tid + 256 * it;
      const int col = f >> 4, mq = f & 15;
      const long c = c0 + col, m = m0 + 4 * mq;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (c < K) {
        const float* p = A + m + M * c;
        if (vecA && m + 3 < M) v = *reinterpret_cast<const float4*>(p);
        else {
          if (m < M) v.x = p[0];
          if (m + 1 < M) v.y = p[1];
          if (m + 2 < M) v.z = p[2];
          if (m + 3 < M) v.w = p[3];
        }
      }
      *reinterpret_cast<float4*>(&As[col * LD + 4 * mq]) = v;
    }
#pragma unroll
    for (int it = 0; it < 2 * NB; ++it) {
      const int f = tid + 256 * it;
      const int j = f >> 4, mq = f & 15;
      const long m = m0 + 4 * mq;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (j < b) {
        const float* p = B + m + M * j;
        if (vecB && m + 3 < M) v = *reinterpret_cast<const float4*>(p);
        else {
          if (m < M) v.x = p[0];
          if (m + 1 < M) v.y = p[1];
          if (m + 2 < M) v.z = p[2];
          if (m + 3 < M) v.w = p[3];
        }
      }
      *reinterpret_cast<float4*>(&Bs[j * LD + 4 * mq]) = v;
    }
    __syncthreads();
    if constexpr (PREC == PREC_F32) {
#pragma unroll 4
      for (int kk = 0; kk < SK_K / 2; ++kk) {
        const float af = As[(32 * w + (lane & 31)) * LD + 2 * kk + (lane >> 5)];
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) {
          const float bf = Bs[(32 * bb + (lane & 31)) * LD + 2 * kk + (lane >> 5)];
          acc[bb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af, bf, acc[bb], 0, 0, 0);
        }
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < SK_K / 16; ++kk) {
        u16x8 af;
        const float* ap = &As[(32 * w + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)];
#pragma unroll
        for (int i = 0; i < 8; ++i) af[i] = to16<PREC>(ap[i]);
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) {
          u16x8 bf;
          const float* bp = &Bs[(32 * bb + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)];
#pragma unroll
          for (int i = 0; i < 8; ++i) bf[i] = to16<PREC>(bp[i]);
          acc[bb] = mfma32x32x16<PREC>(af, bf, acc[bb]);
        }
      }
    }
  }
  // part[split][row c][64]
  float* out = part + (long)split * K * SK_B;
#pragma unroll
  for (int bb = 0; bb < NB; ++bb)
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const long c = c0 + 32 * w + acc32_row(e, lane);
      const int j = 32 * bb + (lane & 31);
      if (c < K) out[c * SK_B + j] = acc[bb][e];
    }
}

// ------------------------------------------------------------------ gemm_ab
// Y(M x b) = A B. Block: 128 rows of A x 64; reduction over columns of A in chunks of
// 64. A tile stored transposed in LDS ([m][c]), B tile as [j][c].
template <int PREC, int NB>
__global__ __launch_bounds__(256) void gemm_ab_kernel(const float* __restrict__ A, long M, long K,
                                                      const float* __restrict__ B, int b, int n_split,
                                                      float* __restrict__ part) {
  constexpr int LD = SK_K + 4;
  __shared__ __attribute__((aligned(16))) float As[SK_TILE * LD];
  __shared__ __attribute__((aligned(16))) float Bs[SK_B * LD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long m0 = (long)blockIdx.x * SK_TILE;
  const int split = blockIdx.y;
  const long n_chunks = (K + SK_K - 1) / SK_K;
  const long k_begin = (long)split * n_chunks / n_split, k_end = (long)(split + 1) * n_chunks / n_split;
  const bool vecA = (M % 4 == 0) && (((uintptr_t)A & 15) == 0);

  f32x16 acc[NB];
#pragma unroll
  for (int bb = 0; bb < NB; ++bb) acc[bb] = f32x16{};
  for (long ch = k_begin; ch < k_end; ++ch) {
    const long cc0 = ch * SK_K;
    __syncthreads();
    // A tile: 64 columns x 128 m; float4 along m (32 per column): 2048 float4 / 256 = 8.
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int f = tid + 256 * it;
      const int k = f >> 5, mq = f & 31;
      const long c = cc0 + k, m = m0 + 4 * mq;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (c < K) {
        const float* p = A + m + M * c;
        if (vecA && m + 3 < M) v = *reinterpret_cast<const float4*>(p);
        else {
          if (m < M) v.x = p[0];
          if (m + 1 < M) v.y = p[1];
          if (m + 2 < M) v.z = p[2];
          if (m + 3 < M) v.w = p[3];
        }
      }
      As[(4 * mq + 0) * LD + k] = v.x;
      As[(4 * mq + 1) * LD + k] = v.y;
      As[(4 * mq + 2) * LD + k] = v.z;
      As[(4 * mq + 3) * LD + k] = v.w;
    }
    // B tile: 64 (j) x 64 (c); B col-major (K x b): B[c + K*j].
#pragma unroll
    for (int it = 0; it < 8 * NB; ++it) {
      const int f = tid + 256 * it;
      const int j = f >> 6, k = f & 63;
      const long c = cc0 + k;
      Bs[j * LD + k] = (j < b && c < K) ? B[c + K * j] : 0.f;
    }
    __syncthreads();
    if constexpr (PREC == PREC_F32) {
#pragma unroll 4
      for (int kk = 0; kk < SK_K / 2; ++kk) {
        const float af = As[(32 * w + (lane & 31)) * LD + 2 * kk + (lane >> 5)];
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) {
          const float bf = Bs[(32 * bb + (lane & 31)) * LD + 2 * kk + (lane >> 5)];
          acc[bb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af, bf, acc[bb], 0, 0, 0);
        }
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < SK_K / 16; ++kk) {
        u16x8 af;
        const float* ap = &As[(32 * w + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)];
#pragma unroll
        for (int i = 0; i < 8; ++i) af[i] = to16<PREC>(ap[i]);
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) {
          u16x8 bf;
          const float* bp = &Bs[(32 * bb + (lane & 31)) * LD + 16 * kk + 8 * (lane >> 5)];
#pragma unroll
          for (int i = 0; i < 8; ++i) bf[i] = to16<PREC>(bp[i]);
          acc[bb] = mfma32x32x16<PREC>(af, bf, acc[bb]);
        }
      }
    }
  }
  float* out = part + (long)split * M * SK_B;
#pragma unroll
  for (int bb = 0; bb < NB; ++bb)
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const long m = m0 + 32 * w + acc32_row(e, lane);
      const int j = 32 * bb + (lane & 31);
      if (m < M) out[m * SK_B + j] = acc[bb][e];
    }
}

// part[split][row][64] -> Y (rows x b, column-major, ld = rows)
__global__ void skinny_reduce_kernel(const float* __restrict__ part, long rows, int b, int n_split,
                                     float* __restrict__ Y) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * SK_B) return;
  const long row = e / SK_B;
  const int j = (int)(e % SK_B);
  if (j >= b) return;
  double s = 0.0;
  for (int k = 0; k < n_split; ++k) s += (double)part[(long)k * rows * SK_B + e];
  Y[row + rows * j] = (float)s;
}

static int pick_split(long tiles, long chunks) {
  long s = 1024 / (tiles > 0 ? tiles : 1);
  if (s < 1) s = 1;
  if (s > chunks) s = chunks;
  if (s > 64) s = 64;
  return (int)s;
}

size_t gemm_ws_floats(long M, long K, int b) {
  (void)b;
  const long tA = (K + SK_TILE - 1) / SK_TILE, cA = (M + SK_K - 1) / SK_K;
  const long tB = (M + SK_TILE - 1) / SK_TILE, cB = (K + SK_K - 1) / SK_K;
  const long s1 = pick_split(tA, cA), s2 = pick_split(tB, cB);
  const long a = s1 * K * SK_B, c = s2 * M * SK_B;
  return (size_t)(a > c ? a : c);
}

int gemm_atb_launch(hipStream_t st, const float* A, long M, long K, const float* B, int b, int prec, float* Y,
                    float* ws, size_t ws_floats) {
  if (b < 1 || b > SK_B || M < 1 || K < 1) return 8;
  const long tiles = (K + SK_TILE - 1) / SK_TILE, chunks = (M + SK_K - 1) / SK_K;
  const int ns = pick_split(tiles, chunks);
  if ((size_t)ns * K * SK_B > ws_floats) return 4;
  dim3 grid((unsigned)tiles, ns);
  if (prec == PREC_BF16) {
    if (b <= 32) hipLaunchKernelGGL((gemm_atb_kernel<PREC_BF16, 1>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
    else hipLaunchKernelGGL((gemm_atb_kernel<PREC_BF16, 2>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
  }
  else if (prec == PREC_F16) {
    if (b <= 32) hipLaunchKernelGGL((gemm_atb_kernel<PREC_F16, 1>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
    else hipLaunchKernelGGL((gemm_atb_kernel<PREC_F16, 2>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
  }
  else {
    if (b <= 32) hipLaunchKernelGGL((gemm_atb_kernel<PREC_F32, 1>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
    else hipLaunchKernelGGL((gemm_atb_kernel<PREC_F32, 2>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
  }
  const long tot = K * SK_B;
  hipLaunchKernelGGL(skinny_reduce_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, ws, K, b, ns, Y);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int gemm_ab_launch(hipStream_t st, const float* A, long M, long K, const float* B, int b, int prec, float* Y,
                   float* ws, size_t ws_floats) {
  if (b < 1 || b > SK_B || M < 1 || K < 1) return 8;
  const long tiles = (M + SK_TILE - 1) / SK_TILE, chunks = (K + SK_K - 1) / SK_K;
  const int ns = pick_split(tiles, chunks);
  if ((size_t)ns * M * SK_B > ws_floats) return 4;
  dim3 grid((unsigned)tiles, ns);
  if (prec == PREC_BF16) {
    if (b <= 32) hipLaunchKernelGGL((gemm_ab_kernel<PREC_BF16, 1>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
    else hipLaunchKernelGGL((gemm_ab_kernel<PREC_BF16, 2>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
  }
  else if (prec == PREC_F16) {
    if (b <= 32) hipLaunchKernelGGL((gemm_ab_kernel<PREC_F16, 1>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
    else hipLaunchKernelGGL((gemm_ab_kernel<PREC_F16, 2>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
  }
  else {
    if (b <= 32) hipLaunchKernelGGL((gemm_ab_kernel<PREC_F32, 1>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
    else hipLaunchKernelGGL((gemm_ab_kernel<PREC_F32, 2>), grid, dim3(256), 0, st, A, M, K, B, b, ns, ws);
  }
  const long tot = M * SK_B;
  hipLaunchKernelGGL(skinny_reduce_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, ws, M, b, ns, Y);
  return hipGetLastError() == hipSuccess ? 0 : 5;
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
__global__ void residual_stream_kernel(const float* __restrict__ L, const float* __restrict__ R, long M, long N,
                                       int k, const float* __restrict__ X, int nts, double* __restrict__ acc_out);

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
  if (full && k <= 16) {
    nts = 32;
    while (nts > 4 && (N % (32L * nts) != 0 || grid.x * (N / (32L * nts)) < 2048)) nts /= 2;
    grid.y = (unsigned)(N / (32L * nts));
  }
  if (n_blocks_out) *n_blocks_out = (int)(grid.x * grid.y);
  if (full) {
    if (k <= 16)
      hipLaunchKernelGGL(residual_stream_kernel, grid, dim3(256), 0, st, L, R, M, N, k, X, nts, acc);
    else if (k <= 32)
      hipLaunchKernelGGL((residual_kernel<32, true>), grid, dim3(256), 0, st, L, R, M, N, k, X, nts, acc);
    else
      hipLaunchKernelGGL((residual_kernel<64, true>), grid, dim3(256), 0, st, L, R, M, N, k, X, nts, acc);
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

// The residual pass for k <= 16 with M a multiple of 128 and N of 32 nts: the same
// block, wave and lane layout as residual_kernel, no bounds checks, and the next
// subtile's X is loaded while the current one is multiplied. X is read once, with
// streaming (nontemporal) loads; each half-wave load covers whole 128-byte lines. The subtile after the last
// is the last again (loaded, never used), so every path issues the same loads and each
// wait covers only the subtile being multiplied; a scheduling barrier keeps the
// compiler from sinking the loads below the products.
__global__ __launch_bounds__(256) void residual_stream_kernel(const float* __restrict__ L, const float* __restrict__ R,
                                                              long M, long N, int k, const float* __restrict__ X,
                                                              int nts, double* __restrict__ acc_out) {
  constexpr int KP = 16;
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
  double se = 0.0, sx = 0.0;
  auto tile = [&](const float* xv, int t) {
    const long n = nbase + 32 * t + li;
    float ar[KP / 2];
#pragma unroll
    for (int j = 0; j < KP / 2; ++j) {
      const int kk = 2 * j + kh;
      const float v = R[n + N * (long)min(kk, k - 1)];
      ar[j] = kk < k ? v : 0.f;
    }
    f32x16 acc = f32x16{};
#pragma unroll
    for (int j = 0; j < KP / 2; ++j) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[j], bl[j], acc, 0, 0, 0);
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
  float xa[16], xb[16];
  load_x(xa, 0);
  for (int t = 0; t < nts; t += 2) {
    load_x(xb, min(t + 1, nts - 1));
    __builtin_amdgcn_sched_barrier(0);
    tile(xa, t);
    load_x(xa, min(t + 2, nts - 1));
    __builtin_amdgcn_sched_barrier(0);
    if (t + 1 < nts) tile(xb, t + 1);
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
