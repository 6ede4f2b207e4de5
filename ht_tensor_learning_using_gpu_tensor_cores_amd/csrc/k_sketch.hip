// Learned per-mode sketch: forward and backward kernels of the Khatri-Rao test matrix
// and of the sketched-range loss, and the Adam update (gfx950).
//
// No counterpart in the reference: its range finder draws a dense Gaussian test matrix
// (fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:419-444). Here
// each mode m owns a sketch S_m (w x n_m, row-major: S_m[j, i] at off_m + j*n_m + i) and
// the test matrix of a node whose matricisation has the column modes m0..m1-1 is
//   Omega[c, j] = prod_m S_m[j, i_m(c)],   c = column-major linear index of (i_m0, ..., i_m1-1).
// The large products of the loss (A Omega, A^T Y, ...) run on the skinny MFMA GEMM kernels
// (k_gemm.hip); this file holds the pieces around them.
#include "ht_common.h"

namespace ht {

struct KRModes {
  int n[HT_KR_MAX];    // sizes of the column modes, fastest first
  long off[HT_KR_MAX]; // offsets of their sketches in the flat sketch array
  int nm;              // number of column modes
  int w;               // sketch width
  long cols;           // product of the n
};

// Omega (cols x w, column-major) from the per-mode sketches.
__global__ __launch_bounds__(256) void kr_omega_kernel(const float* __restrict__ S, KRModes km, float* __restrict__ om) {
  const long total = km.cols * km.w;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long c = idx % km.cols;
    const int j = (int)(idx / km.cols);
    long r = c;
    float v = 1.f;
    for (int k = 0; k < km.nm; ++k) {
      const int i = (int)(r % km.n[k]);
      r /= km.n[k];
      v *= S[km.off[k] + (long)j * km.n[k] + i];
    }
    om[idx] = v;
  }
}

// Backward of kr_omega: grad_m[j, i] += scale * sum_{c : i_m(c) = i} dOm[c, j] prod_{m' != m} S_m'[j, i_m'(c)].
// One thread per (mode, row j, index i).
__global__ __launch_bounds__(256) void kr_grad_kernel(const float* __restrict__ S, const float* __restrict__ dom,
                                                      KRModes km, long total, float scale, float* __restrict__ grad) {
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    long e = idx;
    int k = 0;
    while (k < km.nm - 1 && e >= (long)km.w * km.n[k]) {
      e -= (long)km.w * km.n[k];
      ++k;
    }
    const int j = (int)(e / km.n[k]);
    const int i = (int)(e % km.n[k]);
    const long others = km.cols / km.n[k];
    double acc = 0.0;
    for (long o = 0; o < others; ++o) {
      long r = o, c = 0, stride = 1;
      float pr = 1.f;
      for (int kk = 0; kk < km.nm; ++kk) {
        int d;
        if (kk == k) {
          d = i;
        } else {
          d = (int)(r % km.n[kk]);
          r /= km.n[kk];
          pr *= S[km.off[kk] + (long)j * km.n[kk] + d];
        }
        c += (long)d * stride;
        stride *= km.n[kk];
      }
      acc += (double)dom[c + km.cols * j] * (double)pr;
    }
    float* g = grad + km.off[k] + (long)j * km.n[k] + i;
    *g += (float)(scale * acc);
  }
}

// Partial products A^T B of two tall FP32 matrices (A: rows x wa, B: rows x wb, ld rows),
// FP64 accumulation: block b writes part[b][wa*wb] (column-major) over its share of
// 128-row chunks.
constexpr int TG_ROWS = 128;
constexpr int TG_LD = TG_ROWS + 1;  // odd stride: lanes reading different columns hit different banks
__global__ __launch_bounds__(256) void tall_gram_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                        long rows, int wa, int wb, double* __restrict__ part) {
  __shared__ float sa[64 * TG_LD];
  __shared__ float sb[64 * TG_LD];
  const int tid = threadIdx.x;
  const int np = wa * wb;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  for (long r0 = (long)blockIdx.x * TG_ROWS; r0 < rows; r0 += (long)gridDim.x * TG_ROWS) {
    const long nr = rows - r0 < TG_ROWS ? rows - r0 : TG_ROWS;
    for (int e = tid; e < TG_ROWS * wa; e += 256) {
      const int r = e % TG_ROWS, j = e / TG_ROWS;
      sa[j * TG_LD + r] = r < nr ? A[r0 + r + rows * j] : 0.f;
    }
    for (int e = tid; e < TG_ROWS * wb; e += 256) {
      const int r = e % TG_ROWS, j = e / TG_ROWS;
      sb[j * TG_LD + r] = r < nr ? B[r0 + r + rows * j] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int p = tid + 256 * q;
      if (p < np) {
        const int a = p % wa, b = p / wa;
        const float* pa = sa + a * TG_LD;
        const float* pb = sb + b * TG_LD;
        double s = 0.0;
        for (int r = 0; r < TG_ROWS; ++r) s += (double)pa[r] * (double)pb[r];
        acc[q] += s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int p = tid + 256 * q;
    if (p < np) part[(long)blockIdx.x * np + p] = acc[q];
  }
}

// Partial Y^T Y of a tall FP32 Y (rows x b, ld rows, 16 <= b <= 64) on FP64 MFMA
// (16x16x4): block blk takes 128-row chunks blk, blk + gridDim.x, ..., stages each as FP64
// in LDS (row-major, one conversion per value), and wave w accumulates every upper 16 x 16
// tile pair (I <= J, at most 10) over rows [32 w, 32 w + 32) of each chunk, so the waves
// carry equal MFMA work (FP64 MFMA issues at a quarter of the BF16 rate: it is the unit
// that bounds this kernel). The next chunk is loaded into registers while the current one
// is multiplied. The waves' tiles are summed in a fixed order through LDS and the block's
// b x b partial (column-major, both triangles) goes to part[blk]. The products of FP32
// values are exact in FP64, so the sums are FP64 sums of exact terms, as in
// tall_gram_kernel, whose VALU form spends two conversions per product.
constexpr int TS_ROWS = 128, TS_LD = 65, TS_PER = TS_ROWS * 64 / 256;
__global__ __launch_bounds__(256) void tall_syrk_kernel(const float* __restrict__ Y, long rows, int b,
                                                        double* __restrict__ part) {
  __shared__ double Ys[TS_ROWS * TS_LD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int bt = (b + 15) / 16, np = bt * (bt + 1) / 2;
  const int bp = 16 * bt;  // staged (padded) width
  f64x4 acc[10];
#pragma unroll
  for (int u = 0; u < 10; ++u) acc[u] = f64x4{};
  const long n_chunks = (rows + TS_ROWS - 1) / TS_ROWS;
  float nxt[TS_PER];
  auto load = [&](long ch) {
    const long r0 = ch * TS_ROWS;
#pragma unroll
    for (int q = 0; q < TS_PER; ++q) {
      const int e = tid + 256 * q, row = e % TS_ROWS, col = e / TS_ROWS;
      const long r = r0 + row;
      nxt[q] = (col < b && r < rows) ? Y[r + rows * col] : 0.f;
    }
  };
  long ch = blockIdx.x;
  if (ch < n_chunks) load(ch);
  for (; ch < n_chunks; ch += gridDim.x) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TS_PER; ++q) {
      const int e = tid + 256 * q, row = e % TS_ROWS, col = e / TS_ROWS;
      if (col < bp) Ys[row * TS_LD + col] = (double)nxt[q];
    }
    __syncthreads();
    if (ch + gridDim.x < n_chunks) load(ch + gridDim.x);
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 4) {
      const double* yr = Ys + (32 * w + k0 + (lane >> 4)) * TS_LD + (lane & 15);
      double v[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) v[t] = t < bt ? yr[16 * t] : 0.0;
      int u = 0;
#pragma unroll
      for (int I = 0; I < 4; ++I)
#pragma unroll
        for (int J = I; J < 4; ++J) {
          if (J < bt) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[I], v[J], acc[u], 0, 0, 0);
          ++u;
        }
    }
  }
  // sum the four waves' tiles in LDS (wave order fixed), then write the b x b partial
  __syncthreads();
  double* red = Ys;  // 4 x np x 256 doubles <= 10240 fits the staging array (8320) in two halves
  double* out = part + (long)blockIdx.x * b * b;
  for (int half = 0; half < 2; ++half) {
    const int u0 = 5 * half;
    __syncthreads();
    int u = 0;
#pragma unroll
    for (int I = 0; I < 4; ++I)
#pragma unroll
      for (int J = I; J < 4; ++J) {
        if (u >= u0 && u < u0 + 5 && J < bt) {
#pragma unroll
          for (int e = 0; e < 4; ++e) red[((w * 5 + (u - u0)) * 4 + e) * 64 + lane] = acc[u][e];
        }
        ++u;
      }
    __syncthreads();
    if (w == 0) {
      int uu = 0;
#pragma unroll
      for (int I = 0; I < 4; ++I)
#pragma unroll
        for (int J = I; J < 4; ++J) {
          if (uu >= u0 && uu < u0 + 5 && J < bt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int o = ((uu - u0) * 4 + e) * 64 + lane;
              const double t = (red[o] + red[o + 5 * 256]) + (red[o + 10 * 256] + red[o + 15 * 256]);
              const int i = 16 * I + acc16_row_f64(e, lane), j = 16 * J + (lane & 15);
              if (i < b && j < b) {
                out[i + (long)b * j] = t;
                if (I != J) out[j + (long)b * i] = t;
              }
            }
          }
          ++uu;
        }
    }
  }
}

// out[p] = sum_b part[b][p]: a block takes 64 entries (one per lane, consecutive in
// memory) and its 4 waves split the partials (wave w: w, w + 4, ...); the waves' sums are
// added in a fixed order through LDS, so the result does not depend on timing.
__global__ __launch_bounds__(256) void tall_gram_reduce_kernel(const double* __restrict__ part, int nb, int np,
                                                               double* __restrict__ out) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (p < np) {
#pragma unroll 8
    for (int b = w; b < nb; b += 4) s += part[(long)b * np + p];
  }
  red[w][lane] = s;
  __syncthreads();
  if (w == 0 && p < np) out[p] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// Small FP64 solve of one node's loss (one work-group, w <= 64):
//   Wr = W + ridge * tr(W)/w * I = L L^T,  C1 = Wr^-1,  C2 = C1 E C1,
//   C2s = (2/xsq) (C2 + (ridge/w) tr(C2) I),
//   loss_acc[0] += 1 - tr(C1 E)/xsq,  coef[0] = -2/xsq,  info[0] |= 1 on a non-positive pivot.
__global__ __launch_bounds__(256) void sketch_solve_kernel(const double* __restrict__ Wg, const double* __restrict__ Eg,
                                                           const double* __restrict__ xsq, int w, double ridge,
                                                           double* __restrict__ C1, double* __restrict__ C2s,
                                                           double* __restrict__ coef, double* __restrict__ loss_acc,
                                                           int* __restrict__ info) {
  constexpr int LD = 65;
  __shared__ double L[64 * LD];
  __shared__ double T[64 * LD];
  __shared__ double Wi[64 * LD];
  __shared__ double red[4], red2[4];
  __shared__ double sh_tr;
  const int tid = threadIdx.x;
  for (int e = tid; e < w * w; e += 256) L[(e % w) * LD + e / w] = Wg[e];
  __syncthreads();
  if (tid == 0) {
    double tr = 0.0;
    for (int i = 0; i < w; ++i) tr += L[i * LD + i];
    sh_tr = tr;
  }
  __syncthreads();
  if (tid < w) L[tid * LD + tid] += ridge * sh_tr / w;
  __syncthreads();
  // Cholesky, right-looking, lower triangle
  for (int k = 0; k < w; ++k) {
    if (tid == 0) {
      double d = L[k * LD + k];
      if (!(d > 0.0)) {
        info[0] |= 1;
        d = 1e-300;
      }
      L[k * LD + k] = sqrt(d);
    }
    __syncthreads();
    for (int i = k + 1 + tid; i < w; i += 256) L[i * LD + k] /= L[k * LD + k];
    __syncthreads();
    const int m = w - k - 1;
    for (int e = tid; e < m * m; e += 256) {
      const int i = k + 1 + e % m, j = k + 1 + e / m;
      if (j <= i) L[i * LD + j] -= L[i * LD + k] * L[j * LD + k];
    }
    __syncthreads();
  }
  // T = L^-1 (column j by thread j, forward substitution)
  if (tid < w) {
    const int j = tid;
    for (int i = 0; i < w; ++i) {
      double s = (i == j) ? 1.0 : 0.0;
      for (int m = j; m < i; ++m) s -= L[i * LD + m] * T[m * LD + j];
      T[i * LD + j] = i < j ? 0.0 : s / L[i * LD + i];
    }
  }
  __syncthreads();
  // Wi = L^-T L^-1
  for (int e = tid; e < w * w; e += 256) {
    const int i = e % w, j = e / w;
    double s = 0.0;
    for (int m = i > j ? i : j; m < w; ++m) s += T[m * LD + i] * T[m * LD + j];
    Wi[i * LD + j] = s;
  }
  __syncthreads();
  // T = Wi E
  for (int e = tid; e < w * w; e += 256) {
    const int i = e % w, j = e / w;
    double s = 0.0;
    for (int m = 0; m < w; ++m) s += Wi[i * LD + m] * Eg[m + w * j];
    T[i * LD + j] = s;
  }
  __syncthreads();
  // L = C2 = Wi E Wi (L is free after the inverse)
  for (int e = tid; e < w * w; e += 256) {
    const int i = e % w, j = e / w;
    double s = 0.0;
    for (int m = 0; m < w; ++m) s += T[i * LD + m] * Wi[m * LD + j];
    L[i * LD + j] = s;
  }
  double cap = 0.0, tc2 = 0.0;
  for (int i = tid; i < w; i += 256) cap += T[i * LD + i];
  __syncthreads();
  for (int i = tid; i < w; i += 256) tc2 += L[i * LD + i];
  cap = wave_sum(cap);
  tc2 = wave_sum(tc2);
  if ((tid & 63) == 0) {
    red[tid >> 6] = cap;
    red2[tid >> 6] = tc2;
  }
  __syncthreads();
  const double xs = xsq[0];
  const double c2 = red2[0] + red2[1] + red2[2] + red2[3];
  // the ridge follows tr(Y^T Y), which adds (ridge/w) tr(C2) I to C2 in the gradient
  for (int e = tid; e < w * w; e += 256) {
    const int i = e % w, j = e / w;
    C2s[e] = 2.0 / xs * (L[i * LD + j] + (i == j ? ridge / w * c2 : 0.0));
    C1[e] = Wi[i * LD + j];
  }
  if (tid == 0) {
    const double c = red[0] + red[1] + red[2] + red[3];
    loss_acc[0] += 1.0 - c / xs;
    coef[0] = -2.0 / xs;
  }
}

// out (rows x w, FP32) = A (rows x w, FP32) C (w x w, FP64)  [+ beta[0] * B]. One
// thread per row: its w values of A are loaded at once (every load issued
// unconditionally: columns past w read column w - 1 and are never used; a wave's load
// of a column covers 64 consecutive rows) and kept in registers, and the outputs are
// formed 16 columns at a time in FP64, so A is streamed once. C is read with
// wave-uniform indices (scalar loads; no LDS staging). W: w rounded up to 8.
template <int W>
__global__ __launch_bounds__(256) void tall_small_kernel(const float* __restrict__ A, long rows, int w,
                                                         const double* __restrict__ C, const float* __restrict__ B,
                                                         const double* __restrict__ beta, float* __restrict__ out) {
  const double bt = B ? beta[0] : 0.0;
  for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (long)gridDim.x * blockDim.x) {
    float a[W];
#pragma unroll
    for (int k = 0; k < W; ++k) a[k] = A[r + rows * min(k, w - 1)];
    for (int j0 = 0; j0 < w; j0 += 16) {
      double acc[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) acc[u] = 0.0;
#pragma unroll
      for (int k = 0; k < W; ++k) {
        if (k < w) {
          const double ak = (double)a[k];
#pragma unroll
          for (int u = 0; u < 16; ++u) acc[u] = fma(ak, C[k + w * min(j0 + u, w - 1)], acc[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = j0 + u;
        if (j < w) {
          double v = acc[u];
          if (B) v += bt * (double)B[r + rows * j];
          out[r + rows * j] = (float)v;
        }
      }
    }
  }
}

// Adam on the sketch: g is scaled by `scale` (1 / tensors per step) and cleared. The step
// is skipped (g only cleared) when the step's summed loss is below min_loss_sum: at that
// level the gradient is rounding noise of the FP32 products.
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ S, float* __restrict__ g, float* __restrict__ m,
                                                   float* __restrict__ v, long n, float lr, float b1, float b2,
                                                   float eps, float scale, float bc1, float bc2,
                                                   const double* __restrict__ step_loss, double min_loss_sum) {
  const bool skip = step_loss[0] < min_loss_sum;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    if (skip) {
      g[i] = 0.f;
      continue;
    }
    const float gi = g[i] * scale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    S[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    g[i] = 0.f;
  }
}

__global__ __launch_bounds__(256) void fill_f32_kernel(float* __restrict__ a, long n, float val) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = val;
}
__global__ void fill_f64_kernel(double* __restrict__ a, long n, double val) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = val;
}

// Residual of probe vectors against an orthonormal basis Q (rows x b, FP32):
//   acc[0] += sum_i ||P_i - Q (Q^T P_i)||^2 given QtP = Q^T P (b x np, FP64).
// One thread per row, block partial sums into part[block].
__global__ __launch_bounds__(256) void probe_resid_kernel(const float* __restrict__ Q, const float* __restrict__ P,
                                                          const double* __restrict__ QtP, long rows, int b, int np,
                                                          double* __restrict__ part) {
  __shared__ double qp[64 * 8];
  __shared__ double red[4];
  for (int e = threadIdx.x; e < b * np; e += blockDim.x) qp[e] = QtP[e];
  __syncthreads();
  double s = 0.0;
  for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (long)gridDim.x * blockDim.x) {
    for (int i = 0; i < np; ++i) {
      double v = P[r + rows * i];
      for (int k = 0; k < b; ++k) v -= (double)Q[r + rows * k] * qp[k + b * i];
      s += v * v;
    }
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// ------------------------------------------------------------------ launchers
static unsigned grid_for(long total, long cap = 4096) {
  long b = (total + 255) / 256;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

static int make_modes(const int64_t* n, const int64_t* off, int m0, int m1, int w, KRModes* km) {
  if (m1 - m0 < 1 || m1 - m0 > HT_KR_MAX || w < 1 || w > 64) return 8;
  km->nm = m1 - m0;
  km->w = w;
  km->cols = 1;
  for (int k = 0; k < km->nm; ++k) {
    km->n[k] = (int)n[m0 + k];
    km->off[k] = (long)off[m0 + k];
    km->cols *= n[m0 + k];
  }
  return 0;
}

int kr_omega_launch(hipStream_t st, const float* S, const int64_t* n, const int64_t* off, int m0, int m1, int w,
                    float* om) {
  KRModes km;
  if (make_modes(n, off, m0, m1, w, &km)) return 8;
  hipLaunchKernelGGL(kr_omega_kernel, dim3(grid_for(km.cols * w)), dim3(256), 0, st, S, km, om);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int kr_grad_launch(hipStream_t st, const float* S, const float* dom, const int64_t* n, const int64_t* off, int m0,
                   int m1, int w, float scale, float* grad) {
  KRModes km;
  if (make_modes(n, off, m0, m1, w, &km)) return 8;
  long total = 0;
  for (int k = 0; k < km.nm; ++k) total += (long)w * km.n[k];
  hipLaunchKernelGGL(kr_grad_kernel, dim3(grid_for(total)), dim3(256), 0, st, S, dom, km, total, scale, grad);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int tall_gram_blocks(long rows) {
  long nb = (rows + TG_ROWS - 1) / TG_ROWS;
  return (int)(nb > 1024 ? 1024 : (nb < 1 ? 1 : nb));
}

int tall_gram_launch(hipStream_t st, const float* A, int wa, const float* B, int wb, long rows, double* part,
                     double* out) {
  if (wa < 1 || wa > 64 || wb < 1 || wb > 64 || rows < 1) return 8;
  const int nb = tall_gram_blocks(rows);
  if (A == B && wa == wb && wa >= 16)
    hipLaunchKernelGGL(tall_syrk_kernel, dim3(nb), dim3(256), 0, st, A, rows, wa, part);
  else
    hipLaunchKernelGGL(tall_gram_kernel, dim3(nb), dim3(256), 0, st, A, B, rows, wa, wb, part);
  hipLaunchKernelGGL(tall_gram_reduce_kernel, dim3((wa * wb + 63) / 64), dim3(256), 0, st, part, nb, wa * wb, out);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int sketch_solve_launch(hipStream_t st, const double* W, const double* E, const double* xsq, int w, double ridge,
                        double* C1, double* C2s, double* coef, double* loss_acc, int* info) {
  if (w < 1 || w > 64) return 8;
  hipLaunchKernelGGL(sketch_solve_kernel, dim3(1), dim3(256), 0, st, W, E, xsq, w, ridge, C1, C2s, coef, loss_acc,
                     info);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int tall_small_launch(hipStream_t st, const float* A, long rows, int w, const double* C, const float* B,
                      const double* beta, float* out) {
  if (w < 1 || w > 64) return 8;
  // one row per thread where the grid allows (the occupancy hides the loads' latency)
  const dim3 g(grid_for(rows, 1L << 20)), blk(256);
  switch ((w + 7) / 8) {
    case 1: hipLaunchKernelGGL(tall_small_kernel<8>, g, blk, 0, st, A, rows, w, C, B, beta, out); break;
    case 2: hipLaunchKernelGGL(tall_small_kernel<16>, g, blk, 0, st, A, rows, w, C, B, beta, out); break;
    case 3: hipLaunchKernelGGL(tall_small_kernel<24>, g, blk, 0, st, A, rows, w, C, B, beta, out); break;
    case 4: hipLaunchKernelGGL(tall_small_kernel<32>, g, blk, 0, st, A, rows, w, C, B, beta, out); break;
    case 5: hipLaunchKernelGGL(tall_small_kernel<40>, g, blk, 0, st, A, rows, w, C, B, beta, out); break;
    case 6: hipLaunchKernelGGL(tall_small_kernel<48>, g, blk, 0, st, A, rows, w, C, B, beta, out); break;
    case 7: hipLaunchKernelGGL(tall_small_kernel<56>, g, blk, 0, st, A, rows, w, C, B, beta, out); break;
    default: hipLaunchKernelGGL(tall_small_kernel<64>, g, blk, 0, st, A, rows, w, C, B, beta, out); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int adam_launch(hipStream_t st, float* S, float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                float scale, float bc1, float bc2, const double* step_loss, double min_loss_sum) {
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, st, S, g, m, v, n, lr, b1, b2, eps, scale, bc1,
                     bc2, step_loss, min_loss_sum);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// acc[0] += step[0]; step[0] = 0 (one thread).
__global__ void fold_loss_kernel(double* __restrict__ step, double* __restrict__ acc) {
  if (threadIdx.x == 0) {
    acc[0] += step[0];
    step[0] = 0.0;
  }
}
int fold_loss_launch(hipStream_t st, double* step, double* acc) {
  hipLaunchKernelGGL(fold_loss_kernel, dim3(1), dim3(64), 0, st, step, acc);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int fill_f32_launch(hipStream_t st, float* a, long n, float val) {
  hipLaunchKernelGGL(fill_f32_kernel, dim3(grid_for(n)), dim3(256), 0, st, a, n, val);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int fill_f64_launch(hipStream_t st, double* a, long n, double val) {
  hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, a, n, val);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int probe_resid_blocks(long rows) { return (int)grid_for(rows, 512); }

int probe_resid_launch(hipStream_t st, const float* Q, const float* P, const double* QtP, long rows, int b, int np,
                       double* part, double* out) {
  if (b < 1 || b > 64 || np < 1 || np > 8) return 8;
  const int nb = probe_resid_blocks(rows);
  hipLaunchKernelGGL(probe_resid_kernel, dim3(nb), dim3(256), 0, st, Q, P, QtP, rows, b, np, part);
  hipLaunchKernelGGL(tall_gram_reduce_kernel, dim3(1), dim3(256), 0, st, part, nb, 1, out);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

}  // namespace ht
