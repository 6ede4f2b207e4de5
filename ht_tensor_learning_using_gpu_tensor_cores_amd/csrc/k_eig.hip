// Small dense FP64 kernels of the HT engine: the projected eigen problems and helpers.
//
// jacobi_eig: symmetric eigen decomposition of an n x n (n <= 64) FP64 matrix by the
// cyclic parallel Jacobi method (round-robin pair ordering, one work-group, LDS), sorted
// descending with canonical signs. It is the (k+p)-sized projected solver that replaces
// cusolverDnSsyevd/Sgesvd at every node (third_order_tensor_decompsition/CUDA/Optimized/
// kernel.cu:682-718, fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/
// kernel.cu:267-287, :457-481).
//
// cholqr: Cholesky QR (two passes) of a tall FP64 block with replacement of numerically
// dependent columns by fresh Philox vectors -- the orthonormalisation step of the
// subspace iteration on small Gram matrices (the reference's QR: geqrf + orgqr,
// .../Optimized/kernel.cu:289-346).
#include <algorithm>

#include "ht_common.h"

namespace ht {

constexpr int JMAX = 64;
constexpr int JLD = JMAX + 1;

__global__ __launch_bounds__(256) void jacobi_eig_kernel(const double* __restrict__ A, int n, int lda,
                                                         long strideA, double* __restrict__ w, long stridew,
                                                         double* __restrict__ V, int ldv, long strideV,
                                                         int* __restrict__ info, int max_sweeps, double tol) {
  __shared__ double a[JMAX * JLD];
  __shared__ double v[JMAX * JLD];
  __shared__ double cs[JMAX / 2], sn[JMAX / 2];
  __shared__ int pp[JMAX / 2], qq[JMAX / 2];
  __shared__ double red[8];
  __shared__ int order[JMAX];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int bidx = blockIdx.x;
  A += (long)bidx * strideA;
  w += (long)bidx * stridew;
  V += (long)bidx * strideV;
  const int np = (n + 1) & ~1;

  // Load; the padding index (odd n) gets a large negative diagonal so that it stays
  // decoupled (zero off-diagonals are never rotated) and sorts last.
  double amax = 0.0;
  for (int e = tid; e < n * n; e += 256) {
    const int i = e % n, j = e / n;
    const double x = A[i + (long)lda * j];
    a[i * JLD + j] = x;
    amax = fmax(amax, fabs(x));
  }
  {
    double m = amax;
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (lane == 0) red[wv] = m;
  }
  for (int e = tid; e < np * np; e += 256) {
    const int i = e % np, j = e / np;
    v[i * JLD + j] = (i == j) ? 1.0 : 0.0;
    if (i >= n || j >= n) a[i * JLD + j] = 0.0;
  }
  __syncthreads();
  const double scale = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  if (np != n && tid == 0) a[(np - 1) * JLD + (np - 1)] = -(scale * (double)n + 1.0) * 4.0;
  __syncthreads();

  const int half = np / 2;
  int converged = 0;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int e = tid; e < n * n; e += 256) {
      const int i = e % n, j = e / n;
      const double x = a[i * JLD + j];
      if (i == j) dg += x * x;
      else off += x * x;
    }
    off = wave_sum(off);
    dg = wave_sum(dg);
    __syncthreads();
    if (lane == 0) { red[wv] = off; red[4 + wv] = dg; }
    __syncthreads();
    off = red[0] + red[1] + red[2] + red[3];
    dg = red[4] + red[5] + red[6] + red[7];
    __syncthreads();
    if (off <= tol * tol * dg || off == 0.0) { converged = 1; break; }
    for (int step = 0; step < np - 1; ++step) {
      if (tid < half) {
        int p, q;
        if (tid == 0) { p = np - 1; q = step; }
        else {
          p = (step + tid) % (np - 1);
          q = (step - tid + (np - 1)) % (np - 1);
        }
        if (p > q) { const int t = p; p = q; q = t; }
        const double app = a[p * JLD + p], aqq = a[q * JLD + q], apq = a[p * JLD + q];
        double c = 1.0, s = 0.0;
        if (apq != 0.0 && fabs(apq) > 1e-300) {
          const double tau = (aqq - app) / (2.0 * apq);
          const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          s = t * c;
        }
        pp[tid] = p; qq[tid] = q; cs[tid] = c; sn[tid] = s;
      }
      __syncthreads();
      // A <- A J and V <- V J (columns p, q of every pair).
      for (int e = tid; e < np * half; e += 256) {
        const int r = e % np, i = e / np;
        const int p = pp[i], q = qq[i];
        const double c = cs[i], s = sn[i];
        const double xp = a[r * JLD + p], xq = a[r * JLD + q];
        a[r * JLD + p] = c * xp - s * xq;
        a[r * JLD + q] = s * xp + c * xq;
        const double yp = v[r * JLD + p], yq = v[r * JLD + q];
        v[r * JLD + p] = c * yp - s * yq;
        v[r * JLD + q] = s * yp + c * yq;
      }
      __syncthreads();
      // A <- J^T A (rows p, q).
      for (int e = tid; e < np * half; e += 256) {
        const int r = e % np, i = e / np;
        const int p = pp[i], q = qq[i];
        const double c = cs[i], s = sn[i];
        const double xp = a[p * JLD + r], xq = a[q * JLD + r];
        a[p * JLD + r] = c * xp - s * xq;
        a[q * JLD + r] = s * xp + c * xq;
      }
      __syncthreads();
      // Exact zeros on the rotated pairs (removes rounding residue).
      if (tid < half) {
        const int p = pp[tid], q = qq[tid];
        if (sn[tid] != 0.0) { a[p * JLD + q] = 0.0; a[q * JLD + p] = 0.0; }
      }
      __syncthreads();
    }
  }
  // Sort descending (n <= 64: one thread).
  if (tid == 0) {
    for (int i = 0; i < np; ++i) order[i] = i;
    for (int i = 0; i < np; ++i) {
      int best = i;
      for (int j = i + 1; j < np; ++j)
        if (a[order[j] * JLD + order[j]] > a[order[best] * JLD + order[best]]) best = j;
      const int t = order[i]; order[i] = order[best]; order[best] = t;
    }
    if (info) info[bidx] = converged ? 0 : 1;
  }
  __syncthreads();
  // Output column j = eigenvector order[j], signed so its largest-magnitude entry is > 0.
  for (int j = wv; j < n; j += 4) {
    const int src = order[j];
    double best = -1.0;
    int bi = 0;
    for (int r = lane; r < n; r += 64) {
      const double x = fabs(v[r * JLD + src]);
      if (x > best) { best = x; bi = r; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const double sg = v[bi * JLD + src] < 0.0 ? -1.0 : 1.0;
    for (int r = lane; r < n; r += 64) V[r + (long)ldv * j] = sg * v[r * JLD + src];
    if (lane == 0) w[j] = a[src * JLD + src];
  }
}

// Contiguous problems (lda = ldv = n) take the 1024-thread solver of the batched eigen
// problems (jacobi_full_kernel: the same sorted, sign-canonical result, its steps spread
// over 16 waves); tolerances below 1e-14, which FP64 rounding of a sweep does not reach
// reliably, are raised to 1e-14.
int jacobi_eig_launch(hipStream_t st, const double* A, int n, int lda, double* w, double* V, int ldv, int* info,
                      int max_sweeps, double tol) {
  if (n < 1 || n > JMAX) return 8;
  if (lda == n && ldv == n)
    return jacobi_full_launch(st, A, (long)n * n, n, w, n, V, (long)n * n, 1, info, max_sweeps,
                              tol < 1e-14 ? 1e-14 : tol);
  hipLaunchKernelGGL(jacobi_eig_kernel, dim3(1), dim3(256), 0, st, A, n, lda, 0L, w, 0L, V, ldv, 0L, info,
                     max_sweeps, tol);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int jacobi_eig_batched_launch(hipStream_t st, const double* A, int n, long strideA, double* w, long stridew,
                              double* V, long strideV, int batch, int* info, int max_sweeps, double tol) {
  if (n < 1 || n > JMAX || batch < 1) return 8;
  hipLaunchKernelGGL(jacobi_eig_kernel, dim3(batch), dim3(256), 0, st, A, n, n, strideA, w, stridew, V, n,
                     strideV, info, max_sweeps, tol);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// ------------------------------------------------------------------ dgemm_small
// C (m x n) = alpha op(A) op(B) + beta C, FP64, column-major. 64 x 64 output tile per
// work-group through LDS in steps of 16 along k; thread (tx, ty) of a 16 x 16 grid owns the
// 4 x 4 outputs (tx + 16 a, ty + 16 b), so C is written with consecutive rows per lane.
// FP64 FMA (same peak as FP64 MFMA on this GPU). Loads follow each operand's contiguous
// dimension: A by rows (TA = 0) or by k (TA = 1), B by k (TB = 0) or by columns (TB = 1).
constexpr int DG_T = 64, DG_K = 16;
// Split-K: slice z = blockIdx.z covers k in [z kchunk, (z+1) kchunk) and writes its tile
// to C + z cstride (kchunk = k, cstride = 0 without a split).
template <int TA, int TB, typename TE = double>
__global__ __launch_bounds__(256) void dgemm_tiled_kernel(int m, int n, int k, double alpha,
                                                          const TE* __restrict__ A, int lda,
                                                          const double* __restrict__ B, int ldb, double beta,
                                                          double* __restrict__ C, int ldc, int kchunk, long cstride) {
  __shared__ double As[DG_K][DG_T + 1];
  __shared__ double Bs[DG_K][DG_T + 1];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const long i0 = (long)blockIdx.x * DG_T, j0 = (long)blockIdx.y * DG_T;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(k, kb + kchunk);
  C += (long)blockIdx.z * cstride;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  // Tiles of the next k step are loaded into registers while this one is multiplied.
  // Every load is issued unconditionally with its index clamped into the matrix (a load
  // behind a divergent branch is waited for at the join); entries outside are stored
  // as zero.
  double ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int ii, kk;
      if (TA == 0) { ii = e % DG_T; kk = e / DG_T; } else { kk = e % DG_K; ii = e / DG_K; }
      const long gi = min(i0 + ii, (long)m - 1);
      const int gk = min(k0 + kk, ke - 1);
      ra[q] = (double)(TA == 0 ? A[gi + (long)lda * gk] : A[gk + (long)lda * gi]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int jj, kk;
      if (TB == 0) { kk = e % DG_K; jj = e / DG_K; } else { jj = e % DG_T; kk = e / DG_T; }
      const long gj = min(j0 + jj, (long)n - 1);
      const int gk = min(k0 + kk, ke - 1);
      rb[q] = TB == 0 ? B[gk + (long)ldb * gj] : B[gj + (long)ldb * gk];
    }
  };
  auto store = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int ii, kk;
      if (TA == 0) { ii = e % DG_T; kk = e / DG_T; } else { kk = e % DG_K; ii = e / DG_K; }
      As[kk][ii] = (i0 + ii < m && k0 + kk < ke) ? ra[q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int jj, kk;
      if (TB == 0) { kk = e % DG_K; jj = e / DG_K; } else { jj = e % DG_T; kk = e / DG_T; }
      Bs[kk][jj] = (j0 + jj < n && k0 + kk < ke) ? rb[q] : 0.0;
    }
  };
  if (kb < ke) load(kb);
  for (int k0 = kb; k0 < ke; k0 += DG_K) {
    __syncthreads();
    store(k0);
    __syncthreads();
    if (k0 + DG_K < ke) load(k0 + DG_K);
#pragma unroll
    for (int kk = 0; kk < DG_K; ++kk) {
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[kk][tx + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[kk][ty + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const long i = i0 + tx + 16 * a, j = j0 + ty + 16 * b;
      if (i < m && j < n) {
        double* c = C + i + (long)ldc * j;
        *c = alpha * acc[a][b] + (beta == 0.0 ? 0.0 : beta * (*c));
      }
    }
}

// C[i + ldc j] = alpha sum_z P[z][i + m j] + beta C (deterministic order over z).
__global__ void dgemm_splitk_reduce_kernel(const double* __restrict__ P, int m, int n, int nz, double alpha,
                                           double beta, double* __restrict__ C, int ldc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * n) return;
  const int i = (int)(e % m), j = (int)(e / m);
  double s = 0.0;
  for (int z = 0; z < nz; ++z) s += P[(long)z * m * n + e];
  double* c = C + i + (long)ldc * j;
  *c = alpha * s + (beta == 0.0 ? 0.0 : beta * (*c));
}

template <int TA, int TB>
static void dgemm_dispatch(hipStream_t st, dim3 grid, int m, int n, int k, double alpha, const double* A, int lda,
                           const double* B, int ldb, double beta, double* C, int ldc, int kchunk, long cstride) {
  hipLaunchKernelGGL((dgemm_tiled_kernel<TA, TB>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta, C,
                     ldc, kchunk, cstride);
}

static void dgemm_any(hipStream_t st, int ta, int tb, dim3 grid, int m, int n, int k, double alpha, const double* A,
                      int lda, const double* B, int ldb, double beta, double* C, int ldc, int kchunk, long cstride) {
  if (!ta && !tb) dgemm_dispatch<0, 0>(st, grid, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kchunk, cstride);
  else if (ta && !tb) dgemm_dispatch<1, 0>(st, grid, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kchunk, cstride);
  else if (!ta && tb) dgemm_dispatch<0, 1>(st, grid, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kchunk, cstride);
  else dgemm_dispatch<1, 1>(st, grid, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kchunk, cstride);
}

int dgemm_small_launch(hipStream_t st, int ta, int tb, int m, int n, int k, double alpha, const double* A, int lda,
                       const double* B, int ldb, double beta, double* C, int ldc) {
  if (m < 1 || n < 1) return 0;
  const dim3 grid((unsigned)((m + DG_T - 1) / DG_T), (unsigned)((n + DG_T - 1) / DG_T), 1);
  if (grid.y > 65535) return 8;
  dgemm_any(st, ta, tb, grid, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, k < 1 ? 1 : k, 0);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// A in FP32 (widened as it is loaded): C = alpha A B + beta C, A m x k column-major.
int dgemm_small_f32a_launch(hipStream_t st, int m, int n, int k, double alpha, const float* A, int lda,
                            const double* B, int ldb, double beta, double* C, int ldc) {
  if (m < 1 || n < 1) return 0;
  const dim3 grid((unsigned)((m + DG_T - 1) / DG_T), (unsigned)((n + DG_T - 1) / DG_T), 1);
  if (grid.y > 65535) return 8;
  hipLaunchKernelGGL((dgemm_tiled_kernel<0, 0, float>), grid, dim3(256), 0, st, m, n, k, alpha, A, lda, B, ldb, beta,
                     C, ldc, k < 1 ? 1 : k, 0L);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int dgemm_splitk_slices(int m, int n, int k) {
  const long tiles = (long)((m + DG_T - 1) / DG_T) * ((n + DG_T - 1) / DG_T);
  long z = 512 / tiles;  // about two work-groups per CU
  const long zk = k / 128;  // at least 128 of k per slice (a b x b Gram of 1024 rows: 8 slices)
  if (z > zk) z = zk;
  if (z > 1024) z = 1024;
  return z < 1 ? 1 : (int)z;
}

int dgemm_splitk_launch(hipStream_t st, int ta, int tb, int m, int n, int k, double alpha, const double* A, int lda,
                        const double* B, int ldb, double beta, double* C, int ldc, double* ws) {
  if (m < 1 || n < 1) return 0;
  const int nz = dgemm_splitk_slices(m, n, k);
  if (nz == 1) return dgemm_small_launch(st, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  int kchunk = (k + nz - 1) / nz;
  kchunk = (kchunk + DG_K - 1) / DG_K * DG_K;
  const int z = (k + kchunk - 1) / kchunk;
  const dim3 grid((unsigned)((m + DG_T - 1) / DG_T), (unsigned)((n + DG_T - 1) / DG_T), (unsigned)z);
  dgemm_any(st, ta, tb, grid, m, n, k, 1.0, A, lda, B, ldb, 0.0, ws, m, kchunk, (long)m * n);
  const long tot = (long)m * n;
  hipLaunchKernelGGL(dgemm_splitk_reduce_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, ws, m, n, z,
                     alpha, beta, C, ldc);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// ------------------------------------------------------------------ chol_rinv
// One pass of shifted CholeskyQR (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa,
// SIAM J. Sci. Comput. 2020). From W = Z^T Z (b x b, b <= 64): on the first pass the
// columns are equilibrated (S = diag(W)^-1/2) and W_s = S W S + sigma I with
// sigma = 11 (m b + b (b + 1)) u tr(W_s), which keeps the Cholesky factor of a numerically
// rank-deficient block defined; later passes use S = I, sigma = 0. Writes
// Rinv = S (chol(W_s)^T)^-1 (upper triangular, column-major) so that Z Rinv has
// orthonormal columns up to the pass's accuracy. info[0] |= 1 on a non-positive pivot.
//
// 256 threads on LDS-resident matrices with one barrier per step (compact loops: the
// kernel runs once, so straight-line code would be fetched once and wait on the
// instruction cache throughout). Factorisation step k reads the pivot, subtracts the
// outer product of the unscaled column k divided by the pivot from the trailing lower
// triangle, and after the barrier scales column k; the next step touches neither column
// k nor what the scaling writes. The inverse T = L^-1 is eliminated the same way, a
// column of L per step with T's rows scaled at the end.
__global__ __launch_bounds__(256) void chol_rinv_kernel(const double* __restrict__ W, int b, long m, int first,
                                                        double* __restrict__ Rinv, int* __restrict__ info) {
  __shared__ double L[JMAX * JLD];
  __shared__ double T[JMAX * JLD];
  __shared__ double sc[JMAX];
  __shared__ double rdg[JMAX];
  __shared__ double shift;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  if (tid < b) {
    const double d = W[tid + (long)b * tid];
    sc[tid] = (first && d > 0.0) ? 1.0 / sqrt(d) : 1.0;
  }
  __syncthreads();
  for (int e = tid; e < b * b; e += 256) {
    const int i = e % b, j = e / b;
    L[i * JLD + j] = sc[i] * W[e] * sc[j];
    T[i * JLD + j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  if (tid < 64) {
    double tr = tid < b ? L[tid * JLD + tid] : 0.0;
    for (int o = 32; o > 0; o >>= 1) tr += __shfl_xor(tr, o, 64);
    if (tid == 0) shift = first ? 11.0 * ((double)m * b + (double)b * (b + 1)) * 1.1102230246251565e-16 * tr : 0.0;
  }
  __syncthreads();
  if (tid < b) L[tid * JLD + tid] += shift;
  __syncthreads();
  for (int k = 0; k < b; ++k) {
    double d = L[k * JLD + k];
    if (!(d > 0.0)) {
      if (tid == 0 && info) info[0] |= 1;
      d = 1e-300;
    }
    const double rdd = 1.0 / d;
    const int R = b - k - 1;
    for (int i = k + 1 + tx; i < b; i += 16) {  // the trailing lower triangle
      const double lik = L[i * JLD + k] * rdd;
      for (int j = k + 1 + ty; j <= i; j += 16) L[i * JLD + j] -= lik * L[j * JLD + k];
    }
    __syncthreads();
    const double piv = sqrt(d);
    if (tid < R) L[(k + 1 + tid) * JLD + k] /= piv;
    if (tid == 0) L[k * JLD + k] = piv;
  }
  __syncthreads();
  if (tid < b) rdg[tid] = 1.0 / L[tid * JLD + tid];
  __syncthreads();
  // T starts as I; step q subtracts L[r][q] (T[q][:] / L[q][q]) from the rows r > q, on
  // the columns j <= q (T is lower triangular); row q is final (before its scaling) once
  // step q - 1 is done
  for (int q = 0; q < b; ++q) {
    const double rq = rdg[q];
    for (int r = q + 1 + tx; r < b; r += 16) {
      const double lrq = L[r * JLD + q] * rq;
      for (int j = ty; j <= q; j += 16) T[r * JLD + j] -= lrq * T[q * JLD + j];
    }
    __syncthreads();
  }
  // Rinv[i, j] = sc[i] (L^-T)[i, j] = sc[i] T[j, i] rdg[j]
  for (int e = tid; e < b * b; e += 256) {
    const int i = e % b, j = e / b;
    Rinv[e] = j < i ? 0.0 : sc[i] * T[j * JLD + i] * rdg[j];
  }
}

// Columns of Z (m x b) whose squared norm W[j][j] is at most 1e-28 of the largest are
// replaced by Philox Gaussian vectors and their row and column of W = Z^T Z recomputed
// (one work-group; a block from an exactly singular operator would otherwise keep a zero
// column through every pass).
__global__ __launch_bounds__(256) void refresh_zero_cols_kernel(double* __restrict__ Z, long m, int b,
                                                                double* __restrict__ W, uint64_t seed,
                                                                uint32_t stream) {
  __shared__ double dmax;
  __shared__ double red[4];
  __shared__ double dg[JMAX];  // the diagonal of W (a refresh leaves the others' in place)
  const int tid = threadIdx.x;
  if (tid < b) dg[tid] = W[tid + (long)b * tid];
  __syncthreads();
  if (tid == 0) {
    double d = 0.0;
    for (int i = 0; i < b; ++i) d = fmax(d, dg[i]);
    dmax = d;
  }
  __syncthreads();
  for (int j = 0; j < b; ++j) {
    if (dg[j] > 1e-28 * dmax && dmax > 0.0) continue;  // uniform across the block
    for (long r = tid; r < m; r += 256) Z[r + m * j] = philox_normal(seed, stream, (uint64_t)j * m + r);
    __syncthreads();
    for (int i = 0; i < b; ++i) {
      double s = 0.0;
      for (long r = tid; r < m; r += 256) s += Z[r + m * i] * Z[r + m * j];
      s = wave_sum(s);
      if ((tid & 63) == 0) red[tid >> 6] = s;
      __syncthreads();
      if (tid == 0) {
        const double v = red[0] + red[1] + red[2] + red[3];
        W[i + (long)b * j] = v;
        W[j + (long)b * i] = v;
      }
      __syncthreads();
    }
  }
}

int refresh_zero_cols_launch(hipStream_t st, double* Z, long m, int b, double* W, uint64_t seed, uint32_t stream) {
  if (b < 1 || b > JMAX) return 8;
  hipLaunchKernelGGL(refresh_zero_cols_kernel, dim3(1), dim3(256), 0, st, Z, m, b, W, seed, stream);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int chol_rinv_launch(hipStream_t st, const double* W, int b, long m, int first, double* Rinv, int* info) {
  if (b < 1 || b > JMAX || m < b) return 8;
  hipLaunchKernelGGL(chol_rinv_kernel, dim3(1), dim3(256), 0, st, W, b, m, first, Rinv, info);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// ------------------------------------------------------------------ cholqr
// One work-group. Z (m x b, FP64, ld m) is orthonormalised in place by two Cholesky-QR
// passes after scaling its columns to unit norm; a column whose Cholesky pivot falls
// below 1e-13 of the largest diagonal is replaced by a Philox Gaussian vector and the
// passes restart. W (b x b) is scratch; R (b x b) receives the product of the two
// triangular factors and the column scaling (Z_in = Z_out R).
__global__ __launch_bounds__(256) void cholqr_kernel(double* __restrict__ Z, int m, int b, double* __restrict__ W,
                                                     double* __restrict__ Rout, uint64_t seed, uint32_t stream,
                                                     int* __restrict__ info) {
  __shared__ double R[JMAX * JLD];
  __shared__ double Racc[JMAX * JLD];
  __shared__ double cn[JMAX];
  __shared__ int bad;
  __shared__ double dmax;
  const int tid = threadIdx.x;
  int passes = 0, restarts = 0;
  while (passes < 2 && restarts < 4 * JMAX) {
    __syncthreads();
    if (passes == 0) {
      // Equilibrate: unit columns (the span is unchanged). A block G Q whose columns carry
      // a spread spectrum (norms th_1 .. th_b) is then well conditioned, which keeps the
      // Cholesky factor, and the small directions, accurate. Racc starts as diag(norms).
      if (tid < b) {
        double s = 0.0;
        for (int r = 0; r < m; ++r) s += Z[r + (long)m * tid] * Z[r + (long)m * tid];
        cn[tid] = sqrt(s);
      }
      __syncthreads();
      for (int e = tid; e < m * b; e += 256) {
        const int j = e / m;
        if (cn[j] > 0.0) Z[(long)e] /= cn[j];
      }
      for (int e = tid; e < b * b; e += 256) {
        const int i = e % b, j = e / b;
        Racc[i * JLD + j] = (i == j) ? (cn[i] > 0.0 ? cn[i] : 1.0) : 0.0;
      }
      __syncthreads();
    }
    // W = Z^T Z (upper), one thread per (i <= j).
    for (int e = tid; e < b * b; e += 256) {
      const int i = e % b, j = e / b;
      if (i > j) continue;
      double s = 0.0;
      for (int r = 0; r < m; ++r) s += Z[r + (long)m * i] * Z[r + (long)m * j];
      R[i * JLD + j] = s;
      R[j * JLD + i] = s;
    }
    if (tid == 0) bad = -1;
    __syncthreads();
    if (tid == 0) {
      double d = 0.0;
      for (int i = 0; i < b; ++i) d = fmax(d, R[i * JLD + i]);
      dmax = d;
    }
    __syncthreads();
    // Right-looking Cholesky W = R^T R in LDS.
    for (int j = 0; j < b; ++j) {
      if (tid == 0) {
        const double d = R[j * JLD + j];
        if (!(d > 1e-13 * dmax) || !(dmax > 0.0)) bad = j;
        else R[j * JLD + j] = sqrt(d);
      }
      __syncthreads();
      if (bad >= 0) break;
      const double rjj = R[j * JLD + j];
      for (int c = j + 1 + tid; c < b; c += 256) R[j * JLD + c] /= rjj;
      __syncthreads();
      for (int e = tid; e < (b - j - 1) * (b - j - 1); e += 256) {
        const int c1 = j + 1 + e % (b - j - 1), c2 = j + 1 + e / (b - j - 1);
        if (c1 <= c2) R[c1 * JLD + c2] -= R[j * JLD + c1] * R[j * JLD + c2];
      }
      __syncthreads();
    }
    if (bad >= 0) {
      const int j = bad;
      for (int r = tid; r < m; r += 256)
        Z[r + (long)m * j] = philox_normal(seed, stream, (uint64_t)(restarts + 1) * (uint64_t)m * JMAX + (uint64_t)j * m + r);
      ++restarts;
      passes = 0;
      __syncthreads();
      continue;
    }
    // Z <- Z R^{-1}: each thread solves rows independently (forward substitution).
    for (int r = tid; r < m; r += 256) {
      double x[JMAX];
      for (int j = 0; j < b; ++j) {
        double s = Z[r + (long)m * j];
        for (int l = 0; l < j; ++l) s -= x[l] * R[l * JLD + j];
        x[j] = s / R[j * JLD + j];
      }
      for (int j = 0; j < b; ++j) Z[r + (long)m * j] = x[j];
    }
    __syncthreads();
    // Racc <- R * Racc (upper triangular).
    for (int e = tid; e < b * b; e += 256) {
      const int i = e % b, j = e / b;
      double s = 0.0;
      if (i <= j)
        for (int l = i; l <= j; ++l) s += R[i * JLD + l] * Racc[l * JLD + j];
      W[i + (long)b * j] = s;
    }
    __syncthreads();
    for (int e = tid; e < b * b; e += 256) Racc[(e % b) * JLD + e / b] = W[(e % b) + (long)b * (e / b)];
    ++passes;
  }
  __syncthreads();
  if (Rout)
    for (int e = tid; e < b * b; e += 256) Rout[e] = Racc[(e % b) * JLD + e / b];
  if (tid == 0 && info) *info = (passes == 2) ? restarts : -1;
}

int cholqr_launch(hipStream_t st, double* Z, int m, int b, double* W, double* R, uint64_t seed, uint32_t stream,
                  int* info) {
  if (b < 1 || b > JMAX || m < b) return 8;
  hipLaunchKernelGGL(cholqr_kernel, dim3(1), dim3(256), 0, st, Z, m, b, W, R, seed, stream, info);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// ------------------------------------------------------------------ elementwise helpers
__global__ void cast_f64_f32_kernel(const double* __restrict__ a, float* __restrict__ b, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (float)a[i];
}
__global__ void cast_f32_f64_kernel(const float* __restrict__ a, double* __restrict__ b, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (double)a[i];
}
__global__ void fill_normals_f64_kernel(double* __restrict__ a, long n, uint64_t seed, uint32_t stream) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    a[i] = philox_normal(seed, stream, (uint64_t)i);
}
__global__ void fill_normals_f32_kernel(float* __restrict__ a, long n, uint64_t seed, uint32_t stream) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    a[i] = (float)philox_normal(seed, stream, (uint64_t)i);
}
__global__ void copy_cols_f64_kernel(const double* __restrict__ a, int lda, double* __restrict__ b, int ldb, int m,
                                     int n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * n) return;
  const int i = (int)(e % m), j = (int)(e / m);
  b[i + (long)ldb * j] = a[i + (long)lda * j];
}
__global__ void canon_signs_kernel(double* __restrict__ U, int m, int ldu) {
  __shared__ double bv[4];
  __shared__ int bix[4];
  const int j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double best = -1.0;
  int bi = 0;
  for (int r = tid; r < m; r += 256) {
    const double x = fabs(U[r + (long)ldu * j]);
    if (x > best) { best = x; bi = r; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
  }
  if (lane == 0) { bv[wv] = best; bix[wv] = bi; }
  __syncthreads();
  if (tid == 0) {
    for (int k = 1; k < 4; ++k)
      if (bv[k] > bv[0] || (bv[k] == bv[0] && bix[k] < bix[0])) { bv[0] = bv[k]; bix[0] = bix[k]; }
  }
  __syncthreads();
  const double sg = U[bix[0] + (long)ldu * j] < 0.0 ? -1.0 : 1.0;
  if (sg < 0.0)
    for (int r = tid; r < m; r += 256) U[r + (long)ldu * j] = -U[r + (long)ldu * j];
}
// Canonical signs of tall columns (m > CS_SPLIT rows) in two passes over many blocks: a
// one-block-per-column scan streams the column serially, which for the million-row bases
// of a root child is most of the time of a decomposition. Pass 1: block (j, slice) finds
// the largest |U[r][j]| over its CS_ROWS rows (ties to the smaller row) and records
// (|value|, row, value) in part[j][slice]. Pass 2: block (j, slice) reduces the column's
// records the same way and negates its rows when the recorded value is negative (the
// value is taken from the record, not re-read, so no block reads a row another block is
// negating).
constexpr int CS_ROWS = 8192, CS_SPLIT = 16384;
__global__ __launch_bounds__(256) void canon_max_kernel(const double* __restrict__ U, int m, int ldu,
                                                        double* __restrict__ part, int S) {
  __shared__ double bv[4], bs[4];
  __shared__ int bix[4];
  const int j = blockIdx.x, sl = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int r0 = sl * CS_ROWS, r1 = min(m, r0 + CS_ROWS);
  double best = -1.0, sv = 0.0;
  int bi = 0;
#pragma unroll 8
  for (int r = r0 + tid; r < r1; r += 256) {
    const double x = U[r + (long)ldu * j];
    if (fabs(x) > best) {
      best = fabs(x);
      bi = r;
      sv = x;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64), os = __shfl_xor(sv, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
      sv = os;
    }
  }
  if (lane == 0) {
    bv[wv] = best;
    bix[wv] = bi;
    bs[wv] = sv;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < 4; ++w)
      if (bv[w] > bv[0] || (bv[w] == bv[0] && bix[w] < bix[0])) {
        bv[0] = bv[w];
        bix[0] = bix[w];
        bs[0] = bs[w];
      }
    double* o = part + 3 * ((long)j * S + sl);
    o[0] = bv[0];
    o[1] = (double)bix[0];
    o[2] = bs[0];
  }
}
__global__ __launch_bounds__(256) void canon_flip_kernel(double* __restrict__ U, int m, int ldu,
                                                         const double* __restrict__ part, int S) {
  __shared__ double bv[4], bs[4];
  __shared__ int bix[4];
  const int j = blockIdx.x, sl = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double best = -1.0, sv = 0.0;
  int bi = 0;
  for (int e = tid; e < S; e += 256) {
    const double* o = part + 3 * ((long)j * S + e);
    const int oi = (int)o[1];
    if (o[0] > best || (o[0] == best && oi < bi)) {
      best = o[0];
      bi = oi;
      sv = o[2];
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64), os = __shfl_xor(sv, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
      sv = os;
    }
  }
  if (lane == 0) {
    bv[wv] = best;
    bix[wv] = bi;
    bs[wv] = sv;
  }
  __syncthreads();
  double s0 = bs[0], b0 = bv[0];
  int i0 = bix[0];
  for (int w = 1; w < 4; ++w)
    if (bv[w] > b0 || (bv[w] == b0 && bix[w] < i0)) {
      b0 = bv[w];
      i0 = bix[w];
      s0 = bs[w];
    }
  if (s0 < 0.0) {
    const int r0 = sl * CS_ROWS, r1 = min(m, r0 + CS_ROWS);
    for (int r = r0 + tid; r < r1; r += 256) U[r + (long)ldu * j] = -U[r + (long)ldu * j];
  }
}
template <typename T>
__global__ void sumsq_stage1(const T* __restrict__ a, long n, double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double x = (double)a[i];
    s += x * x;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
__global__ void sum_stage2(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

static unsigned grid_for(long n) {
  long b = (n + 255) / 256;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Several FP64 -> FP32 copies (optionally sqrt(max(x, 0)) first) in one launch: blockIdx.y
// picks the copy.
struct CastTable {
  CastItem it[CAST_MAX];
};
__global__ void cast_multi_kernel(CastTable tab) {
  const CastItem c = tab.it[blockIdx.y];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += (long)gridDim.x * blockDim.x) {
    const double x = c.src[i];
    c.dst[i] = (float)(c.sqrt_clamp ? sqrt(fmax(x, 0.0)) : x);
  }
}
int cast_multi_launch(hipStream_t st, const CastItem* items, int n) {
  for (int s0 = 0; s0 < n; s0 += CAST_MAX) {
    CastTable tab;
    const int cnt = std::min(CAST_MAX, n - s0);
    long mx = 1;
    for (int i = 0; i < cnt; ++i) {
      tab.it[i] = items[s0 + i];
      mx = std::max(mx, tab.it[i].n);
    }
    // up to 8 blocks per CU for the largest item (a million-row basis of a 1024^3 run is
    // 256 MB: a thread loops over it, and 256 blocks kept too few loads in flight)
    hipLaunchKernelGGL(cast_multi_kernel, dim3(std::min<unsigned>(grid_for(mx), 2048), cnt), dim3(256), 0, st, tab);
    if (hipGetLastError() != hipSuccess) return 5;
  }
  return 0;
}
int cast_f64_f32_launch(hipStream_t st, const double* a, float* b, long n) {
  hipLaunchKernelGGL(cast_f64_f32_kernel, dim3(grid_for(n)), dim3(256), 0, st, a, b, n);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
int cast_f32_f64_launch(hipStream_t st, const float* a, double* b, long n) {
  hipLaunchKernelGGL(cast_f32_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, a, b, n);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
int fill_normals_f64_launch(hipStream_t st, double* a, long n, uint64_t seed, uint32_t stream) {
  hipLaunchKernelGGL(fill_normals_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, a, n, seed, stream);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
int fill_normals_f32_launch(hipStream_t st, float* a, long n, uint64_t seed, uint32_t stream) {
  hipLaunchKernelGGL(fill_normals_f32_kernel, dim3(grid_for(n)), dim3(256), 0, st, a, n, seed, stream);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
int copy_cols_f64_launch(hipStream_t st, const double* a, int lda, double* b, int ldb, int m, int n) {
  const long tot = (long)m * n;
  if (tot <= 0) return 0;
  hipLaunchKernelGGL(copy_cols_f64_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, a, lda, b, ldb, m, n);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
size_t canon_signs_ws_doubles(int m, int k) {
  return m > CS_SPLIT ? (size_t)3 * k * ((m + CS_ROWS - 1) / CS_ROWS) : 0;
}
int canon_signs_launch(hipStream_t st, double* U, int m, int k, int ldu, double* ws) {
  if (k < 1 || m < 1) return 0;
  if (m <= CS_SPLIT) {
    hipLaunchKernelGGL(canon_signs_kernel, dim3(k), dim3(256), 0, st, U, m, ldu);
  } else {
    if (!ws) return 8;
    const int S = (m + CS_ROWS - 1) / CS_ROWS;
    hipLaunchKernelGGL(canon_max_kernel, dim3(k, S), dim3(256), 0, st, U, m, ldu, ws, S);
    hipLaunchKernelGGL(canon_flip_kernel, dim3(k, S), dim3(256), 0, st, U, m, ldu, ws, S);
  }
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
// out must hold 1 + 1024 doubles (part scratch after the result).
int sumsq_f64_launch(hipStream_t st, const double* a, long n, double* out) {
  unsigned g = grid_for(n);
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(sumsq_stage1<double>, dim3(g), dim3(256), 0, st, a, n, out + 1);
  hipLaunchKernelGGL(sum_stage2, dim3(1), dim3(256), 0, st, out + 1, (int)g, out);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
int sumsq_f32_launch(hipStream_t st, const float* a, long n, double* out) {
  unsigned g = grid_for(n);
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(sumsq_stage1<float>, dim3(g), dim3(256), 0, st, a, n, out + 1);
  hipLaunchKernelGGL(sum_stage2, dim3(1), dim3(256), 0, st, out + 1, (int)g, out);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// ------------------------------------------------------------------ Ritz residuals
// out[i] = ||GU[:,i] - w[i] U[:,i]||_2 for i < k (one block per column).
__global__ void ritz_resid_kernel(const double* __restrict__ GU, const double* __restrict__ U,
                                  const double* __restrict__ w, int n, double* __restrict__ out) {
  __shared__ double red[4];
  const int i = blockIdx.x;
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double d = GU[r + (long)n * i] - w[i] * U[r + (long)n * i];
    s += d * d;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[i] = sqrt(red[0] + red[1] + red[2] + red[3]);
}
int ritz_resid_launch(hipStream_t st, const double* GU, const double* U, const double* w, int n, int k,
                      double* out) {
  if (k < 1) return 0;
  hipLaunchKernelGGL(ritz_resid_kernel, dim3(k), dim3(256), 0, st, GU, U, w, n, out);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// ------------------------------------------------------------------ FP64 Gram of a view
// G(i,j) = sum_{p,s} x[p + P i + P M s] x[p + P j + P M s] for small tensors (M <= 4096).
// One 16 x 16 tile (I <= J) per block of 4 waves on FP64 MFMA 16x16x4: wave w takes the
// reduction steps k0 = 4 (w + 4 u), eight per batch with all sixteen loads issued before
// the products (indices past K are clamped and their values zeroed); the four partial
// tiles are added through LDS and the tile is written with its mirror.
__global__ __launch_bounds__(256) void gram_pms_f64_kernel(const double* __restrict__ x, PMS v, double* __restrict__ G,
                                                           float* __restrict__ G32) {
  const int I = blockIdx.x, J = blockIdx.y;
  if (I > J) return;
  __shared__ double part[4][256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i16 = lane & 15, kk = lane >> 4;
  const long M = v.M, P = v.P, K = v.P * v.S;
  const long i = 16L * I + i16, j = 16L * J + i16;
  const bool iok = i < M, jok = j < M;
  const long ic = iok ? i : M - 1, jc = jok ? j : M - 1;
  auto at = [&](long row, long k) {  // x[p + P row + P M s], k = p + P s
    const long s = k / P, q = k - s * P;
    return x[q + P * row + P * M * s];
  };
  f64x4 acc = {};
  constexpr int UB = 8;
  for (long k0 = 4L * w; k0 < K; k0 += 16L * UB) {
    double a[UB], b[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      const long k = k0 + 16L * u + kk;
      const long kc = k < K ? k : K - 1;
      a[u] = at(ic, kc);
      b[u] = at(jc, kc);
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      const bool ok = k0 + 16L * u + kk < K;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ok && iok ? a[u] : 0.0, ok && jok ? b[u] : 0.0, acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) part[w][acc16_row_f64(e, lane) * 16 + i16] = acc[e];
  __syncthreads();
  const int r = tid >> 4, c = tid & 15;
  const long gi = 16L * I + r, gj = 16L * J + c;
  if (gi < M && gj < M) {
    const double sum = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
    G[gi + M * gj] = sum;
    G[gj + M * gi] = sum;
    if (G32) {
      G32[gi + M * gj] = (float)sum;
      G32[gj + M * gi] = (float)sum;
    }
  }
}
int gram_pms_f64_launch(hipStream_t st, const double* x, PMS v, double* G, float* G32) {
  const unsigned nt = (unsigned)((v.M + 15) / 16);
  hipLaunchKernelGGL(gram_pms_f64_kernel, dim3(nt, nt), dim3(256), 0, st, x, v, G, G32);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// ------------------------------------------------------------------ misc small kernels
__global__ void sum_pairs_kernel(const double* __restrict__ acc, int n, double* __restrict__ out) {
  __shared__ double r0[4], r1[4];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) { a += acc[2 * i]; b += acc[2 * i + 1]; }
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) { r0[threadIdx.x >> 6] = a; r1[threadIdx.x >> 6] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = r0[0] + r0[1] + r0[2] + r0[3];
    out[1] = r1[0] + r1[1] + r1[2] + r1[3];
  }
}
int sum_strided_launch(hipStream_t st, const double* acc, int n, double* out) {
  hipLaunchKernelGGL(sum_pairs_kernel, dim3(1), dim3(256), 0, st, acc, n, out);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
__global__ void sqrt_clamp_kernel(const double* __restrict__ w, int k, double* __restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) s[i] = sqrt(fmax(w[i], 0.0));
}
int sqrt_clamp_launch(hipStream_t st, const double* w, int k, double* s) {
  if (k < 1) return 0;
  hipLaunchKernelGGL(sqrt_clamp_kernel, dim3((k + 255) / 256), dim3(256), 0, st, w, k, s);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
// U[:, j] = A[:, j] / sqrt(max(w_j, 0)) (zero where that is 0) with the canonical sign (its
// largest-magnitude entry, the first on ties, made positive); A[r, j] is read at
// in[r rs + j cs], which may alias U. One block per column, m <= 4096: the column is
// held in registers before anything is written.
__global__ __launch_bounds__(256) void cols_unit_sign_kernel(const double* in, long rs, long cs, double* U, int m,
                                                             int ldu, const double* __restrict__ w, int k,
                                                             int* __restrict__ ok) {
  __shared__ double bv[4], bs[4];
  __shared__ int bix[4];
  const int j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (ok && j == 0 && tid == 0) *ok = (w[0] > 0.0 && w[k - 1] > 1e-24 * w[0]) ? 1 : 0;
  const double sw = sqrt(fmax(w[j], 0.0));
  double x[16];
  double best = -1.0, sgn = 1.0;
  int bi = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = tid + 256 * i;
    x[i] = 0.0;
    if (r < m) {
      const double a = in[r * rs + j * cs];
      x[i] = sw > 0.0 ? a / sw : 0.0;
      if (fabs(x[i]) > best) {
        best = fabs(x[i]);
        sgn = x[i];
        bi = r;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64), os = __shfl_xor(sgn, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      sgn = os;
      bi = oi;
    }
  }
  if (lane == 0) {
    bv[wv] = best;
    bs[wv] = sgn;
    bix[wv] = bi;
  }
  __syncthreads();
  double b0 = bv[0], s0 = bs[0];
  int i0 = bix[0];
  for (int k = 1; k < 4; ++k)
    if (bv[k] > b0 || (bv[k] == b0 && bix[k] < i0)) {
      b0 = bv[k];
      s0 = bs[k];
      i0 = bix[k];
    }
  const double sg = s0 < 0.0 ? -1.0 : 1.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = tid + 256 * i;
    if (r < m) U[r + (long)ldu * j] = sg * x[i];
  }
}
int cols_unit_sign_launch(hipStream_t st, const double* in, long rs, long cs, double* U, int m, int k, int ldu,
                          const double* w, int* ok) {
  if (m < 1 || m > 4096 || k < 1) return 8;
  hipLaunchKernelGGL(cols_unit_sign_kernel, dim3(k), dim3(256), 0, st, in, rs, cs, U, m, ldu, w, k, ok);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
__global__ void scale_cols_inv_kernel(double* __restrict__ A, int m, int n, const double* __restrict__ s) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * n) return;
  const int j = (int)(e / m);
  A[e] = s[j] > 0.0 ? A[e] / s[j] : 0.0;
}
int scale_cols_inv_launch(hipStream_t st, double* A, int m, int n, const double* s) {
  const long tot = (long)m * n;
  if (tot <= 0) return 0;
  hipLaunchKernelGGL(scale_cols_inv_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, A, m, n, s);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

}  // namespace ht
