// Block Householder QR device routines of the TSQR tree (256 threads = 4 waves).
// The block is rows x b in LDS, row-major with an odd stride ld >= b. Column work is
// lane = column; the norm of the pivot column is lane = row with a wave reduction, and the
// row loops keep four independent accumulators so that LDS reads overlap.
#pragma once
#include "ht_common.h"

namespace ht {

constexpr int TQ_B = 64;           // widest block
constexpr int TQ_FANROWS = 4;      // a node holds at most 4 b rows

__device__ __forceinline__ int tq_ld(int b) { return b | 1; }

// Householder QR of A (rows x b, ld); on return the upper triangle holds R and the strict
// lower part the reflector tails (unit diagonal implied); tau[j] the coefficients.
// red holds 4 x TQ_B doubles, nrm 4 doubles.
__device__ void block_householder(double* A, double* tau, double* red, double* nrm, int rows, int b, int ld) {
  const int tid = threadIdx.x, c = tid & 63, q = tid >> 6;
  const int ng = blockDim.x >> 6;
  const int rq = (rows + ng - 1) / ng;  // <= 64 rows per wave
  const int r_lo = q * rq, r_hi = min(rows, r_lo + rq);
  __shared__ double sh_beta, sh_scale, sh_tau;
  for (int j = 0; j < b && j < rows; ++j) {
    // ||A[j+1:, j]||^2: one row per lane, then across the waves
    {
      double s = 0.0;
      for (int r = max(r_lo, j + 1) + c; r < r_hi; r += 64) {
        const double x = A[r * ld + j];
        s += x * x;
      }
      s = wave_sum(s);
      if (c == 0) nrm[q] = s;
    }
    __syncthreads();
    if (tid == 0) {
      double sub = 0.0;
      for (int g = 0; g < ng; ++g) sub += nrm[g];
      const double alpha = A[j * ld + j];
      if (sub == 0.0) {
        sh_tau = 0.0; sh_beta = alpha; sh_scale = 0.0;
      } else {
        const double nr = sqrt(alpha * alpha + sub);
        const double beta = alpha >= 0.0 ? -nr : nr;
        sh_tau = (beta - alpha) / beta;
        sh_scale = 1.0 / (alpha - beta);
        sh_beta = beta;
      }
      tau[j] = sh_tau;
    }
    __syncthreads();
    const double tj = sh_tau;
    for (int r = max(r_lo, j + 1) + c; r < r_hi; r += 64) A[r * ld + j] *= sh_scale;
    __syncthreads();
    if (tid == 0) A[j * ld + j] = sh_beta;
    // w_c = sum_{r >= j} v_r A[r][c] for c > j (v_j = 1)
    if (c > j && c < b && tj != 0.0) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      int r = max(r_lo, j);
      if (r == j && r < r_hi) { s0 = A[j * ld + c]; ++r; }
      for (; r + 3 < r_hi; r += 4) {
        s0 += A[r * ld + j] * A[r * ld + c];
        s1 += A[(r + 1) * ld + j] * A[(r + 1) * ld + c];
        s2 += A[(r + 2) * ld + j] * A[(r + 2) * ld + c];
        s3 += A[(r + 3) * ld + j] * A[(r + 3) * ld + c];
      }
      for (; r < r_hi; ++r) s0 += A[r * ld + j] * A[r * ld + c];
      red[q * TQ_B + c] = (s0 + s1) + (s2 + s3);
    }
    __syncthreads();
    if (c > j && c < b && tj != 0.0) {
      double wsum = 0.0;
      for (int g = 0; g < ng; ++g) wsum += red[g * TQ_B + c];
      const double wc = tj * wsum;
      int r = max(r_lo, j);
      if (r == j && r < r_hi) { A[j * ld + c] -= wc; ++r; }
      for (; r + 3 < r_hi; r += 4) {
        const double v0 = A[r * ld + j], v1 = A[(r + 1) * ld + j], v2 = A[(r + 2) * ld + j], v3 = A[(r + 3) * ld + j];
        const double a0 = A[r * ld + c], a1 = A[(r + 1) * ld + c], a2 = A[(r + 2) * ld + c], a3 = A[(r + 3) * ld + c];
        A[r * ld + c] = a0 - v0 * wc;
        A[(r + 1) * ld + c] = a1 - v1 * wc;
        A[(r + 2) * ld + c] = a2 - v2 * wc;
        A[(r + 3) * ld + c] = a3 - v3 * wc;
      }
      for (; r < r_hi; ++r) A[r * ld + c] -= A[r * ld + j] * wc;
    }
    __syncthreads();
  }
}

// Reflector accessors: tree nodes keep FP64 reflectors in global memory (row-major,
// stride ldv), leaves keep them in place in W (FP32, column-major, ld m).
struct VGlobal64 {
  const double* V;
  int ldv;
  __device__ double operator()(int r, int j) const { return V[r * ldv + j]; }
};
struct VLeaf32 {
  const float* W;
  long r0, m;
  __device__ double operator()(int r, int j) const { return (double)W[r0 + r + m * j]; }
};

// M (rows x b in LDS, ld) <- H_0 ... H_{b-1} M with the reflectors given by V.
template <class VA>
__device__ void apply_q(VA V, const double* tau, double* M, double* red, int rows, int b, int ld) {
  const int tid = threadIdx.x, c = tid & 63, q = tid >> 6;
  const int ng = blockDim.x >> 6;
  const int rq = (rows + ng - 1) / ng;
  const int r_lo = q * rq, r_hi = min(rows, r_lo + rq);
  for (int j = min(b, rows) - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj == 0.0) continue;  // uniform across the block
    if (c < b) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      int r = max(r_lo, j);
      if (r == j && r < r_hi) { s0 = M[j * ld + c]; ++r; }
      for (; r + 3 < r_hi; r += 4) {
        s0 += V(r, j) * M[r * ld + c];
        s1 += V(r + 1, j) * M[(r + 1) * ld + c];
        s2 += V(r + 2, j) * M[(r + 2) * ld + c];
        s3 += V(r + 3, j) * M[(r + 3) * ld + c];
      }
      for (; r < r_hi; ++r) s0 += V(r, j) * M[r * ld + c];
      red[q * TQ_B + c] = (s0 + s1) + (s2 + s3);
    }
    __syncthreads();
    if (c < b) {
      double wsum = 0.0;
      for (int g = 0; g < ng; ++g) wsum += red[g * TQ_B + c];
      const double wc = tj * wsum;
      int r = max(r_lo, j);
      if (r == j && r < r_hi) { M[j * ld + c] -= wc; ++r; }
      for (; r + 3 < r_hi; r += 4) {
        const double v0 = V(r, j), v1 = V(r + 1, j), v2 = V(r + 2, j), v3 = V(r + 3, j);
        M[r * ld + c] -= v0 * wc;
        M[(r + 1) * ld + c] -= v1 * wc;
        M[(r + 2) * ld + c] -= v2 * wc;
        M[(r + 3) * ld + c] -= v3 * wc;
      }
      for (; r < r_hi; ++r) M[r * ld + c] -= V(r, j) * wc;
    }
    __syncthreads();
  }
}

}  // namespace ht
