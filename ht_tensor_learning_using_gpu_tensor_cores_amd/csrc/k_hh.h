// Block Householder QR device routines shared by TSQR and the small solvers.
// Thread layout: 256 threads; lane = column c (< 64), wave q = quarter of the rows.
#pragma once
#include "ht_common.h"

namespace ht {

constexpr int TQ_ROWS = 256;
constexpr int TQ_B = 64;

// Householder QR of the rows x b matrix in LDS A[r*TQ_B + c]; on return the upper
// triangle holds R, the strict lower part the reflector tails (unit diagonal implied).
__device__ void block_householder(double* A, double* tau, double* red, int rows, int b, int ld = TQ_B) {
  const int tid = threadIdx.x, c = tid & 63, q = tid >> 6;
  const int ng = blockDim.x >> 6;  // row groups (red holds ng x TQ_B partials)
  const int rq = (rows + ng - 1) / ng;
  const int r_lo = q * rq, r_hi = min(rows, r_lo + rq);
  __shared__ double sh_beta, sh_scale, sh_tau;
  for (int j = 0; j < b && j < rows; ++j) {
    // norm^2 of column j from row j down
    if (c == j) {
      double s = 0.0;
      for (int r = max(r_lo, j + 1); r < r_hi; ++r) s += A[r * ld + j] * A[r * ld + j];
      red[q * TQ_B + j] = s;
    }
    __syncthreads();
    if (tid == 0) {
      double sub = 0.0;
      for (int g = 0; g < ng; ++g) sub += red[g * TQ_B + j];
      const double alpha = A[j * ld + j];
      if (sub == 0.0) {
        sh_tau = 0.0; sh_beta = alpha; sh_scale = 0.0;
      } else {
        const double nrm = sqrt(alpha * alpha + sub);
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        sh_tau = (beta - alpha) / beta;
        sh_scale = 1.0 / (alpha - beta);
        sh_beta = beta;
      }
      tau[j] = sh_tau;
    }
    __syncthreads();
    const double tj = sh_tau;
    if (c == j) {
      for (int r = max(r_lo, j + 1); r < r_hi; ++r) A[r * ld + j] *= sh_scale;
    }
    __syncthreads();
    if (tid == 0) A[j * ld + j] = sh_beta;
    // w_c = sum_{r>=j} v_r A[r][c] for c > j
    if (c > j && c < b && tj != 0.0) {
      double s = 0.0;
      for (int r = max(r_lo, j); r < r_hi; ++r) {
        const double vr = (r == j) ? 1.0 : A[r * ld + j];
        s += vr * A[r * ld + c];
      }
      red[q * TQ_B + c] = s;
    }
    __syncthreads();
    if (c > j && c < b && tj != 0.0) {
      double wsum = 0.0;
      for (int g = 0; g < ng; ++g) wsum += red[g * TQ_B + c];
      const double wc = tj * wsum;
      for (int r = max(r_lo, j); r < r_hi; ++r) {
        const double vr = (r == j) ? 1.0 : A[r * ld + j];
        A[r * ld + c] -= vr * wc;
      }
    }
    __syncthreads();
  }
}

// Reflector accessors: tree nodes keep FP64 reflectors in global memory (row-major
// rows x TQ_B), leaves keep them in place in W (FP32, column-major, ld m).
struct VGlobal64 {
  const double* V;
  __device__ double operator()(int r, int j) const { return V[r * TQ_B + j]; }
};
struct VLds {
  const double* V;
  int ld;
  __device__ double operator()(int r, int j) const { return V[r * ld + j]; }
};
struct VLeaf32 {
  const float* W;
  long r0, m;
  __device__ double operator()(int r, int j) const { return (double)W[r0 + r + m * j]; }
};

// M (rows x b in LDS) <- H_0 ... H_{b-1} M with the reflectors given by V.
template <class VA>
__device__ void apply_q(VA V, const double* tau, double* M, double* red, int rows, int b, int ld = TQ_B) {
  const int tid = threadIdx.x, c = tid & 63, q = tid >> 6;
  const int ng = blockDim.x >> 6;
  const int rq = (rows + ng - 1) / ng;
  const int r_lo = q * rq, r_hi = min(rows, r_lo + rq);
  for (int j = min(b, rows) - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj == 0.0) continue;  // uniform across the block
    if (c < b) {
      double s = 0.0;
      for (int r = max(r_lo, j); r < r_hi; ++r) {
        const double vr = (r == j) ? 1.0 : V(r, j);
        s += vr * M[r * ld + c];
      }
      red[q * TQ_B + c] = s;
    }
    __syncthreads();
    if (c < b) {
      double wsum = 0.0;
      for (int g = 0; g < ng; ++g) wsum += red[g * TQ_B + c];
      const double wc = tj * wsum;
      for (int r = max(r_lo, j); r < r_hi; ++r) {
        const double vr = (r == j) ? 1.0 : V(r, j);
        M[r * ld + c] -= vr * wc;
      }
    }
    __syncthreads();
  }
}

}  // namespace ht
