// Batched small symmetric eigen solvers of the HT engine (FP64, n <= 256, block <= 64).
//
// The node problems are small: leaf Grams (n x n, n = mode size) and the Grams of the
// leaf core's matricisations (k_l k_r x k_l k_r). Their top-k eigenpairs replace the
// reference's cusolverDnSsyevd / Sgesvd calls (third_order_tensor_decompsition/CUDA/
// Optimized/kernel.cu:682-718; fourth_order_tensor_decomposition/single_GPU/CUDA/
// Optimized/kernel.cu:267-287, :457-481). One round of the solver is
//   power_kernel  Y = G^q W          column blocks of 8 in parallel, each power followed
//                                    by a column normalisation (the span is unchanged)
//   hhqr_kernel   Q = orth(Y)        Householder QR in FP64 (stable for any conditioning,
//                                    rank-deficient blocks included)
//   power_kernel  Z = G Q            (q = 1, no normalisation)
//   rr_kernel     T = Q^T Z = S Th S^T by parallel Jacobi; U = Q S, the Ritz residuals
//                 ||G u_i - th_i u_i|| and a converged flag per problem
// and the host only reads the flags between rounds.
#include "k_hh.h"

namespace ht {

constexpr int SJ = 64;
constexpr int SJLD = SJ + 1;

// Cyclic parallel Jacobi on the np x np symmetric matrix a (LDS, ld SJLD), accumulating
// the rotations into v. Round-robin pair ordering; each step rotates np/2 disjoint pairs
// and updates every 2x2 block (rows of pair i, columns of pair j) as R_i^T A_ij R_j in a
// single phase. Returns 1 when the off-diagonal mass fell below tol^2 * diagonal mass.
__device__ int jacobi_sweeps(double* a, double* v, int n, int np, int max_sweeps, double tol) {
  __shared__ double cs[SJ / 2], sn[SJ / 2];
  __shared__ int pp[SJ / 2], qq[SJ / 2];
  __shared__ double red[16];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nt = blockDim.x;
  const int half = np / 2;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int e = tid; e < n * n; e += nt) {
      const int i = e % n, j = e / n;
      const double x = a[i * SJLD + j];
      if (i == j) dg += x * x;
      else off += x * x;
    }
    off = wave_sum(off);
    dg = wave_sum(dg);
    __syncthreads();
    if (lane == 0) { red[wv] = off; red[8 + wv] = dg; }
    __syncthreads();
    off = 0.0;
    dg = 0.0;
    for (int w = 0; w < nt / 64; ++w) { off += red[w]; dg += red[8 + w]; }
    __syncthreads();
    if (off <= tol * tol * dg || off == 0.0) return 1;
    for (int step = 0; step < np - 1; ++step) {
      if (tid < half) {
        int p, q;
        if (tid == 0) { p = np - 1; q = step; }
        else {
          p = (step + tid) % (np - 1);
          q = (step - tid + (np - 1)) % (np - 1);
        }
        if (p > q) { const int t = p; p = q; q = t; }
        const double app = a[p * SJLD + p], aqq = a[q * SJLD + q], apq = a[p * SJLD + q];
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
      for (int e = tid; e < half * half; e += nt) {
        const int bi = e % half, bj = e / half;
        const int p = pp[bi], q = qq[bi], P = pp[bj], Q = qq[bj];
        const double c = cs[bi], s = sn[bi], C = cs[bj], S = sn[bj];
        const double apP = a[p * SJLD + P], apQ = a[p * SJLD + Q];
        const double aqP = a[q * SJLD + P], aqQ = a[q * SJLD + Q];
        // right: columns (P, Q); left: rows (p, q)
        const double xpP = C * apP - S * apQ, xpQ = S * apP + C * apQ;
        const double xqP = C * aqP - S * aqQ, xqQ = S * aqP + C * aqQ;
        double ypP = c * xpP - s * xqP, yqP = s * xpP + c * xqP;
        double ypQ = c * xpQ - s * xqQ, yqQ = s * xpQ + c * xqQ;
        if (bi == bj && s != 0.0) { ypQ = 0.0; yqP = 0.0; }
        a[p * SJLD + P] = ypP; a[p * SJLD + Q] = ypQ;
        a[q * SJLD + P] = yqP; a[q * SJLD + Q] = yqQ;
      }
      for (int e = tid; e < np * half; e += nt) {
        const int r = e % np, j = e / np;
        const int P = pp[j], Q = qq[j];
        const double C = cs[j], S = sn[j];
        const double vP = v[r * SJLD + P], vQ = v[r * SJLD + Q];
        v[r * SJLD + P] = C * vP - S * vQ;
        v[r * SJLD + Q] = S * vP + C * vQ;
      }
      __syncthreads();
    }
  }
  return 0;
}

// Load an n x n matrix into a (padded to even np with a decoupled, large negative
// diagonal entry that sorts last), set v = I.
__device__ void jacobi_load(double* a, double* v, const double* A, int lda, int n, int np, bool symmetrise) {
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double amax = 0.0;
  for (int e = tid; e < np * np; e += blockDim.x) {
    const int i = e % np, j = e / np;
    double x = 0.0;
    if (i < n && j < n) {
      x = A[i + (long)lda * j];
      if (symmetrise) x = 0.5 * (x + A[j + (long)lda * i]);
    }
    a[i * SJLD + j] = x;
    v[i * SJLD + j] = (i == j) ? 1.0 : 0.0;
    amax = fmax(amax, fabs(x));
  }
  for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o, 64));
  if (lane == 0) red[wv] = amax;
  __syncthreads();
  if (np != n && tid == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)blockDim.x / 64; ++w) m = fmax(m, red[w]);
    a[(np - 1) * SJLD + (np - 1)] = -(m * (double)n + 1.0) * 4.0;
  }
  __syncthreads();
}

// Descending order of the diagonal of a (one thread).
__device__ void jacobi_order(const double* a, int np, int* order) {
  if (threadIdx.x == 0) {
    for (int i = 0; i < np; ++i) order[i] = i;
    for (int i = 0; i < np; ++i) {
      int best = i;
      for (int j = i + 1; j < np; ++j)
        if (a[order[j] * SJLD + order[j]] > a[order[best] * SJLD + order[best]]) best = j;
      const int t = order[i]; order[i] = order[best]; order[best] = t;
    }
  }
  __syncthreads();
}

// Full eigen decomposition of n x n (n <= 64) symmetric matrices, batched.
__global__ __launch_bounds__(256) void jacobi_full_kernel(const double* __restrict__ A, long sA, int n, double* __restrict__ w,
                                                          long sw, double* __restrict__ V, long sV, int* __restrict__ info,
                                                          int max_sweeps, double tol) {
  __shared__ double a[SJ * SJLD];
  __shared__ double v[SJ * SJLD];
  __shared__ int order[SJ];
  const int bidx = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  A += bidx * sA;
  w += bidx * sw;
  V += bidx * sV;
  const int np = (n + 1) & ~1;
  jacobi_load(a, v, A, n, n, np, false);
  const int conv = jacobi_sweeps(a, v, n, np, max_sweeps, tol);
  jacobi_order(a, np, order);
  for (int j = wv; j < n; j += 4) {
    const int src = order[j];
    double best = -1.0;
    int bi = 0;
    for (int r = lane; r < n; r += 64) {
      const double x = fabs(v[r * SJLD + src]);
      if (x > best) { best = x; bi = r; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const double sg = v[bi * SJLD + src] < 0.0 ? -1.0 : 1.0;
    for (int r = lane; r < n; r += 64) V[r + (long)n * j] = sg * v[r * SJLD + src];
    if (lane == 0) w[j] = a[src * SJLD + src];
  }
  if (tid == 0 && info) info[bidx] = conv ? 0 : 1;
}

// Y = G^q W for a batch of problems (n <= 256): two columns per work-group of 1024
// threads; thread (row r, part p) sums the p-th quarter of the row's dot products, the
// quarters are added through LDS, and each power is followed by a column normalisation.
constexpr int PW_COLS = 2;
__global__ __launch_bounds__(1024) void power_kernel(const double* __restrict__ G, long sG, const double* __restrict__ W,
                                                     long sW, double* __restrict__ Y, long sY, int n, int b, int q,
                                                     int normalise) {
  __shared__ double cur[256 * PW_COLS];
  __shared__ double part[4][256 * PW_COLS];
  __shared__ double red[16][PW_COLS];
  __shared__ double scl[PW_COLS];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int r = tid & 255, p = tid >> 8;
  const int c0 = blockIdx.x * PW_COLS, bidx = blockIdx.y;
  const int nc = min(PW_COLS, b - c0);
  G += bidx * sG;
  W += bidx * sW;
  Y += bidx * sY;
  for (int e = tid; e < 256 * PW_COLS; e += 1024) {
    const int rr = e / PW_COLS, c = e % PW_COLS;
    cur[e] = (rr < n && c < nc) ? W[rr + (long)n * (c0 + c)] : 0.0;
  }
  __syncthreads();
  const int l0 = (p * n) / 4, l1 = ((p + 1) * n) / 4;
  for (int it = 0; it < q; ++it) {
    double a0 = 0.0, a1 = 0.0;
    if (r < n) {
      const double* gp = G + r;
#pragma unroll 16
      for (int l = l0; l < l1; ++l) {
        const double g = gp[(long)n * l];
        a0 += g * cur[l * PW_COLS + 0];
        a1 += g * cur[l * PW_COLS + 1];
      }
    }
    part[p][r * PW_COLS + 0] = a0;
    part[p][r * PW_COLS + 1] = a1;
    __syncthreads();
    double y = 0.0;
    if (tid < 256 * PW_COLS) {
      y = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    }
    if (normalise) {
      // tid < 512: row tid / 2, column tid % 2; reduce squares per column
      const double y2 = (tid < 256 * PW_COLS) ? y * y : 0.0;
      double s0 = (tid % 2 == 0) ? y2 : 0.0, s1 = (tid % 2 == 1) ? y2 : 0.0;
      s0 = wave_sum(s0);
      s1 = wave_sum(s1);
      if (lane == 0) { red[wv][0] = s0; red[wv][1] = s1; }
      __syncthreads();
      if (tid < PW_COLS) {
        double s = 0.0;
        for (int g = 0; g < 16; ++g) s += red[g][tid];
        scl[tid] = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
      }
      __syncthreads();
      if (tid < 256 * PW_COLS) y *= scl[tid % PW_COLS];
    }
    if (tid < 256 * PW_COLS) cur[tid] = y;
    __syncthreads();
  }
  for (int e = tid; e < n * nc; e += 1024) {
    const int rr = e % n, c = e / n;
    Y[rr + (long)n * (c0 + c)] = cur[rr * PW_COLS + c];
  }
}

// Q = orth(Y) (n x b, b <= n, n * b <= HH_MAX) by Householder QR. The reflectors and the
// explicitly formed Q both stay in LDS (row stride b).
constexpr int HH_MAX = 9400;
__global__ __launch_bounds__(1024) void hhqr_kernel(const double* __restrict__ Y, long sY, double* __restrict__ Q,
                                                    long sQ, int n, int b) {
  __shared__ double smem[2 * HH_MAX];
  __shared__ double red[16 * TQ_B];
  __shared__ double tau[TQ_B];
  const int tid = threadIdx.x, bidx = blockIdx.x;
  Y += bidx * sY;
  Q += bidx * sQ;
  double* A = smem;
  double* M = smem + HH_MAX;
  for (int e = tid; e < n * b; e += 1024) {
    const int r = e % n, c = e / n;
    A[r * b + c] = Y[r + (long)n * c];
  }
  __syncthreads();
  block_householder(A, tau, red, n, b, b);
  for (int e = tid; e < n * b; e += 1024) {
    const int r = e / b, c = e % b;
    M[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  apply_q(VLds{A, b}, tau, M, red, n, b, b);
  for (int e = tid; e < n * b; e += 1024) {
    const int r = e % n, c = e / n;
    Q[r + (long)n * c] = M[r * b + c];
  }
}

// Rayleigh-Ritz: T = sym(Q^T Z), Jacobi(T) = S Th S^T; Uf = Q S (all b Ritz vectors,
// descending), U = first k (canonical signs), w = Th[:k], converged[bidx] = 1 when every
// top-k Ritz residual ||Z s_i - th_i Q s_i|| <= tol * |th_1|. Q and Z are streamed
// through LDS in chunks of RR_ROWS rows.
constexpr int RR_ROWS = 32;
__global__ __launch_bounds__(256) void rr_kernel(const double* __restrict__ Q, const double* __restrict__ Z, long sQ,
                                                 int n, int b, int k, double* __restrict__ Uf, double* __restrict__ U,
                                                 long sU, double* __restrict__ w, long sw, double* __restrict__ wall,
                                                 double tol, int* __restrict__ converged, int max_sweeps) {
  __shared__ double a[SJ * SJLD];
  __shared__ double v[SJ * SJLD];
  __shared__ double qc[RR_ROWS * SJ];
  __shared__ double zc[RR_ROWS * SJ];
  __shared__ int order[SJ];
  __shared__ double res[SJ];
  __shared__ double sgn[SJ];
  const int tid = threadIdx.x, bidx = blockIdx.x;
  Q += bidx * sQ;
  Z += bidx * sQ;
  Uf += bidx * sQ;
  U += bidx * sU;
  w += bidx * sw;
  wall += (long)bidx * SJ;
  const int np = (b + 1) & ~1;
  const int npairs = b * (b + 1) / 2;
  // per-thread accumulators for up to 9 (i <= j) pairs (b <= 64: 2080 pairs / 256)
  double t1[9], t2[9];
#pragma unroll
  for (int u = 0; u < 9; ++u) { t1[u] = 0.0; t2[u] = 0.0; }
  for (int r0 = 0; r0 < n; r0 += RR_ROWS) {
    const int nr = min(RR_ROWS, n - r0);
    __syncthreads();
    for (int e = tid; e < RR_ROWS * b; e += 256) {
      const int r = e % RR_ROWS, c = e / RR_ROWS;
      qc[r * SJ + c] = (r < nr) ? Q[r0 + r + (long)n * c] : 0.0;
      zc[r * SJ + c] = (r < nr) ? Z[r0 + r + (long)n * c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 9; ++u) {
      const int pidx = tid + 256 * u;
      if (pidx < npairs) {
        // pidx -> (i <= j): column-major upper triangle enumeration
        int j = (int)((sqrt(8.0 * pidx + 1.0) - 1.0) * 0.5);
        while (j * (j + 1) / 2 > pidx) --j;
        while ((j + 1) * (j + 2) / 2 <= pidx) ++j;
        const int i = pidx - j * (j + 1) / 2;
        double s1 = 0.0, s2 = 0.0;
        for (int r = 0; r < nr; ++r) {
          s1 += qc[r * SJ + i] * zc[r * SJ + j];
          s2 += qc[r * SJ + j] * zc[r * SJ + i];
        }
        t1[u] += s1;
        t2[u] += s2;
      }
    }
  }
  for (int e = tid; e < np * np; e += 256) {
    const int i = e % np, j = e / np;
    a[i * SJLD + j] = 0.0;
    v[i * SJLD + j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 9; ++u) {
    const int pidx = tid + 256 * u;
    if (pidx < npairs) {
      int j = (int)((sqrt(8.0 * pidx + 1.0) - 1.0) * 0.5);
      while (j * (j + 1) / 2 > pidx) --j;
      while ((j + 1) * (j + 2) / 2 <= pidx) ++j;
      const int i = pidx - j * (j + 1) / 2;
      const double t = 0.5 * (t1[u] + t2[u]);
      a[i * SJLD + j] = t;
      a[j * SJLD + i] = t;
    }
  }
  __syncthreads();
  if (np != b && tid == 0) {
    double m = 0.0;
    for (int i = 0; i < b; ++i) m = fmax(m, fabs(a[i * SJLD + i]));
    a[(np - 1) * SJLD + (np - 1)] = -(m * (double)b + 1.0) * 4.0;
  }
  __syncthreads();
  jacobi_sweeps(a, v, b, np, max_sweeps, 1e-14);
  jacobi_order(a, np, order);
  if (tid < SJ) res[tid] = 0.0;
  __syncthreads();
  // Uf = Q S (descending order) and G u = Z S, chunk by chunk; residuals and the
  // largest-magnitude entry of each of the first k columns on the way.
  const int lane = tid & 63, wv = tid >> 6;
  for (int r0 = 0; r0 < n; r0 += RR_ROWS) {
    const int nr = min(RR_ROWS, n - r0);
    __syncthreads();
    for (int e = tid; e < RR_ROWS * b; e += 256) {
      const int r = e % RR_ROWS, c = e / RR_ROWS;
      qc[r * SJ + c] = (r < nr) ? Q[r0 + r + (long)n * c] : 0.0;
      zc[r * SJ + c] = (r < nr) ? Z[r0 + r + (long)n * c] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < RR_ROWS * b; e += 256) {
      const int r = e % RR_ROWS, j = e / RR_ROWS;
      if (r >= nr) continue;
      const int src = order[j];
      double u = 0.0, gu = 0.0;
      for (int l = 0; l < b; ++l) {
        const double sv = v[l * SJLD + src];
        u += qc[r * SJ + l] * sv;
        gu += zc[r * SJ + l] * sv;
      }
      Uf[r0 + r + (long)n * j] = u;
      if (j < k) {
        const double d = gu - a[src * SJLD + src] * u;
        atomicAdd(&res[j], d * d);
      }
    }
  }
  __syncthreads();
  // canonical signs of the first k columns (read back from Uf)
  for (int j = wv; j < k; j += 4) {
    double best = -1.0;
    int bi = 0;
    for (int r = lane; r < n; r += 64) {
      const double x = fabs(Uf[r + (long)n * j]);
      if (x > best) { best = x; bi = r; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if (lane == 0) sgn[j] = Uf[bi + (long)n * j] < 0.0 ? -1.0 : 1.0;
  }
  __syncthreads();
  for (int e = tid; e < n * k; e += 256) {
    const int r = e % n, j = e / n;
    U[r + (long)n * j] = sgn[j] * Uf[r + (long)n * j];
  }
  if (tid < b) wall[tid] = a[order[tid] * SJLD + order[tid]];
  if (tid < k) w[tid] = a[order[tid] * SJLD + order[tid]];
  if (tid == 0) {
    const double th1 = fabs(a[order[0] * SJLD + order[0]]);
    double rmax = 0.0;
    for (int j = 0; j < k; ++j) rmax = fmax(rmax, sqrt(res[j]));
    converged[bidx] = (rmax <= tol * th1 || th1 == 0.0) ? 1 : 0;
  }
}

int jacobi_full_launch(hipStream_t st, const double* A, long sA, int n, double* w, long sw, double* V, long sV,
                       int batch, int* info, int max_sweeps, double tol) {
  if (n < 1 || n > SJ || batch < 1) return 8;
  hipLaunchKernelGGL(jacobi_full_kernel, dim3(batch), dim3(256), 0, st, A, sA, n, w, sw, V, sV, info, max_sweeps, tol);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int power_launch(hipStream_t st, const double* G, long sG, const double* W, long sW, double* Y, long sY, int n, int b,
                 int q, int normalise, int batch) {
  if (n < 1 || n > 256 || b < 1 || b > 64 || batch < 1) return 8;
  hipLaunchKernelGGL(power_kernel, dim3((b + PW_COLS - 1) / PW_COLS, batch), dim3(1024), 0, st, G, sG, W, sW, Y, sY, n, b,
                     q, normalise);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int hhqr_launch(hipStream_t st, const double* Y, long sY, double* Vscratch, double* Q, long sQ, int n, int b, int batch) {
  (void)Vscratch;
  if (n < b || (long)n * b > HH_MAX || b < 1 || b > TQ_B || batch < 1) return 8;
  hipLaunchKernelGGL(hhqr_kernel, dim3(batch), dim3(1024), 0, st, Y, sY, Q, sQ, n, b);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int rr_launch(hipStream_t st, const double* Q, const double* Z, long sQ, int n, int b, int k, double* Uf, double* U,
              long sU, double* w, long sw, double* wall, double tol, int* converged, int batch) {
  if (n < 1 || b < 1 || b > SJ || k < 1 || k > b || batch < 1) return 8;
  hipLaunchKernelGGL(rr_kernel, dim3(batch), dim3(256), 0, st, Q, Z, sQ, n, b, k, Uf, U, sU, w, sw, wall, tol, converged,
                     20);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

}  // namespace ht
