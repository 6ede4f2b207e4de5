// Batched small symmetric eigen solvers of the HT engine (FP64, n <= 256, block <= 64).
//
// The node problems are small: leaf Grams (n x n, n = mode size) and the Grams of the
// leaf core's matricisations (k_l k_r x k_l k_r). Their top-k eigenpairs replace the
// reference's cusolverDnSsyevd / Sgesvd calls (third_order_tensor_decompsition/CUDA/
// Optimized/kernel.cu:682-718; fourth_order_tensor_decomposition/single_GPU/CUDA/
// Optimized/kernel.cu:267-287, :457-481). One round of the block solver is
//   power_kernel  Y = f(G) W         f = G^q (normalised powers) or, from the second round
//                                    on, the degree-q Chebyshev filter that damps [0, cut]
//                                    (cut = smallest Ritz value of the previous round)
//   qr_kernel     Q = orth(Y)        Householder QR in FP64, rows in registers
//   power_kernel  Z = G Q
//   rr_kernel     T = sym(Q^T Z) = S Th S^T by parallel Jacobi; U = Q S, the Ritz
//                 residuals ||G u_i - th_i u_i||, canonical signs and a converged flag
// and the host only reads the flags between rounds.
//
// Block layout: the n x b blocks W, Y, Q, Z and Uf are row-major (ld = b), one problem
// after another. The row-distributed kernels (qr, rr) run 16 waves; thread (wave g,
// lane c) holds column c of rows r = i * 16 + g, i < 16, in registers.
#include "k_jacobi.h"
#include <type_traits>

namespace ht {

// Full eigen decomposition of n x n (n <= 64) symmetric matrices, batched (one block of
// 1024 threads each); columns of V in descending eigenvalue order with canonical signs
// (largest-magnitude entry positive).
__global__ __launch_bounds__(1024) void jacobi_full_kernel(const double* __restrict__ A, long sA, int n, double* __restrict__ w,
                                                          long sw, double* __restrict__ V, long sV, int* __restrict__ info,
                                                          int max_sweeps, double tol) {
  __shared__ double a0[SJ * SJLD];
  __shared__ double v[SJ * SJLD];
  __shared__ int order[SJ];
  const int bidx = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  A += bidx * sA;
  w += bidx * sw;
  V += bidx * sV;
  double* a = a0;
  int conv;
  int np;
  if (n <= 16) {
    // padded to 16 (decoupled diagonal entries below every eigenvalue) for the one-wave
    // register Jacobi
    np = 16;
    for (int e = tid; e < 256; e += SNT) {
      const int i = e % 16, j = e / 16;
      a[i * SJLD + j] = (i < n && j < n) ? A[i + (long)n * j] : 0.0;
      v[i * SJLD + j] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    if (n < 16 && tid == 0) {
      double m = 0.0;
      for (int i = 0; i < n; ++i) m = fmax(m, fabs(a[i * SJLD + i]));
      for (int i = n; i < 16; ++i) a[i * SJLD + i] = -(m * (double)n + 1.0) * 4.0;
    }
    __syncthreads();
    const double pad = n < 16 ? a[15 * SJLD + 15] : -INFINITY;
    __shared__ int conv_s;
    if (tid < 64) {
      const int c = jacobi16_wave(a, v, max_sweeps, tol, pad);
      if (tid == 0) conv_s = c;
    }
    __syncthreads();
    conv = conv_s;
  } else {
    np = (n + 1) & ~1;
    jacobi_load(a, v, A, n, n, np, false);
    conv = jacobi_sweeps(a, v, n, np, max_sweeps, tol);
  }
  jacobi_order(a, np, order);
  for (int j = wv; j < n; j += SW) {
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

// Y = f(G) W for a batch of problems (n <= 256, row-major blocks, ld b): PW_COLS columns
// per work-group of 1024 threads; thread (row pair rp, part p) sums the p-th eighth of the
// dot products of rows 2 rp and 2 rp + 1, and the eighths are added through LDS. Two rows
// per thread let each value of the current block (one 16-byte LDS read per two columns)
// serve two rows: the loop is bound by LDS instruction issue, not by G's L2 traffic.
//   mode 0: q powers, every fourth one (and the last) followed by a column normalisation
//           (the span is unchanged);
//   mode 1: one product, unnormalised;
//   mode 2: the degree-q Chebyshev filter on [0, e], e = cut[problem]:
//           T_0 = W, T_1 = L W, T_{j+1} = 2 L T_j - T_{j-1}, L = (2 G - e I) / e, with
//           (T_j, T_{j-1}) rescaled together by the column norm of T_j every fourth step.
//           Falls back to mode 0 when e <= 0.
// With T != nullptr (mode 1), the block also forms T[l][c] = sum_r W[r][l] Y[r][c] for its
// columns (row-major b x b per problem, stride 64 * 64): the Rayleigh quotient Q^T G Q
// when W = Q.
template <int PW_COLS, typename TG>
__global__ __launch_bounds__(1024) void power_kernel(const TG* __restrict__ G, long sG, const double* __restrict__ W,
                                                     double* __restrict__ Y, long sWY, int n, int b, int q, int mode,
                                                     const double* __restrict__ cut, double* __restrict__ T,
                                                     uint64_t gseed, uint32_t gstream) {
  __shared__ __attribute__((aligned(16))) double cur[256 * PW_COLS];
  __shared__ double prv[256 * PW_COLS];
  __shared__ double part[8][256 * PW_COLS];
  __shared__ double red[SW][PW_COLS];
  __shared__ double scl[PW_COLS];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c0 = blockIdx.x * PW_COLS, bidx = blockIdx.y;
  const int nc = min(PW_COLS, b - c0);
  G += bidx * sG;
  const bool gen = W == nullptr;  // start block generated here (Philox normals)
  if (!gen) W += bidx * sWY;
  Y += bidx * sWY;
  double e = 0.0;
  if (mode == 2) {
    e = cut[bidx];
    if (!(e > 0.0)) mode = 0;
  }
  for (int x = tid; x < 256 * PW_COLS; x += 1024) {
    const int rr = x / PW_COLS, c = x % PW_COLS;
    cur[x] = (rr < n && c < nc)
                 ? (gen ? philox_normal(gseed, gstream, (uint64_t)(bidx * sWY + (long)rr * b + c0 + c))
                        : W[(long)rr * b + c0 + c])
                 : 0.0;
    prv[x] = 0.0;
  }
  __syncthreads();
  const int rp = tid & 127, p8 = tid >> 7;
  const int ra = 2 * rp, rb = 2 * rp + 1;
  const int l0 = (p8 * n) / 8, l1 = ((p8 + 1) * n) / 8;
  const bool own = tid < 256 * PW_COLS;  // tid = row * PW_COLS + column
  static_assert(PW_COLS == 2 || PW_COLS == 4, "one or two 16-byte reads per l");
  for (int it = 0; it < q; ++it) {
    double acc0[PW_COLS], acc1[PW_COLS];
#pragma unroll
    for (int c = 0; c < PW_COLS; ++c) acc0[c] = acc1[c] = 0.0;
    if (ra < n) {
      const TG* ga = G + ra;
      const TG* gb = G + (rb < n ? rb : ra);
      // eight steps of G per batch, all loads issued before their products (index clamped
      // to the range, past it the value is not used). For even n the thread's two rows are
      // adjacent in every column: one 2-element load, so a wave's load covers whole lines.
      using TG2 = typename std::conditional<std::is_same<TG, float>::value, float2, double2>::type;
      const bool pair = (n & 1) == 0;
      for (int l = l0; l < l1; l += 8) {
        TG g0v[8], g1v[8];
        if (pair) {
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const long lc = min(l + u, l1 - 1);
            const TG2 gg = *reinterpret_cast<const TG2*>(ga + (long)n * lc);
            g0v[u] = gg.x;
            g1v[u] = gg.y;
          }
        } else {
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const long lc = min(l + u, l1 - 1);
            g0v[u] = ga[(long)n * lc];
            g1v[u] = gb[(long)n * lc];
          }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (l + u < l1) {
            const double g0 = (double)g0v[u], g1 = (double)g1v[u];
            const double2 c01 = *reinterpret_cast<const double2*>(&cur[(l + u) * PW_COLS]);
            acc0[0] += g0 * c01.x; acc0[1] += g0 * c01.y;
            acc1[0] += g1 * c01.x; acc1[1] += g1 * c01.y;
            if constexpr (PW_COLS == 4) {
              const double2 c23 = *reinterpret_cast<const double2*>(&cur[(l + u) * PW_COLS + 2]);
              acc0[2] += g0 * c23.x; acc0[3] += g0 * c23.y;
              acc1[2] += g1 * c23.x; acc1[3] += g1 * c23.y;
            }
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < PW_COLS; ++c) {
      part[p8][ra * PW_COLS + c] = acc0[c];
      part[p8][rb * PW_COLS + c] = rb < n ? acc1[c] : 0.0;  // rows past n stay zero
    }
    __syncthreads();
    double y = 0.0, x = 0.0;
    if (own) {
#pragma unroll
      for (int pp = 0; pp < 8; ++pp) y += part[pp][tid];
      x = cur[tid];
      if (mode == 2) {
        const double ly = (2.0 * y - e * x) / e;
        y = (it == 0) ? ly : 2.0 * ly - prv[tid];
      }
    }
    // column normalisation every fourth step and on the last: it only keeps the values in
    // range (the span is unchanged), and four steps grow them by at most T_4(x_1) (mode 2)
    // or theta_1^4 (mode 0), far inside the FP64 range
    if (mode != 1 && ((it & 3) == 3 || it == q - 1)) {
      double s[PW_COLS];
#pragma unroll
      for (int c = 0; c < PW_COLS; ++c) s[c] = wave_sum((own && (tid % PW_COLS) == c) ? y * y : 0.0);
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < PW_COLS; ++c) red[wv][c] = s[c];
      }
      __syncthreads();
      if (tid < PW_COLS) {
        double t = 0.0;
        for (int w = 0; w < SW; ++w) t += red[w][tid];
        scl[tid] = t > 0.0 ? 1.0 / sqrt(t) : 0.0;
      }
      __syncthreads();
      if (own) {
        const double f = scl[tid % PW_COLS];
        y *= f;
        x *= f;
      }
    }
    if (own) {
      prv[tid] = x;
      cur[tid] = y;
    }
    __syncthreads();
  }
  for (int x = tid; x < n * nc; x += 1024) {
    const int rr = x / nc, c = x % nc;
    Y[(long)rr * b + c0 + c] = cur[rr * PW_COLS + c];
  }
  if (T) {
    const int l = tid & 63, c = (tid >> 6) & 3, pp = tid >> 8;
    double t = 0.0;
    if (l < b && c < nc && c < PW_COLS) {
      const int r0 = (pp * n) / 4, r1 = ((pp + 1) * n) / 4;
      for (int rr = r0; rr < r1; ++rr) t += W[(long)rr * b + l] * cur[rr * PW_COLS + c];
    }
    if (c < PW_COLS) part[pp][l * PW_COLS + c] = t;
    __syncthreads();
    if (tid < 64 * PW_COLS) {
      const int l2 = tid / PW_COLS, c2 = tid % PW_COLS;
      if (l2 < b && c2 < nc)
        T[(long)bidx * 64 * 64 + l2 * b + c0 + c2] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    }
  }
}

// Modes 0 and 2 of power_kernel with G in registers for the whole launch, loaded once, so
// a step costs the work-group its FMAs and one LDS reduction instead of another stream of
// G through L2 (the repeated stream is what bounds power_kernel). Wave p (of 16) holds
// columns [p L, p L + L) of G (L = NR / 16), lane l its rows l + 64 i, i < RPT = NR / 64:
// RPT * L values per thread. Every lane of a wave reads the same block row (an LDS
// broadcast) and each read feeds RPT FMAs per column. The 16 partial sums of a row meet in
// LDS. A step is bound by the work-group's FP64 FMAs and LDS traffic, so each work-group
// takes PW = 1 column: b work-groups, each with the least work per step. TG = float (the
// FP32 copy of G) or double; NR = 256 (n <= 256: 64 values per thread) or 128 (n <= 128:
// 16).
template <typename TG, int NR, int PW>
__global__ __launch_bounds__(1024) void power_reg_kernel(const TG* __restrict__ G, long sG, const double* __restrict__ W,
                                                         double* __restrict__ Y, long sWY, int n, int b, int q, int mode,
                                                         const double* __restrict__ cut, uint64_t gseed,
                                                         uint32_t gstream) {
  constexpr int NPART = 16, L = NR / NPART, RPT = NR / 64, CH = RPT * PW >= 4 ? 4 : 8;
  static_assert(L % CH == 0, "chunks of CH block rows");
  __shared__ double cur[NR * PW];
  __shared__ double prv[NR * PW];
  __shared__ double part[NPART][NR * PW];
  __shared__ double red[SW][PW];
  __shared__ double scl[PW];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c0 = blockIdx.x * PW, bidx = blockIdx.y;
  const int nc = min(PW, b - c0);
  G += bidx * sG;
  const bool gen = W == nullptr;  // start block generated here (Philox normals)
  if (!gen) W += bidx * sWY;
  Y += bidx * sWY;
  double e = 0.0;
  if (mode == 2) {
    e = cut[bidx];
    if (!(e > 0.0)) mode = 0;
  }
  const int l0 = wv * L;
  TG g[RPT][L];
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = lane + 64 * i;
#pragma unroll
    for (int j = 0; j < L; ++j) g[i][j] = (r < n && l0 + j < n) ? G[r + (long)n * (l0 + j)] : TG(0);
  }
  for (int x = tid; x < NR * PW; x += 1024) {
    const int rr = x / PW, c = x % PW;
    cur[x] = (rr < n && c < nc)
                 ? (gen ? philox_normal(gseed, gstream, (uint64_t)(bidx * sWY + (long)rr * b + c0 + c))
                        : W[(long)rr * b + c0 + c])
                 : 0.0;
    prv[x] = 0.0;
  }
  __syncthreads();
  const bool own = tid < NR * PW;  // tid = row * PW + column
  for (int it = 0; it < q; ++it) {
    double acc[PW][RPT];
#pragma unroll
    for (int c = 0; c < PW; ++c)
#pragma unroll
      for (int i = 0; i < RPT; ++i) acc[c][i] = 0.0;
    // chunks of CH block rows, the scheduler kept from hoisting later chunks' reads (the
    // resident values of G leave room for one chunk in flight)
#pragma unroll
    for (int j0 = 0; j0 < L; j0 += CH) {
      double cc[CH][PW];
#pragma unroll
      for (int u = 0; u < CH; ++u)
#pragma unroll
        for (int c = 0; c < PW; ++c) cc[u][c] = cur[(l0 + j0 + u) * PW + c];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          // (an empty asm on the value keeps its widening inside the step: hoisted out of
          // the step loop, the widened values would need twice the registers)
          TG gv = g[i][j0 + u];
          asm volatile("" : "+v"(gv));
          const double gj = (double)gv;
#pragma unroll
          for (int c = 0; c < PW; ++c) acc[c][i] = fma(gj, cc[u][c], acc[c][i]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = lane + 64 * i;
#pragma unroll
      for (int c = 0; c < PW; ++c) part[wv][r * PW + c] = acc[c][i];
    }
    __syncthreads();
    double y = 0.0, x = 0.0;
    if (own) {
#pragma unroll
      for (int pp = 0; pp < NPART; ++pp) y += part[pp][tid];
      x = cur[tid];
      if (mode == 2) {
        const double ly = (2.0 * y - e * x) / e;
        y = (it == 0) ? ly : 2.0 * ly - prv[tid];
      }
    }
    // column normalisation every fourth step and on the last (the span is unchanged)
    if ((it & 3) == 3 || it == q - 1) {
      double sq[PW];
#pragma unroll
      for (int c = 0; c < PW; ++c) sq[c] = wave_sum((own && (tid % PW) == c) ? y * y : 0.0);
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < PW; ++c) red[wv][c] = sq[c];
      }
      __syncthreads();
      if (tid < PW) {
        double t = 0.0;
        for (int w = 0; w < SW; ++w) t += red[w][tid];
        scl[tid] = t > 0.0 ? 1.0 / sqrt(t) : 0.0;
      }
      __syncthreads();
      if (own) {
        const double f = scl[tid % PW];
        y *= f;
        x *= f;
      }
    }
    if (own) {
      prv[tid] = x;
      cur[tid] = y;
    }
    __syncthreads();
  }
  for (int x = tid; x < n * nc; x += 1024) {
    const int rr = x / nc, c = x % nc;
    Y[(long)rr * b + c0 + c] = cur[rr * PW + c];
  }
}

// Lane packing of the row-distributed kernels: a wave holds BP columns (lane c < BP) of
// RS = 64 / BP row sets (h = lane / BP); thread (wave g, c, h) holds rows
// r = (i * SW + g) * RS + h, i < SR, so one block covers n <= SW * RS * SR rows.
template <int BP>
__device__ __forceinline__ double col_bcast(double x, int j, int h) {
  double r = lane_bcast(x, j);
#pragma unroll
  for (int q = 1; q < 64 / BP; ++q) {
    const double y = lane_bcast(x, j + q * BP);
    r = (h == q) ? y : r;
  }
  return r;
}
// x[L % 32] + x[32 + L % 32] in every lane L (v_permlane32_swap exchanges the upper half
// of one register with the lower half of the other).
__device__ __forceinline__ double halves_sum(double x) {
  const int lo = __double2loint(x), hi = __double2hiint(x);
  const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double(rh[0], rl[0]) + __hiloint2double(rh[1], rl[1]);
}
template <int BP>
__device__ __forceinline__ double rowset_sum(double s) {
#pragma unroll
  for (int o = BP; o < 32; o <<= 1) s += __shfl_xor(s, o, 64);
  if constexpr (BP <= 32) s = halves_sum(s);
  return s;
}

// Orders the LDS accesses of one wave (lanes read what other lanes of the same wave
// wrote before): LDS operations of a wave complete in order; the fences keep the
// compiler from moving loads or stores across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int QR_LDS = 8192;  // n * b limit of the staged blocks (doubles)

// Q = orth(Y) (row-major n x b, b <= n, b <= BP, n <= SW * (64/BP) * SR, n * b <= QR_LDS),
// rows in registers (thread (g, c, h) holds column c of its rows), the block staged in LDS.
//   First choice (when try_chol), CholeskyQR2: B = X^T X (one (l, c) dot product per
//   thread over the staged rows), B = R^T R and R^{-1} by wave 0 alone (right-looking
//   Cholesky, then back substitution with one column of R^{-1} per lane), X <- X R^{-1};
//   twice. It is taken only when every first-pass pivot exceeds 1e-12 * max diag(B)
//   (cond(Y) below ~1e6), where two passes give orthogonality to rounding.
//   Otherwise Householder QR: per column j the lanes of column j form (tau, beta) and the
//   reflector, store the reflector in the staging block (which then keeps all of them),
//   every thread adds its rows' share of w = v^T A, and the owners of columns > j apply
//   A -= v (tau w). Q is then formed explicitly as H_0 ... H_{b-1} I.
// Number of qr_kernel calls that fell back from Cholesky QR to Householder (diagnostics).
__device__ unsigned long long g_qr_householder = 0;

int qr_fallback_count(unsigned long long* out) {
  return hipMemcpyFromSymbol(out, HIP_SYMBOL(g_qr_householder), sizeof(*out)) == hipSuccess ? 0 : 5;
}

// Cholesky B = R^T R of the b x b SPD matrix B (LDS, ld SJLD, b <= 64) and X = R^{-1}, by
// the whole block (SNT threads) with one barrier per step: a dependent step costs a few
// hundred cycles whoever runs it, and spread over 16 waves its work is one or two items
// per thread. Step k writes row k of R to R (its divisions read the unscaled row k of
// B, which the trailing update of the same step only reads), updates the trailing
// square B[i][j] -= B[k][i] B[k][j] / B[k][k], i, j > k, and forms column k of X above
// the diagonal from X R = I: X[j][k] = -(sum_{i = j..k-1} X[j][i] R[i][k]) / R[k][k]
// (rows of R and columns of X from earlier steps; 16 threads per entry and a shuffle
// reduction), stored in B[j][k], j < k, which the factorisation no longer reads. A last
// pass sets the diagonal 1 / R[k][k] and zeroes the lower triangle. On success R (upper,
// zeros below) is in R and X (upper) in B; returns 0 (uniformly) when a pivot is not
// above piv_tol.
__device__ int chol_rinv_block(double* B, double* R, int b, double piv_tol) {
  __shared__ double rdiag_inv[SJ];  // 1 / R[k][k] of the steps done
  const int tid = threadIdx.x;
  const int jx = tid >> 4, part = tid & 15;
  for (int k = 0; k < b; ++k) {
    const double d = B[k * SJLD + k];
    if (!(d > piv_tol)) return 0;
    // 1/sqrt(d) from v_rsq_f64 and two Newton steps (~1 ulp): no divide or square root
    // on the step's dependent path
    const double rinv = rsq_nr(d), rkk = d * rinv, dinv = rinv * rinv;
    if (tid == 0) rdiag_inv[k] = rinv;
    {
      double t = 0.0;
      if (jx < k)
        for (int i = jx + part; i < k; i += 16)
          t += (i == jx ? rdiag_inv[jx] : B[jx * SJLD + i]) * R[i * SJLD + k];
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
      if (part == 0 && jx < k) B[jx * SJLD + k] = -t * rinv;
    }
    const int m = b - k - 1;
    for (int e = tid; e < b + m * m; e += SNT) {
      if (e < b) {
        R[k * SJLD + e] = e > k ? B[k * SJLD + e] * rinv : (e == k ? rkk : 0.0);
      } else {
        const int q = e - b, ii = q / m, jj = q - ii * m;
        const int i = k + 1 + ii, j = k + 1 + jj;
        B[i * SJLD + j] -= B[k * SJLD + i] * B[k * SJLD + j] * dinv;
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < b * b; e += SNT) {
    const int r = e / b, c = e % b;
    if (r > c) B[r * SJLD + c] = 0.0;
    else if (r == c) B[r * SJLD + c] = rdiag_inv[r];
  }
  __syncthreads();
  return 1;
}

// CholeskyQR2 of Y (row-major n x b, n % 64 == 0, b % 16 == 0, b <= BP, n * b <= QR_LDS)
// for a block of SNT threads, with the products on FP64 MFMA (16x16x4):
//   B = X^T X     waves own (16 x 16 tile, row range) pairs, partial tiles summed via LDS;
//   B = R^T R, R^{-1}  the whole block (chol_rinv_block; R^{-1} lands in B);
//   X <- X R^{-1} output tiles spread over the waves.
// Twice; the first pass requires every pivot above 1e-12 * max diag(B), the second above
// 1e-3 * max diag(B). The second pass starts from a nearly orthonormal X: when its Gram
// is B = I + E with ||E||_F <= 1e-3 (the usual case) it uses W = B^{-1/2} from the
// binomial series I - E/2 + 3/8 E^2 - 5/16 E^3 + 35/128 E^4 (next term below 3e-16)
// instead, three b x b MFMA products and no dependent scalar steps; X W is orthonormal to
// rounding and spans the same columns. Returns true with Q written, false (Q untouched)
// when a pivot test failed. Ys holds n * b doubles, B and Ri SJ * SJLD each.
template <int BP>
__device__ bool cholqr2_mfma(const double* __restrict__ Y, double* __restrict__ Q, double* Ys, double* B, double* Ri,
                             int n, int b) {
  __shared__ double red[SW];
  const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
  const int i16 = lane & 15, kk = lane >> 4;
  const int nt = b / 16, T = nt * nt, KQ = T >= 16 ? 1 : 16 / T;
  for (int e = tid; e < n * b; e += SNT) Ys[e] = Y[e];
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    // B = Ys^T Ys: wave g < T * KQ takes tile t = g % T over rows [kq * n / KQ, +n / KQ);
    // partial tiles go to Ri (KQ * T * 256 <= 4096 doubles), then are summed into B.
    double* P = Ri;
    if (g < T * KQ) {
      const int t = g % T, kq = g / T, I = t / nt, J = t % nt;
      const int rows = n / KQ;
      f64x4 acc = {};
      for (int st = 0; st < rows / 4; ++st) {
        const int r = kq * rows + 4 * st + kk;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ys[r * b + 16 * I + i16], Ys[r * b + 16 * J + i16], acc, 0, 0, 0);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) P[(g * 16 + acc16_row_f64(e, lane)) * 16 + i16] = acc[e];
    }
    __syncthreads();
    for (int e = tid; e < b * b; e += SNT) {
      const int r = e / b, c = e % b;
      const int t = (r / 16) * nt + c / 16;
      double sum = 0.0;
      for (int kq = 0; kq < KQ; ++kq) sum += P[((kq * T + t) * 16 + r % 16) * 16 + c % 16];
      B[r * SJLD + c] = sum;
    }
    __syncthreads();
    bool series = false;
    if (pass == 1) {
      double e2 = 0.0;
      for (int e = tid; e < b * b; e += SNT) {
        const int r = e / b, c = e % b;
        const double x = B[r * SJLD + c] - (r == c ? 1.0 : 0.0);
        e2 += x * x;
      }
      e2 = wave_sum(e2);
      if (lane == 0) red[g] = e2;
      __syncthreads();
      double tot = 0.0;
#pragma unroll
      for (int w = 0; w < SW; ++w) tot += red[w];
      series = tot <= 1e-6;  // uniform
    }
    const double* W = B;  // right factor of the update X <- X W
    if (series) {
      // E = B - I (in B); Horner: P = 35/128 E - 5/16 I, P = E P + 3/8 I, P = E P - 1/2 I,
      // P = E P + I (in Ri)
      for (int e = tid; e < b * b; e += SNT) {
        const int r = e / b, c = e % b;
        const double x = B[r * SJLD + c] - (r == c ? 1.0 : 0.0);
        B[r * SJLD + c] = x;
        Ri[r * SJLD + c] = (35.0 / 128.0) * x + (r == c ? -5.0 / 16.0 : 0.0);
      }
      __syncthreads();
      const double cs[3] = {3.0 / 8.0, -0.5, 1.0};
#pragma unroll
      for (int h = 0; h < 3; ++h) {
        f64x4 acc = {};
        const int I = g / nt, J = g % nt;
        if (g < T)
          for (int st = 0; st < b / 4; ++st)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(B[(16 * I + i16) * SJLD + 4 * st + kk],
                                                      Ri[(4 * st + kk) * SJLD + 16 * J + i16], acc, 0, 0, 0);
        __syncthreads();
        if (g < T) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int r = 16 * I + acc16_row_f64(e, lane), c = 16 * J + i16;
            Ri[r * SJLD + c] = acc[e] + (r == c ? cs[h] : 0.0);
          }
        }
        __syncthreads();
      }
      W = Ri;
    } else {
      double dmax = 0.0;
      for (int k = 0; k < b; ++k) dmax = fmax(dmax, B[k * SJLD + k]);
      const double piv_tol = (pass == 0 ? 1e-12 : 1e-3) * dmax;
      if (!chol_rinv_block(B, Ri, b, piv_tol)) return false;  // uniform: every thread read the same pivot
    }
    // X <- X W: output tile (RT, CT); W = R^{-1} is upper triangular (k < 16 (CT + 1)),
    // the series factor is full. n b <= 8192 makes at most 32 tiles, two per wave.
    const int ntile = (n / 16) * nt;
    f64x4 res[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int tile = g + 16 * u;
      res[u] = f64x4{};
      if (tile < ntile) {
        const int RT = tile / nt, CT = tile % nt;
        const int ksteps = series ? b / 4 : 4 * (CT + 1);
        for (int st = 0; st < ksteps; ++st)
          res[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ys[(16 * RT + i16) * b + 4 * st + kk],
                                                        W[(4 * st + kk) * SJLD + 16 * CT + i16], res[u], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int tile = g + 16 * u;
      if (tile < ntile) {
        const int RT = tile / nt, CT = tile % nt;
#pragma unroll
        for (int e = 0; e < 4; ++e) Ys[(16 * RT + acc16_row_f64(e, lane)) * b + 16 * CT + i16] = res[u][e];
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < n * b; e += SNT) Q[e] = Ys[e];
  return true;
}

template <int BP, int SR>
__global__ __launch_bounds__(1024) void qr_kernel(const double* __restrict__ Y, double* __restrict__ Q, long sYQ, int n,
                                                  int b, int try_chol) {
  constexpr int RS = 64 / BP;
  __shared__ double Ys[QR_LDS];
  __shared__ double B[SJ * SJLD];
  __shared__ double Ri[SJ * SJLD];
  __shared__ double nrm[SW];
  __shared__ double red[2][SW][BP];
  __shared__ double taus[64];
  __shared__ double scal[2];
  __shared__ double alpha_s;
  __shared__ int bad;
  const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
  const int c = lane % BP, h = lane / BP;
  Y += blockIdx.x * sYQ;
  Q += blockIdx.x * sYQ;
  double a[SR];
#pragma unroll
  for (int i = 0; i < SR; ++i) {
    const int r = (i * SW + g) * RS + h;
    a[i] = (r < n && c < b) ? Y[(long)r * b + c] : 0.0;
  }
  // ---- CholeskyQR2
  bool chol = try_chol != 0;
  // (n % 64 == 0 and n b <= 8192 leave n <= 128 for b > 32: the SR = 16 packing never
  // takes this path)
  if constexpr (SR <= 8) {
    if (chol && n % 64 == 0 && b % 16 == 0) {
      if (cholqr2_mfma<BP>(Y, Q, Ys, B, Ri, n, b)) return;
      chol = false;
    }
  }
  for (int pass = 0; pass < 2 && chol; ++pass) {
#pragma unroll
    for (int i = 0; i < SR; ++i) {
      const int r = (i * SW + g) * RS + h;
      if (r < n && c < b) Ys[r * b + c] = a[i];
    }
    __syncthreads();
    for (int e = tid; e < b * b; e += SNT) {
      const int l = e / b, cc = e % b;
      double t = 0.0;
      for (int r = 0; r < n; ++r) t += Ys[r * b + l] * Ys[r * b + cc];
      B[l * SJLD + cc] = t;
    }
    __syncthreads();
    if (g == 0) {
      int ok = 1;
      double dmax = 0.0;
      for (int k = 0; k < b; ++k) dmax = fmax(dmax, B[k * SJLD + k]);
      const double piv_tol = (pass == 0 ? 1e-12 : 1e-3) * dmax;
      for (int k = 0; k < b; ++k) {
        const double d = B[k * SJLD + k];
        if (!(d > piv_tol)) {
          ok = 0;
          break;
        }
        const double inv = 1.0 / sqrt(d);
        const int m = b - k - 1;
        for (int e = lane; e < m * m; e += 64) {
          const int i = k + 1 + e % m, j = k + 1 + e / m;
          if (i <= j) B[i * SJLD + j] -= (B[k * SJLD + i] * inv) * (B[k * SJLD + j] * inv);
        }
        wave_sync();
        for (int j = k + 1 + lane; j < b; j += 64) B[k * SJLD + j] *= inv;
        if (lane == 0) B[k * SJLD + k] = d * inv;
        wave_sync();
      }
      if (ok && lane < b) {
        // column `lane` of R^{-1} (upper): back substitution
        for (int i = b - 1; i >= 0; --i) {
          double t = (i == lane) ? 1.0 : 0.0;
          for (int k = i + 1; k <= lane; ++k) t -= B[i * SJLD + k] * Ri[k * SJLD + lane];
          Ri[i * SJLD + lane] = (i <= lane) ? t / B[i * SJLD + i] : 0.0;
        }
      }
      if (lane == 0) bad = !ok;
    }
    __syncthreads();
    if (bad) {
      chol = false;
      break;
    }
    // X <- X R^{-1}
#pragma unroll
    for (int i = 0; i < SR; ++i) {
      const int r = (i * SW + g) * RS + h;
      double t = 0.0;
      if (r < n && c < b)
        for (int l = 0; l <= c; ++l) t += Ys[r * b + l] * Ri[l * SJLD + c];
      a[i] = t;
    }
    __syncthreads();
  }
  if (chol) {
#pragma unroll
    for (int i = 0; i < SR; ++i) {
      const int r = (i * SW + g) * RS + h;
      if (r < n && c < b) Q[(long)r * b + c] = a[i];
    }
    return;
  }
  // ---- Householder QR
  if (tid == 0) atomicAdd(&g_qr_householder, 1ull);
#pragma unroll
  for (int i = 0; i < SR; ++i) {
    const int r = (i * SW + g) * RS + h;
    a[i] = (r < n && c < b) ? Y[(long)r * b + c] : 0.0;
  }
  {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < SR; ++i) {
      const int r = (i * SW + g) * RS + h;
      t += (c == 0 && r > 0) ? a[i] * a[i] : 0.0;
      if (c == 0 && r == 0) alpha_s = a[i];
    }
    t = rowset_sum<BP>(t);
    if (lane == 0) nrm[g] = t;
  }
  for (int j = 0; j < b; ++j) {
    __syncthreads();
    if (c == j) {
      double sub = 0.0;
#pragma unroll
      for (int w = 0; w < SW; ++w) sub += nrm[w];
      const double alpha = alpha_s;
      double tau = 0.0, scale = 0.0, beta = alpha;
      if (sub > 0.0) {
        const double nr = sqrt(alpha * alpha + sub);
        beta = alpha >= 0.0 ? -nr : nr;
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
#pragma unroll
      for (int i = 0; i < SR; ++i) {
        const int r = (i * SW + g) * RS + h;
        if (r < n) Ys[r * b + j] = r > j ? a[i] * scale : (r == j ? 1.0 : 0.0);
      }
      if (g == 0 && h == 0) {
        taus[j] = tau;
        scal[0] = tau;
        scal[1] = beta;
      }
    }
    __syncthreads();
    const double tau = scal[0], beta = scal[1];
    double vv[SR];
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < SR; ++i) {
      const int r = (i * SW + g) * RS + h;
      vv[i] = r < n ? Ys[r * b + j] : 0.0;
      t += vv[i] * a[i];
    }
    t = rowset_sum<BP>(t);
    const int par = j & 1;
    if (h == 0) red[par][g][c] = t;
    __syncthreads();
    if (c > j && c < b && tau != 0.0) {
      double wsum = 0.0;
#pragma unroll
      for (int w = 0; w < SW; ++w) wsum += red[par][w][c];
      const double wc = tau * wsum;
#pragma unroll
      for (int i = 0; i < SR; ++i) a[i] -= vv[i] * wc;
    }
    if (c == j) {
#pragma unroll
      for (int i = 0; i < SR; ++i) {
        const int r = (i * SW + g) * RS + h;
        if (r == j) a[i] = beta;
      }
    }
    // partial norm of column j + 1 below the diagonal, and its diagonal entry
    double t2 = 0.0;
#pragma unroll
    for (int i = 0; i < SR; ++i) {
      const int r = (i * SW + g) * RS + h;
      t2 += (c == j + 1 && r > j + 1) ? a[i] * a[i] : 0.0;
      if (c == j + 1 && r == j + 1) alpha_s = a[i];
    }
    t2 = rowset_sum<BP>(t2);
    if (lane == ((j + 1) % BP) && j + 1 < BP) nrm[g] = t2;
  }
  double m[SR];
#pragma unroll
  for (int i = 0; i < SR; ++i) m[i] = ((i * SW + g) * RS + h == c && c < b) ? 1.0 : 0.0;
  __syncthreads();
  for (int j = b - 1; j >= 0; --j) {
    const double tau = taus[j];
    double vv[SR];
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < SR; ++i) {
      const int r = (i * SW + g) * RS + h;
      vv[i] = r < n ? Ys[r * b + j] : 0.0;
      t += vv[i] * m[i];
    }
    t = rowset_sum<BP>(t);
    const int par = j & 1;
    if (h == 0) red[par][g][c] = t;
    __syncthreads();
    if (c < b && tau != 0.0) {
      double wsum = 0.0;
#pragma unroll
      for (int w = 0; w < SW; ++w) wsum += red[par][w][c];
      const double wc = tau * wsum;
#pragma unroll
      for (int i = 0; i < SR; ++i) m[i] -= vv[i] * wc;
    }
  }
#pragma unroll
  for (int i = 0; i < SR; ++i) {
    const int r = (i * SW + g) * RS + h;
    if (r < n && c < b) Q[(long)r * b + c] = m[i];
  }
}

// Rayleigh-Ritz from T = Q^T G Q (row-major b x b, formed with Z = G Q): Jacobi on
// sym(T) = S Th S^T; Uf = Q S (all b Ritz vectors, descending, row-major), U = first k
// (column-major, canonical signs), w = Th[:k], wall = Th, cut = Th[b-1], converged = 1
// when every top-k Ritz residual ||Z s_i - th_i Q s_i|| <= max(tol |th_1|, tol_rel |th_i|), pairs with
// |th_i| <= floor * |th_1| exempt (a cluster below what the Gram resolves); rres (may be
// null) receives that largest residual over |th_1|. Q and then Z are
// staged in LDS; thread (g, c, h) forms column c of the Ritz vectors on its rows.
template <int BP, int SR>
__global__ __launch_bounds__(1024) void rr_kernel(const double* __restrict__ Q, const double* __restrict__ Z, long sQ,
                                                  const double* __restrict__ T, int n, int b, int k,
                                                  double* __restrict__ Uf, double* __restrict__ U, long sU,
                                                  double* __restrict__ w, long sw, double* __restrict__ wall,
                                                  double* __restrict__ cut, double tol, double tol_rel,
                                                  double floor, int* __restrict__ converged,
                                                  double* __restrict__ rres, int max_sweeps) {
  constexpr int RS = 64 / BP;
  __shared__ double a[SJ * SJLD];
  __shared__ double v[SJ * SJLD];
  __shared__ double Ms[QR_LDS];
  __shared__ int order[SJ];
  __shared__ double res[SJ];
  __shared__ double bests[SW][BP];
  __shared__ int besti[SW][BP];
  __shared__ double sgn[SJ];
  const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6, bidx = blockIdx.x;
  const int c = lane % BP, h = lane / BP;
  Q += bidx * sQ;
  Z += bidx * sQ;
  Uf += bidx * sQ;
  T += (long)bidx * 64 * 64;
  U += bidx * sU;
  w += bidx * sw;
  wall += (long)bidx * SJ;
  // blocks of up to 16 columns are padded to 16 and solved by one wave in registers (a
  // one-wave 32 x 32 version, four blocks per lane, measured slower than the block-wide
  // Jacobi below: its lane permutes outweigh the barriers saved)
  constexpr int WN = BP == 16 ? 16 : 0;
  const int np = WN ? WN : (b + 1) & ~1;
  for (int e = tid; e < np * np; e += SNT) {
    const int i = e % np, j = e / np;
    a[i * SJLD + j] = (i < b && j < b) ? 0.5 * (T[i * b + j] + T[j * b + i]) : 0.0;
    v[i * SJLD + j] = (i == j) ? 1.0 : 0.0;
  }
  if (tid < SJ) res[tid] = 0.0;
  for (int e = tid; e < n * b; e += SNT) Ms[e] = Q[e];
  __syncthreads();
  if (np != b && tid == 0) {
    double m = 0.0;
    for (int i = 0; i < b; ++i) m = fmax(m, fabs(a[i * SJLD + i]));
    for (int i = b; i < np; ++i) a[i * SJLD + i] = -(m * (double)b + 1.0) * 4.0;
  }
  __syncthreads();
  // the Ritz residuals are tested against tol * theta_1: an off-diagonal remainder 1e2
  // times smaller (relative, in the Frobenius norm; never below 1e-14) moves them by at
  // most ~sqrt(b) 1e-2 tol theta_1, well inside the test
  const double jtol = fmax(1e-14, 1e-2 * tol);
  if constexpr (WN == 16) {
    const double pad = b < 16 ? a[15 * SJLD + 15] : -INFINITY;
    if (tid < 64) jacobi16_wave(a, v, max_sweeps, jtol, pad);
    __syncthreads();
  } else {
    jacobi_sweeps(a, v, b, np, max_sweeps, jtol);
  }
  jacobi_order(a, np, order);
  // Ritz vectors u = Q S[:, order[c]] and G u = Z S[:, order[c]] on this thread's rows.
  const int src = c < b ? order[c] : 0;
  const double th = c < b ? a[src * SJLD + src] : 0.0;
  // (the row index is the inner loop: SR independent accumulators keep the LDS loads
  // in flight; rows past n read row 0 and are not used)
  double u[SR];
#pragma unroll
  for (int i = 0; i < SR; ++i) u[i] = 0.0;
  if (c < b) {
    for (int l = 0; l < b; ++l) {
      const double sl = v[l * SJLD + src];
#pragma unroll
      for (int i = 0; i < SR; ++i) {
        const int r = (i * SW + g) * RS + h;
        u[i] += Ms[(r < n ? r : 0) * b + l] * sl;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < n * b; e += SNT) Ms[e] = Z[e];
  __syncthreads();
  double gu[SR];
#pragma unroll
  for (int i = 0; i < SR; ++i) gu[i] = 0.0;
  if (c < b) {
    for (int l = 0; l < b; ++l) {
      const double sl = v[l * SJLD + src];
#pragma unroll
      for (int i = 0; i < SR; ++i) {
        const int r = (i * SW + g) * RS + h;
        gu[i] += Ms[(r < n ? r : 0) * b + l] * sl;
      }
    }
  }
  double rsum = 0.0;
#pragma unroll
  for (int i = 0; i < SR; ++i) {
    const int r = (i * SW + g) * RS + h;
    const double d = (r < n && c < b) ? gu[i] - th * u[i] : 0.0;
    rsum += d * d;
  }
  rsum = rowset_sum<BP>(rsum);
  // largest-magnitude entry of this thread's rows (first occurrence), then Uf
  double bv = -1.0, bs = 0.0;
  int bi = 0;
#pragma unroll
  for (int i = 0; i < SR; ++i) {
    const int r = (i * SW + g) * RS + h;
    if (r < n) {
      const double x = fabs(u[i]);
      if (x > bv) { bv = x; bs = u[i]; bi = r; }
      if (c < b) Uf[(long)r * b + c] = u[i];
    }
  }
#pragma unroll
  for (int o = BP; o < 64; o <<= 1) {
    const double ov = __shfl_xor(bv, o, 64);
    const double os = __shfl_xor(bs, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bs = os; bi = oi; }
  }
  if (h == 0) {
    bests[g][c] = bs;
    besti[g][c] = bi;
    if (c < b) atomicAdd(&res[c], rsum);
  }
  __syncthreads();
  if (tid < b) {
    double best = -1.0, sb = 0.0;
    int ib = 0;
    for (int ww = 0; ww < SW; ++ww) {
      const double x = fabs(bests[ww][tid]);  // -1 sentinel rows only when n < 1
      const int ix = besti[ww][tid];
      if (x > best || (x == best && ix < ib)) { best = x; ib = ix; sb = bests[ww][tid]; }
    }
    sgn[tid] = sb < 0.0 ? -1.0 : 1.0;
  }
  __syncthreads();
  if (c < k) {
    const double sg = sgn[c];
#pragma unroll
    for (int i = 0; i < SR; ++i) {
      const int r = (i * SW + g) * RS + h;
      if (r < n) U[r + (long)n * c] = sg * u[i];
    }
  }
  if (g == 0 && h == 0 && c < b) {
    wall[c] = th;
    if (c < k) w[c] = th;
  }
  if (tid == 0) {
    const double th1 = fabs(a[order[0] * SJLD + order[0]]);
    double rmax = 0.0;
    int ok = 1;
    for (int j = 0; j < k; ++j) {
      const double thj = fabs(a[order[j] * SJLD + order[j]]);
      if (thj > floor * th1) {
        const double rj = sqrt(res[j]);
        rmax = fmax(rmax, rj);
        if (!(rj <= fmax(tol * th1, tol_rel * thj))) ok = 0;
      }
    }
    converged[bidx] = (ok || th1 == 0.0) ? 1 : 0;
    if (rres) rres[bidx] = th1 > 0.0 ? rmax / th1 : 0.0;
    cut[bidx] = a[order[b - 1] * SJLD + order[b - 1]];
  }
}

int jacobi_full_launch(hipStream_t st, const double* A, long sA, int n, double* w, long sw, double* V, long sV,
                       int batch, int* info, int max_sweeps, double tol) {
  if (n < 1 || n > SJ || batch < 1) return 8;
  hipLaunchKernelGGL(jacobi_full_kernel, dim3(batch), dim3(SNT), 0, st, A, sA, n, w, sw, V, sV, info, max_sweeps, tol);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// G32, when given, is an FP32 copy of G that modes 0 and 2 read instead; mode 1 always
// reads G. Filter steps keep G in registers (power_reg_kernel, one column per work-group)
// for FP32 G and for FP64 G with n <= 128; the rest take power_kernel, two columns per
// work-group.
int power_launch(hipStream_t st, const double* G, long sG, const double* W, double* Y, long sWY, int n, int b, int q,
                 int mode, const double* cut, double* T, int batch, const float* G32, uint64_t gseed,
                 uint32_t gstream) {
  if (n < 1 || n > 256 || b < 1 || b > 64 || q < 1 || mode < 0 || mode > 2 || batch < 1) return 8;
  if (mode == 2 && !cut) return 8;
  if (T && (mode != 1 || q != 1 || !W)) return 8;
  const dim3 grid((b + 1) / 2, batch), grid1(b, batch);
  // filter steps (modes 0 and 2) with G held in registers, one column per work-group
  if (mode != 1 && G32) {
    if (n <= 128)
      hipLaunchKernelGGL((power_reg_kernel<float, 128, 1>), grid1, dim3(1024), 0, st, G32, sG, W, Y, sWY, n, b, q,
                         mode, cut, gseed, gstream);
    else
      hipLaunchKernelGGL((power_reg_kernel<float, 256, 1>), grid1, dim3(1024), 0, st, G32, sG, W, Y, sWY, n, b, q,
                         mode, cut, gseed, gstream);
  } else if (mode != 1 && n <= 128) {
    hipLaunchKernelGGL((power_reg_kernel<double, 128, 1>), grid1, dim3(1024), 0, st, G, sG, W, Y, sWY, n, b, q, mode,
                       cut, gseed, gstream);
  } else {
    // one product (mode 1, with T), or FP64 filter steps on n > 128 (128 values per thread
    // would not fit the register file)
    hipLaunchKernelGGL((power_kernel<2, double>), grid, dim3(1024), 0, st, G, sG, W, Y, sWY, n, b, q, mode, cut, T,
                       gseed, gstream);
  }
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// Launches pick the smallest lane packing (BP, SR) with b <= BP and
// n <= 16 * (64 / BP) * SR: <16,2>, <32,4>, <32,8>, <64,8>, <64,16>.
int qr_launch(hipStream_t st, const double* Y, double* Q, long sYQ, int n, int b, int try_chol, int batch) {
  if (n < 1 || n > SW * SR || b < 1 || b > 64 || b > n || n * b > QR_LDS || batch < 1) return 8;
  const dim3 grid(batch), block(SNT);
  if (b <= 16 && n <= 128) hipLaunchKernelGGL((qr_kernel<16, 2>), grid, block, 0, st, Y, Q, sYQ, n, b, try_chol);
  else if (b <= 32 && n <= 128) hipLaunchKernelGGL((qr_kernel<32, 4>), grid, block, 0, st, Y, Q, sYQ, n, b, try_chol);
  else if (b <= 32) hipLaunchKernelGGL((qr_kernel<32, 8>), grid, block, 0, st, Y, Q, sYQ, n, b, try_chol);
  else if (n <= 128) hipLaunchKernelGGL((qr_kernel<64, 8>), grid, block, 0, st, Y, Q, sYQ, n, b, try_chol);
  else hipLaunchKernelGGL((qr_kernel<64, 16>), grid, block, 0, st, Y, Q, sYQ, n, b, try_chol);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

int rr_launch(hipStream_t st, const double* Q, const double* Z, long sQ, const double* T, int n, int b, int k,
              double* Uf, double* U, long sU, double* w, long sw, double* wall, double* cut, double tol,
              double tol_rel, double floor, int* converged, double* rres, int batch) {
  if (n < 1 || n > SW * SR || b < 1 || b > SJ || b > n || n * b > QR_LDS || k < 1 || k > b || batch < 1) return 8;
  const dim3 grid(batch), block(SNT);
  const int sweeps = 20;
#define HT_RR(BPv, SRv)                                                                                          \
  hipLaunchKernelGGL((rr_kernel<BPv, SRv>), grid, block, 0, st, Q, Z, sQ, T, n, b, k, Uf, U, sU, w, sw, wall, cut, \
                     tol, tol_rel, floor, converged, rres, sweeps)
  if (b <= 16 && n <= 128) HT_RR(16, 2);
  else if (b <= 32 && n <= 128) HT_RR(32, 4);
  else if (b <= 32) HT_RR(32, 8);
  else if (n <= 128) HT_RR(64, 8);
  else HT_RR(64, 16);
#undef HT_RR
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

}  // namespace ht
