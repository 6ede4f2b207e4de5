// Jacobi eigen solver device routines (FP64, n <= 64) shared by the small solvers and
// the batched kernel.
#pragma once
#include "ht_common.h"

namespace ht {

constexpr int SJ = 64;        // largest Jacobi problem
constexpr int SJLD = SJ + 1;  // LDS row stride of the Jacobi matrices
constexpr int SW = 16;        // waves of the 1024-thread kernels
constexpr int SR = 16;        // rows per thread there: n <= SW * SR
constexpr int SNT = SW * 64;  // threads of the 1024-thread kernels

__device__ __forceinline__ double lane_bcast(double x, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
  return __hiloint2double(hi, lo);
}

// x from lane l - 1 (ROW_SR1) or l + 1 (ROW_SL1) of the same 16-lane row, by DPP moves;
// the row's first (last) lane keeps 0.
constexpr int DPP_ROW_SL1 = 0x101, DPP_ROW_SR1 = 0x111;
template <int CTRL>
__device__ __forceinline__ double dpp_row_shift(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Round-robin pairing for np (even) indices: step s pairs (np-1, s) and
// ((s+t) mod (np-1), (s-t) mod (np-1)), t = 1 .. np/2-1; every pair meets once per sweep.
// Pair t of step s as (p, q), p < q, computed where it is needed (no table, no LDS read on
// the step's dependent path).
__device__ __forceinline__ void jacobi_pair(int s, int t, int np, int& p, int& q) {
  const int m = np - 1;
  int x, y;
  if (t == 0) {
    x = m;
    y = s;
  } else {
    x = s + t;
    if (x >= m) x -= m;
    y = s - t;
    if (y < 0) y += m;
  }
  p = min(x, y);
  q = max(x, y);
}

// 1/x and 1/sqrt(x) in FP64 from the hardware approximations (v_rcp_f64, v_rsq_f64) and
// two Newton steps each (quadratic convergence from ~2^-22 leaves ~1 ulp).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}
__device__ __forceinline__ double rsq_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  r = fma(0.5 * r, fma(-x * r, r, 1.0), r);
  r = fma(0.5 * r, fma(-x * r, r, 1.0), r);
  return r;
}

// Jacobi rotation (c, s) that annihilates a_pq (Rutishauser's form: t = tan(phi) of the
// smaller angle, c = 1/sqrt(1 + t^2), s = t c; c^2 + s^2 = 1 to rounding).
__device__ __forceinline__ void jacobi_rotation(double app, double aqq, double apq, double& c, double& s) {
  c = 1.0;
  s = 0.0;
  if (apq == 0.0) return;
  const double tau = (aqq - app) * rcp_nr(2.0 * apq);
  const double at = fabs(tau);
  double t;
  if (at > 1e150) {
    t = 0.5 * rcp_nr(at);
  } else {
    const double h = 1.0 + tau * tau;
    t = rcp_nr(at + h * rsq_nr(h));
  }
  if (tau < 0.0) t = -t;
  c = rsq_nr(1.0 + t * t);
  s = t * c;
}

// Cyclic parallel Jacobi on the np x np symmetric matrix a (LDS, ld SJLD), accumulating
// the rotations into v, for a block of NT threads (256 or 1024). Each step rotates np/2
// disjoint pairs (phase A, one thread per pair) and then updates every 2x2 block (rows of
// pair i, columns of pair j) as R_i^T A_ij R_j and the two columns of v per pair (phase
// B, over one combined item space). Returns 1 when the off-diagonal mass fell below
// tol^2 * diagonal mass. Pairs come from jacobi_pair.
// With TAB the per-thread items are kept in `tab` (an LDS array of 3072 entries) instead
// of in registers, which small blocks need to stay within their register budget.
template <int NT = SNT, bool TAB = false>
__device__ __forceinline__ int jacobi_sweeps(double* a, double* v, int n, int np, int max_sweeps, double tol,
                                             unsigned short* tab = nullptr) {
  constexpr int NW = NT / 64;
  constexpr int NI = (3072 + NT - 1) / NT;  // items per thread: (np/2)^2 + np * np/2 <= 3072
  static_assert(NI <= 12, "block too small");
  __shared__ double2 rcs[SJ / 2];
  __shared__ double red[2][NW];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int half = np / 2, nblk = half * half, nitem = nblk + np * half;
  // items of this thread (e = tid + NT u), packed as kind | x0 << 2 | x1 << 8:
  // kind 1 = 2x2 block (x0, x1) = (bi, bj); kind 2 = row x0 of v for pair x1; 0 = none
  auto item = [&](int e) {
    if (e < nblk) return 1 | ((e % half) << 2) | ((e / half) << 8);
    if (e < nitem) return 2 | (((e - nblk) % np) << 2) | (((e - nblk) / np) << 8);
    return 0;
  };
  int it[TAB ? 1 : NI];
  if constexpr (TAB) {
    for (int e = tid; e < nitem; e += NT) tab[e] = (unsigned short)item(e);
    __syncthreads();
  } else {
#pragma unroll
    for (int u = 0; u < NI; ++u) it[u] = item(tid + NT * u);
  }
  auto apply = [&](int packed, int step) {
    const int kind = packed & 3, x0 = (packed >> 2) & 63, x1 = packed >> 8;
    if (kind == 1) {
      const int bi = x0, bj = x1;
      int p, q, P, Q;
      jacobi_pair(step, bi, np, p, q);
      jacobi_pair(step, bj, np, P, Q);
      const double2 cs = rcs[bi], CS = rcs[bj];
      const double c = cs.x, s = cs.y, C = CS.x, S = CS.y;
      const double apP = a[p * SJLD + P], apQ = a[p * SJLD + Q];
      const double aqP = a[q * SJLD + P], aqQ = a[q * SJLD + Q];
      // right: columns (P, Q); left: rows (p, q)
      const double xpP = C * apP - S * apQ, xpQ = S * apP + C * apQ;
      const double xqP = C * aqP - S * aqQ, xqQ = S * aqP + C * aqQ;
      double ypP = c * xpP - s * xqP, yqP = s * xpP + c * xqP;
      double ypQ = c * xpQ - s * xqQ, yqQ = s * xpQ + c * xqQ;
      if (bi == bj && s != 0.0) {
        ypQ = 0.0;
        yqP = 0.0;
      }
      a[p * SJLD + P] = ypP;
      a[p * SJLD + Q] = ypQ;
      a[q * SJLD + P] = yqP;
      a[q * SJLD + Q] = yqQ;
    } else if (kind == 2) {
      const int r = x0, j = x1;
      int P, Q;
      jacobi_pair(step, j, np, P, Q);
      const double2 CS = rcs[j];
      const double vP = v[r * SJLD + P], vQ = v[r * SJLD + Q];
      v[r * SJLD + P] = CS.x * vP - CS.y * vQ;
      v[r * SJLD + Q] = CS.y * vP + CS.x * vQ;
    }
  };
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off = 0.0, dg = 0.0;
    for (int e = tid; e < n * n; e += NT) {
      const int i = e % n, j = e / n;
      const double x = a[i * SJLD + j];
      if (i == j) dg += x * x;
      else off += x * x;
    }
    off = wave_sum(off);
    dg = wave_sum(dg);
    if (lane == 0) {
      red[0][wv] = off;
      red[1][wv] = dg;
    }
    __syncthreads();
    off = 0.0;
    dg = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      off += red[0][w];
      dg += red[1][w];
    }
    __syncthreads();
    if (off <= tol * tol * dg || off == 0.0) return 1;
    for (int step = 0; step < np - 1; ++step) {
      if (tid < half) {
        int p, q;
        jacobi_pair(step, tid, np, p, q);
        double c, s;
        jacobi_rotation(a[p * SJLD + p], a[q * SJLD + q], a[p * SJLD + q], c, s);
        rcs[tid] = make_double2(c, s);
      }
      __syncthreads();
      if constexpr (TAB) {
#pragma unroll 1
        for (int e = tid; e < nitem; e += NT) apply(tab[e], step);
      } else {
#pragma unroll
        for (int u = 0; u < NI; ++u) apply(it[u], step);
      }
      __syncthreads();
    }
  }
  return 0;
}

// Cyclic Jacobi on a 16 x 16 symmetric matrix in the registers of one wave, for problems
// too small to amortise block-wide barriers. Lane l = 8 I + J holds the 2 x 2 blocks (rows
// 2I, 2I+1; columns 2J, 2J+1) of A and of the rotation product V (V starts as I). A step
// rotates the eight pairs (2i, 2i+1): the diagonal lanes form the rotations, every lane
// fetches those of its row and column pair with lane permutes and applies R_I^T A R_J and
// V R_J. Then the indices move one place along the ring t1 .. t7, b7 .. b0 of the circle
// round-robin (t0 fixed; ti = 2i, bi = 2i+1) by a row and a column permutation of A and a
// column permutation of V (rows from lanes l -+ 8, columns from lanes l -+ 1 by DPP row
// shifts), so fifteen steps pair every two indices once. Diagonal entries
// equal to `pad` (decoupled padding) are left out of the convergence test, which is the
// one of jacobi_sweeps. On return a and v (LDS, ld SJLD) hold the result in the slot
// order (columns of v follow the diagonal of a); returns 1 when converged.
__device__ __forceinline__ int jacobi16_wave(double* a, double* v, int max_sweeps, double tol, double pad) {
  const int l = threadIdx.x & 63, I = l >> 3, J = l & 7;
  const bool dgl = I == J;
  double a00 = a[(2 * I) * SJLD + 2 * J], a01 = a[(2 * I) * SJLD + 2 * J + 1];
  double a10 = a[(2 * I + 1) * SJLD + 2 * J], a11 = a[(2 * I + 1) * SJLD + 2 * J + 1];
  double v00 = dgl ? 1.0 : 0.0, v01 = 0.0, v10 = 0.0, v11 = dgl ? 1.0 : 0.0;
  const int up = l >= 8 ? l - 8 : l, dn = l < 56 ? l + 8 : l;
  int conv = 0;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off = dgl ? a01 * a01 + a10 * a10 : a00 * a00 + a01 * a01 + a10 * a10 + a11 * a11;
    double dg = dgl ? (a00 != pad ? a00 * a00 : 0.0) + (a11 != pad ? a11 * a11 : 0.0) : 0.0;
    // one lane's sums decide for the whole wave (butterfly sums can differ in the last bit)
    off = lane_bcast(wave_sum(off), 0);
    dg = lane_bcast(wave_sum(dg), 0);
    if (off <= tol * tol * dg || off == 0.0) {
      conv = 1;
      break;
    }
    for (int step = 0; step < 15; ++step) {
      double c, s;
      jacobi_rotation(a00, a11, a01, c, s);  // used from the diagonal lanes only
      const double cI = __shfl(c, 9 * I, 64), sI = __shfl(s, 9 * I, 64);
      const double cJ = __shfl(c, 9 * J, 64), sJ = __shfl(s, 9 * J, 64);
      const double x00 = cJ * a00 - sJ * a01, x01 = sJ * a00 + cJ * a01;
      const double x10 = cJ * a10 - sJ * a11, x11 = sJ * a10 + cJ * a11;
      a00 = cI * x00 - sI * x10;
      a10 = sI * x00 + cI * x10;
      a01 = cI * x01 - sI * x11;
      a11 = sI * x01 + cI * x11;
      if (dgl && sI != 0.0) {
        a01 = 0.0;
        a10 = 0.0;
      }
      const double w00 = cJ * v00 - sJ * v01, w01 = sJ * v00 + cJ * v01;
      const double w10 = cJ * v10 - sJ * v11, w11 = sJ * v10 + cJ * v11;
      // rows: new t_I = t_{I-1} (t_1 = b_0, t_0 stays); new b_I = b_{I+1} (b_7 = t_7)
      {
        const double s0 = I == 0 ? a10 : a00, s1 = I == 0 ? a11 : a01;
        const double u00 = __shfl(s0, up, 64), u01 = __shfl(s1, up, 64);
        const double d10 = __shfl(a10, dn, 64), d11 = __shfl(a11, dn, 64);
        const double n10 = I == 7 ? a00 : d10, n11 = I == 7 ? a01 : d11;
        if (I != 0) {
          a00 = u00;
          a01 = u01;
        }
        a10 = n10;
        a11 = n11;
      }
      // columns, the same ring (A and V)
      {
        const double s0 = J == 0 ? a01 : a00, s1 = J == 0 ? a11 : a10;
        const double l00 = dpp_row_shift<DPP_ROW_SR1>(s0), l10 = dpp_row_shift<DPP_ROW_SR1>(s1);
        const double r01 = dpp_row_shift<DPP_ROW_SL1>(a01), r11 = dpp_row_shift<DPP_ROW_SL1>(a11);
        const double n01 = J == 7 ? a00 : r01, n11 = J == 7 ? a10 : r11;
        if (J != 0) {
          a00 = l00;
          a10 = l10;
        }
        a01 = n01;
        a11 = n11;
      }
      {
        const double s0 = J == 0 ? w01 : w00, s1 = J == 0 ? w11 : w10;
        const double l00 = dpp_row_shift<DPP_ROW_SR1>(s0), l10 = dpp_row_shift<DPP_ROW_SR1>(s1);
        const double r01 = dpp_row_shift<DPP_ROW_SL1>(w01), r11 = dpp_row_shift<DPP_ROW_SL1>(w11);
        v00 = J == 0 ? w00 : l00;
        v10 = J == 0 ? w10 : l10;
        v01 = J == 7 ? w00 : r01;
        v11 = J == 7 ? w10 : r11;
      }
    }
  }
  a[(2 * I) * SJLD + 2 * J] = a00;
  a[(2 * I) * SJLD + 2 * J + 1] = a01;
  a[(2 * I + 1) * SJLD + 2 * J] = a10;
  a[(2 * I + 1) * SJLD + 2 * J + 1] = a11;
  v[(2 * I) * SJLD + 2 * J] = v00;
  v[(2 * I) * SJLD + 2 * J + 1] = v01;
  v[(2 * I + 1) * SJLD + 2 * J] = v10;
  v[(2 * I + 1) * SJLD + 2 * J + 1] = v11;
  return conv;
}

// Load an n x n matrix into a (padded to even np with a decoupled, large negative
// diagonal entry that sorts last), set v = I.
__device__ __forceinline__ void jacobi_load(double* a, double* v, const double* A, int lda, int n, int np, bool symmetrise) {
  __shared__ double red[SW];
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

// Descending order of the diagonal of a: thread i places index i at its rank (ties by
// index), so order is a permutation.
__device__ __forceinline__ void jacobi_order(const double* a, int np, int* order) {
  const int i = threadIdx.x;
  if (i < np) {
    const double di = a[i * SJLD + i];
    int rank = 0;
    for (int j = 0; j < np; ++j) {
      const double dj = a[j * SJLD + j];
      rank += (dj > di || (dj == di && j < i)) ? 1 : 0;
    }
    order[rank] = i;
  }
  __syncthreads();
}

}  // namespace ht
