// Batched HT of many small, unevenly sized 3rd-order tensors in one persistent launch.
//
// Tree {1},{2,3} (node 0 root, 1 = {1}, 2 = {2,3}, 3 = {2}, 4 = {3}) with the reference's
// plain semantics (every non-root node's basis is the exact truncated SVD of its own
// matricisation: third_order_tensor_decompsition/CUDA/Baseline/htd_unopt.cu:65-88):
//   U1, U2, U3   top eigenvectors of the leaf Grams X_(m) X_(m)^T (FP32 MFMA Grams over
//                LDS-staged slabs and fibre chunks, FP64 Jacobi);
//   U23          the complementary node's basis = right singular vectors of X_(1),
//                X_(1)^T U1 diag(1/sigma), formed slab by slab and never stored;
//   B23          (U2 (x) U3)^T U23 (k2 x k3 x k23), B_root = U1^T X_(1) (U2 (x) U3) B23;
//   rel. error   ||X - X_HT||^2 = ||X||^2 - 2 ||B_root||^2 + tr(B_root B23^T B23 B_root^T)
//                (U1 and U2 (x) U3 orthonormal), all terms in FP64.
// One work-group of 256 threads per tensor; the grid is persistent (blocks stride over
// a host-sorted order, largest tensors first). n <= 64, ranks <= 8.
#include "k_jacobi.h"

namespace ht {

constexpr int BT_NT = 256;                 // threads per work-group
constexpr int BT_RMAX = 8;                 // largest rank
constexpr int BT_NMAX = 64;                // largest mode size
constexpr int BT_SLD = BT_NMAX + 1;        // LDS leading dimension of staged slabs/chunks
constexpr int BT_SCR = 3 * 64 * 64 + 3 * 64 * BT_RMAX;  // doubles of scratch per block

// LDS plan (bytes): phase A (Grams) 2 x 64 x 65 floats; phase B (Jacobi) a, v 2 x 64 x 65
// doubles + item table; phase C (projections) slab + U1..U3 + T, V + W + B23 accumulators.
constexpr int BT_LDS_DOUBLES = 2 * SJ * SJLD + 1024;
// Jacobi item table (3072 uint16) right after a and v, inside the same LDS array.
static_assert(2 * SJ * SJLD * 8 + 3072 * 2 <= BT_LDS_DOUBLES * 8, "LDS plan");

template <int KC>
__global__ __launch_bounds__(BT_NT, 2) void batch_ht_kernel(const float* __restrict__ x, const BatchItem* __restrict__ items,
                                                         const int* __restrict__ order, int count,
                                                         float* __restrict__ out, double* __restrict__ rel_err,
                                                         double* __restrict__ scratch) {
  __shared__ double lds[BT_LDS_DOUBLES];
  __shared__ int ordj[SJ];
  __shared__ double red[2][4];
  __shared__ double sgn_s[BT_RMAX];
  __shared__ double best_s[BT_RMAX];
  __shared__ int besti_s[BT_RMAX];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double* G = scratch + (long)blockIdx.x * BT_SCR;  // G1, G2, G3 (64 x 64 each)
  double* Uc = G + 3 * 64 * 64;                      // U1 (64 x 8, top max(k1, k23)), U2, U3
  for (int it = blockIdx.x; it < count; it += gridDim.x) {
    const BatchItem d = items[order[it]];
    const int n1 = d.n[0], n2 = d.n[1], n3 = d.n[2];
    const int k1 = d.k[1], k23 = d.k[2], k2 = d.k[3], k3 = d.k[4];
    const int k1x = max(k1, k23);
    const float* X = x + d.x_off;
    float* O = out + d.out_off;
    const long slab = (long)n1 * n2;
    // ---------------- phase A: Grams (FP32 MFMA 32x32x2, accumulators per wave tile)
    float* S = reinterpret_cast<float*>(lds);  // two buffers of 64 x 65 floats
    const int ti = w >> 1, tj = w & 1;         // this wave's 32 x 32 tile of each Gram
    const int li = lane & 31, lk = lane >> 5;
    f32x16 g1 = f32x16{}, g2 = f32x16{};
    double xx = 0.0;
    auto load_slab = [&](int i3, float* dst) {
      const float* src = X + slab * i3;
      for (int e = tid; e < n1 * n2; e += BT_NT) {
        const int i1 = e % n1, i2 = e / n1;
        const float v = src[e];
        dst[i1 + BT_SLD * i2] = v;
        xx += (double)v * v;
      }
    };
    __syncthreads();
    load_slab(0, S);
    __syncthreads();
    for (int i3 = 0; i3 < n3; ++i3) {
      float* cur = S + (i3 & 1) * (BT_SLD * 64);
      float* nxt = S + ((i3 + 1) & 1) * (BT_SLD * 64);
      if (i3 + 1 < n3) load_slab(i3 + 1, nxt);
      // G1 += S S^T (rows i1, k = i2); G2 += S^T S (rows i2, k = i1)
      {
        const int ra = 32 * ti + li, rb = 32 * tj + li;
        if (32 * ti < n1 && 32 * tj < n1) {
          for (int k0 = 0; k0 < n2; k0 += 2) {
            const int kk = k0 + lk;
            const float a = (ra < n1 && kk < n2) ? cur[ra + BT_SLD * kk] : 0.f;
            const float b = (rb < n1 && kk < n2) ? cur[rb + BT_SLD * kk] : 0.f;
            g1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, g1, 0, 0, 0);
          }
        }
        if (32 * ti < n2 && 32 * tj < n2) {
          for (int k0 = 0; k0 < n1; k0 += 2) {
            const int kk = k0 + lk;
            const float a = (ra < n2 && kk < n1) ? cur[kk + BT_SLD * ra] : 0.f;
            const float b = (rb < n2 && kk < n1) ? cur[kk + BT_SLD * rb] : 0.f;
            g2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, g2, 0, 0, 0);
          }
        }
      }
      __syncthreads();
    }
    // G3 from fibre chunks: C[k][i3] = X[k + n1 n2 i3], k in a chunk of KC positions
    f32x16 g3 = f32x16{};
    {
      float* C = S;  // KC x n3 with leading dimension KC + 1 per column i3
      const int cld = KC + 1;
      for (long k0 = 0; k0 < slab; k0 += KC) {
        const int kn = (int)min((long)KC, slab - k0);
        for (int e = tid; e < KC * n3; e += BT_NT) {
          const int kk = e % KC, i3 = e / KC;
          C[kk + cld * i3] = kk < kn ? X[k0 + kk + slab * i3] : 0.f;
        }
        __syncthreads();
        const int ra = 32 * ti + li, rb = 32 * tj + li;
        if (32 * ti < n3 && 32 * tj < n3) {
          for (int q = 0; q < KC; q += 2) {
            const int kk = q + lk;
            const float a = ra < n3 ? C[kk + cld * ra] : 0.f;
            const float b = rb < n3 ? C[kk + cld * rb] : 0.f;
            g3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, g3, 0, 0, 0);
          }
        }
        __syncthreads();
      }
    }
    // write the Gram tiles (FP64) to scratch, and ||X||^2
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int r = 32 * ti + acc32_row(e, lane), c = 32 * tj + li;
      G[r + 64 * c] = g1[e];
      G[64 * 64 + r + 64 * c] = g2[e];
      G[2 * 64 * 64 + r + 64 * c] = g3[e];
    }
    xx = wave_sum(xx);
    if (lane == 0) red[0][w] = xx;
    __syncthreads();
    const double xsq = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    __threadfence_block();
    // ---------------- phase B: Jacobi per leaf Gram, top eigenvectors (canonical signs)
    double* a = lds;
    double* v = lds + SJ * SJLD;
    for (int m = 0; m < 3; ++m) {
      const int n = m == 0 ? n1 : (m == 1 ? n2 : n3), np = (n + 1) & ~1;
      const int km = m == 0 ? k1x : (m == 1 ? k2 : k3);
      __syncthreads();
      jacobi_load(a, v, G + m * 64 * 64, 64, n, np, true);
      jacobi_sweeps<BT_NT, true>(a, v, n, np, 30, 1e-14,
                           reinterpret_cast<unsigned short*>(lds + 2 * SJ * SJLD));
      jacobi_order(a, np, ordj);
      // canonical signs of the first km columns; one wave per column
      for (int j = w; j < km; j += 4) {
        const int src = ordj[j];
        double best = -1.0;
        int bi = 0;
        for (int r = lane; r < n; r += 64) {
          const double xv = fabs(v[r * SJLD + src]);
          if (xv > best) { best = xv; bi = r; }
        }
        for (int o = 32; o > 0; o >>= 1) {
          const double ob = __shfl_xor(best, o, 64);
          const int oi = __shfl_xor(bi, o, 64);
          if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        const double sg = v[bi * SJLD + src] < 0.0 ? -1.0 : 1.0;
        for (int r = lane; r < n; r += 64) Uc[m * 64 * BT_RMAX + r + 64 * j] = sg * v[r * SJLD + src];
      }
      // outputs: U_t (FP32) and the node's singular values
      const int leaf = m == 0 ? 1 : (m == 1 ? 3 : 4);
      const int kl = m == 0 ? k1 : km;
      __syncthreads();
      for (int e = tid; e < n * kl; e += BT_NT) {
        const int r = e % n, j = e / n;
        O[d.u_off[leaf] + e] = (float)Uc[m * 64 * BT_RMAX + r + 64 * j];
      }
      if (tid < kl) O[d.s_off[leaf] + tid] = (float)sqrt(fmax(a[ordj[tid] * SJLD + ordj[tid]], 0.0));
      if (m == 0 && tid < k23) O[d.s_off[2] + tid] = (float)sqrt(fmax(a[ordj[tid] * SJLD + ordj[tid]], 0.0));
      if (m == 0 && tid < k23) {
        const double s = sqrt(fmax(a[ordj[tid] * SJLD + ordj[tid]], 0.0));
        best_s[tid] = s;  // sigma of node {2,3} (reused below)
      }
      __syncthreads();
    }
    // ---------------- phase C: U23 on the fly, W = X x2 U2^T x3 U3^T, B23 accumulators
    float* Sl = reinterpret_cast<float*>(lds);                 // n1 x n2 slab, ld BT_SLD
    double* U1 = lds + (BT_SLD * 64 + 1) / 2 + 1;              // 64 x 8
    double* U2 = U1 + 64 * BT_RMAX;
    double* U3 = U2 + 64 * BT_RMAX;
    double* T = U3 + 64 * BT_RMAX;                              // n1 x k2
    double* V = T + 64 * BT_RMAX;                               // n2 x k23 (U23 rows x sigma)
    double* E = V + 64 * BT_RMAX;                               // k2 x k23
    double* Wc = E + BT_RMAX * BT_RMAX;                         // n1 x k2 x k3
    double* B23 = Wc + 64 * BT_RMAX * BT_RMAX;                  // k2 x k3 x k23
    double* sig = B23 + BT_RMAX * BT_RMAX * BT_RMAX;            // k23
    for (int e = tid; e < 64 * BT_RMAX; e += BT_NT) {
      U1[e] = Uc[e];
      U2[e] = Uc[64 * BT_RMAX + e];
      U3[e] = Uc[2 * 64 * BT_RMAX + e];
    }
    for (int e = tid; e < n1 * k2 * k3; e += BT_NT) Wc[e] = 0.0;
    for (int e = tid; e < k2 * k3 * k23; e += BT_NT) B23[e] = 0.0;
    if (tid < k23) sig[tid] = best_s[tid];
    // running argmax |U23[:, c]| (row = i2 + n2 i3, first occurrence) for the sign of U23
    if (tid < BT_RMAX) { best_s[tid] = -1.0; besti_s[tid] = 0; sgn_s[tid] = 1.0; }
    __syncthreads();
    for (int i3 = 0; i3 < n3; ++i3) {
      const float* src = X + slab * i3;
      for (int e = tid; e < n1 * n2; e += BT_NT) Sl[(e % n1) + BT_SLD * (e / n1)] = src[e];
      __syncthreads();
      // T[i1][a] = sum_i2 S[i1][i2] U2[i2][a];  V[i2][c] = sum_i1 S[i1][i2] U1[i1][c]
      for (int e = tid; e < n1 * k2 + n2 * k23; e += BT_NT) {
        if (e < n1 * k2) {
          const int i1 = e % n1, aa = e / n1;
          double t = 0.0;
          for (int i2 = 0; i2 < n2; ++i2) t += (double)Sl[i1 + BT_SLD * i2] * U2[i2 + 64 * aa];
          T[i1 + 64 * aa] = t;
        } else {
          const int f = e - n1 * k2, i2 = f % n2, c = f / n2;
          double t = 0.0;
          for (int i1 = 0; i1 < n1; ++i1) t += (double)Sl[i1 + BT_SLD * i2] * U1[i1 + 64 * c];
          V[i2 + 64 * c] = t;
        }
      }
      __syncthreads();
      // W[i1][a][b] += T[i1][a] U3[i3][b];  E[a][c] = sum_i2 U2[i2][a] V[i2][c]
      for (int e = tid; e < n1 * k2 * k3 + k2 * k23; e += BT_NT) {
        if (e < n1 * k2 * k3) {
          const int i1 = e % n1, ab = e / n1, aa = ab % k2, bb = ab / k2;
          Wc[e] += T[i1 + 64 * aa] * U3[i3 + 64 * bb];
        } else {
          const int f = e - n1 * k2 * k3, aa = f % k2, c = f / k2;
          double t = 0.0;
          for (int i2 = 0; i2 < n2; ++i2) t += U2[i2 + 64 * aa] * V[i2 + 64 * c];
          E[aa + BT_RMAX * c] = t;
        }
      }
      // sign tracking of U23 columns: max |V[i2][c]| over this slab (one thread per column)
      if (tid < k23) {
        double bv = best_s[tid];
        int bi = besti_s[tid];
        double bs = sgn_s[tid];
        for (int i2 = 0; i2 < n2; ++i2) {
          const double val = V[i2 + 64 * tid];
          if (fabs(val) > bv) { bv = fabs(val); bi = i2 + n2 * i3; bs = val < 0.0 ? -1.0 : 1.0; }
        }
        best_s[tid] = bv;
        besti_s[tid] = bi;
        sgn_s[tid] = bs;
      }
      __syncthreads();
      // B23[a][b][c] += E[a][c] U3[i3][b] / sigma_c
      for (int e = tid; e < k2 * k3 * k23; e += BT_NT) {
        const int aa = e % k2, bc = e / k2, bb = bc % k3, c = bc / k3;
        const double s = sig[c];
        B23[e] += s > 0.0 ? E[aa + BT_RMAX * c] * U3[i3 + 64 * bb] / s : 0.0;
      }
      __syncthreads();
    }
    // canonical sign of U23 columns applied to B23 (and so to B_root)
    for (int e = tid; e < k2 * k3 * k23; e += BT_NT) B23[e] *= sgn_s[e / (k2 * k3)];
    __syncthreads();
    // Y[i1][c] = sum_p W[i1][p] B23[p][c] (p = a + k2 b), B_root = U1^T Y (k1 x k23)
    double* Y = T;  // n1 x k23 (reuses T)
    for (int e = tid; e < n1 * k23; e += BT_NT) {
      const int i1 = e % n1, c = e / n1;
      double t = 0.0;
      for (int p = 0; p < k2 * k3; ++p) t += Wc[i1 + n1 * p] * B23[p + k2 * k3 * c];
      Y[i1 + 64 * c] = t;
    }
    __syncthreads();
    double* Br = E;  // k1 x k23
    for (int e = tid; e < k1 * k23; e += BT_NT) {
      const int aa = e % k1, c = e / k1;
      double t = 0.0;
      for (int i1 = 0; i1 < n1; ++i1) t += U1[i1 + 64 * aa] * Y[i1 + 64 * c];
      Br[aa + BT_RMAX * c] = t;
    }
    __syncthreads();
    // outputs B_root (k1 x k23 x 1), B23 (k2 x k3 x k23); error terms
    for (int e = tid; e < k1 * k23; e += BT_NT) O[d.b_off[0] + e] = (float)Br[(e % k1) + BT_RMAX * (e / k1)];
    for (int e = tid; e < k2 * k3 * k23; e += BT_NT) O[d.b_off[2] + e] = (float)B23[e];
    if (tid == 0) {
      double nb = 0.0, tr = 0.0;
      for (int e = 0; e < k1 * k23; ++e) {
        const double bv = Br[(e % k1) + BT_RMAX * (e / k1)];
        nb += bv * bv;
      }
      // tr(Br M Br^T), M = B23mat^T B23mat (k23 x k23)
      for (int c1 = 0; c1 < k23; ++c1)
        for (int c2 = 0; c2 < k23; ++c2) {
          double mcc = 0.0;
          for (int p = 0; p < k2 * k3; ++p) mcc += B23[p + k2 * k3 * c1] * B23[p + k2 * k3 * c2];
          double bb = 0.0;
          for (int aa = 0; aa < k1; ++aa) bb += Br[aa + BT_RMAX * c1] * Br[aa + BT_RMAX * c2];
          tr += mcc * bb;
        }
      const double num = fmax(xsq - 2.0 * nb + tr, 0.0);
      rel_err[order[it]] = xsq > 0.0 ? sqrt(num / xsq) : 0.0;
    }
    __syncthreads();
  }
}

size_t batch_scratch_bytes(int blocks) { return (size_t)blocks * BT_SCR * sizeof(double); }

int batch_ht_launch(hipStream_t st, const float* x, const void* items, const int* order, int count, float* out,
                    double* rel_err, double* scratch, int blocks, int kc) {
  if (count < 1 || blocks < 1) return 8;
  if (kc == 32)
    hipLaunchKernelGGL(batch_ht_kernel<32>, dim3(blocks), dim3(BT_NT), 0, st, x, (const BatchItem*)items, order,
                       count, out, rel_err, scratch);
  else
    hipLaunchKernelGGL(batch_ht_kernel<64>, dim3(blocks), dim3(BT_NT), 0, st, x, (const BatchItem*)items, order,
                       count, out, rel_err, scratch);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

}  // namespace ht
