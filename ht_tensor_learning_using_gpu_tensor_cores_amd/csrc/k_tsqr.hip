// TSQR of tall-skinny FP32 blocks (m x b, b <= 64) as a Householder reduction tree.
//
// The reference's tsqr_svd splits A into 4 row blocks on the host with a CPU loop over
// every element, QRs each block after a host-to-device copy, combines [R1;R2], [R3;R4]
// and the root, then assembles Q with GEMMs (third_order_tensor_decompsition/CUDA/TSQR/
// kernel.cu:500-677). Here the whole tree runs on the device:
//   leaves  -- balanced blocks of 2b..4b rows, one work-group each, Householder QR with
//              FP64 arithmetic in LDS sized to the block; the reflectors overwrite the
//              block in W (FP32), R goes up;
//   tree    -- 4-ary: the stacked R factors of up to four children (<= 4b x b) are
//              QR-factored in FP64 by one work-group; the root R is FP64;
//   Q       -- formed explicitly top-down: each node applies its reflectors to [E; 0]
//              with E the b x b block its parent handed down, leaves write Q rows.
// Thread layout inside a block QR: lane = column c, wave q = quarter of the rows; a node's
// LDS is 4b x (b|1) doubles (52 KB at b = 40), so several work-groups share a CU.
#include "k_hh.h"

namespace ht {

constexpr int TQ_FAN = 4;

struct TsqrLevel {
  long n_nodes;
  double* R;    // [n_nodes][b*b] (column-major b x b)
  double* V;    // [n_nodes][4b * ld] reflectors (row-major, stride ld), tree levels only
  double* tau;  // [n_nodes][b]
  double* E;    // [n_nodes][b*b] block handed down by the parent (column-major)
  int* rows;    // [n_nodes] stacked rows of the node
};

// Dynamic LDS of one node: the block (4b x ld doubles), 4 x 64 partials, tau, 4 norms.
static size_t tq_lds_bytes(int b) { return ((size_t)TQ_FANROWS * b * (b | 1) + 4 * TQ_B + TQ_B + 4) * sizeof(double); }

// Leaf QR. With n_leaf = ceil(m / 4b) blocks, block i holds rows [i m / n_leaf,
// (i+1) m / n_leaf): at most 4b and, when n_leaf > 1, at least 2b rows, so every reflector
// lives in rows the block owns and is stored in place in W (FP32). A single block (m <= 4b)
// keeps its reflectors in FP64 (V1, needed when m < b).
__global__ __launch_bounds__(256) void tsqr_leaf_kernel(float* __restrict__ W, long m, int b, long n_leaf,
                                                        double* __restrict__ Rout, double* __restrict__ tauout,
                                                        double* __restrict__ V1) {
  extern __shared__ double lds[];
  const int ld = tq_ld(b), cap = TQ_FANROWS * b;
  double* A = lds;
  double* red = A + (long)cap * ld;
  double* tau = red + 4 * TQ_B;
  double* nrm = tau + TQ_B;
  const int tid = threadIdx.x;
  const long r0 = (long)blockIdx.x * m / n_leaf;
  const long r1 = (long)(blockIdx.x + 1) * m / n_leaf;
  const int nr = (int)(r1 - r0);
  const int rows = max(nr, b);
  // column-major FP32 in global: consecutive threads take consecutive rows of one column
  for (int e = tid; e < rows * b; e += 256) {
    const int r = e % rows, c = e / rows;
    A[r * ld + c] = r < nr ? (double)W[r0 + r + m * c] : 0.0;
  }
  __syncthreads();
  block_householder(A, tau, red, nrm, rows, b, ld);
  double* R = Rout + (long)blockIdx.x * b * b;
  for (int e = tid; e < b * b; e += 256) {
    const int i = e % b, j = e / b;
    R[e] = (i <= j) ? A[i * ld + j] : 0.0;
  }
  if (n_leaf == 1) {
    for (int e = tid; e < rows * ld; e += 256) V1[e] = A[e];
  } else {
    for (int e = tid; e < nr * b; e += 256) {
      const int r = e % nr, c = e / nr;
      W[r0 + r + m * c] = (float)A[r * ld + c];
    }
  }
  if (tid < b) tauout[(long)blockIdx.x * b + tid] = tau[tid];
}

// Tree node QR: node i stacks the R factors of children FAN*i .. FAN*i+FAN-1.
__global__ __launch_bounds__(256) void tsqr_tree_kernel(const double* __restrict__ Rchild, long n_child, int b,
                                                        double* __restrict__ Rout, double* __restrict__ Vout,
                                                        double* __restrict__ tauout, int* __restrict__ rows_out) {
  extern __shared__ double lds[];
  const int ld = tq_ld(b), cap = TQ_FANROWS * b;
  double* A = lds;
  double* red = A + (long)cap * ld;
  double* tau = red + 4 * TQ_B;
  double* nrm = tau + TQ_B;
  const int tid = threadIdx.x;
  const long c0 = (long)blockIdx.x * TQ_FAN;
  const int nch = (int)min((long)TQ_FAN, n_child - c0);
  const int rows = nch * b;
  for (int e = tid; e < rows * b; e += 256) {
    const int r = e / b, c = e % b;
    const int ch = r / b, i = r % b;
    A[r * ld + c] = Rchild[(c0 + ch) * b * b + i + (long)b * c];
  }
  __syncthreads();
  block_householder(A, tau, red, nrm, rows, b, ld);
  double* R = Rout + (long)blockIdx.x * b * b;
  for (int e = tid; e < b * b; e += 256) {
    const int i = e % b, j = e / b;
    R[e] = (i <= j) ? A[i * ld + j] : 0.0;
  }
  double* V = Vout + (long)blockIdx.x * cap * ld;
  for (int e = tid; e < rows * ld; e += 256) V[e] = A[e];
  if (tid < b) tauout[(long)blockIdx.x * b + tid] = tau[tid];
  if (tid == 0) rows_out[blockIdx.x] = rows;
}

// Top-down through a tree level: node i receives E_i (b x b) and hands the blocks of
// H_i [E_i; 0] to its children.
__global__ __launch_bounds__(256) void tsqr_down_kernel(const double* __restrict__ V, const double* __restrict__ tauv,
                                                        const int* __restrict__ rows_in, const double* __restrict__ E,
                                                        int b, int identity, double* __restrict__ Echild) {
  extern __shared__ double lds[];
  const int ld = tq_ld(b), cap = TQ_FANROWS * b;
  double* M = lds;
  double* red = M + (long)cap * ld;
  double* tau = red + 4 * TQ_B;
  const int tid = threadIdx.x;
  const long node = blockIdx.x;
  const int rows = rows_in[node];
  for (int e = tid; e < rows * b; e += 256) {
    const int r = e / b, c = e % b;
    double val = 0.0;
    if (r < b) val = identity ? (r == c ? 1.0 : 0.0) : E[node * b * b + r + (long)b * c];
    M[r * ld + c] = val;
  }
  if (tid < b) tau[tid] = tauv[node * b + tid];
  __syncthreads();
  apply_q(VGlobal64{V + node * cap * ld, ld}, tau, M, red, rows, b, ld);
  const int nch = rows / b;
  for (int e = tid; e < nch * b * b; e += 256) {
    const int ch = e / (b * b), rem = e % (b * b);
    const int i = rem % b, c = rem / b;
    Echild[(node * TQ_FAN + ch) * b * b + i + (long)b * c] = M[(ch * b + i) * ld + c];
  }
}

// Leaves: Q rows = H_leaf [E_leaf; 0].
__global__ __launch_bounds__(256) void tsqr_leafq_kernel(const float* __restrict__ W, long m, int b, long n_leaf,
                                                         const double* __restrict__ tauv, const double* __restrict__ E,
                                                         int identity, const double* __restrict__ V1,
                                                         float* __restrict__ Q) {
  extern __shared__ double lds[];
  const int ld = tq_ld(b), cap = TQ_FANROWS * b;
  double* M = lds;
  double* red = M + (long)cap * ld;
  double* tau = red + 4 * TQ_B;
  const int tid = threadIdx.x;
  const long r0 = (long)blockIdx.x * m / n_leaf;
  const long r1 = (long)(blockIdx.x + 1) * m / n_leaf;
  const int nr = (int)(r1 - r0);
  const int rows = max(nr, b);
  for (int e = tid; e < rows * b; e += 256) {
    const int r = e / b, c = e % b;
    double val = 0.0;
    if (r < b) val = identity ? (r == c ? 1.0 : 0.0) : E[(long)blockIdx.x * b * b + r + (long)b * c];
    M[r * ld + c] = val;
  }
  if (tid < b) tau[tid] = tauv[(long)blockIdx.x * b + tid];
  __syncthreads();
  if (n_leaf == 1) apply_q(VGlobal64{V1, ld}, tau, M, red, rows, b, ld);
  else apply_q(VLeaf32{W, r0, m}, tau, M, red, rows, b, ld);
  for (int e = tid; e < nr * b; e += 256) {
    const int r = e % nr, c = e / nr;
    Q[r0 + r + m * c] = (float)M[r * ld + c];
  }
}

static long ceil_div(long a, long b) { return (a + b - 1) / b; }

size_t tsqr_ws_bytes(long m, int b) {
  if (b < 1) b = 1;
  const long cap = (long)TQ_FANROWS * b, ld = b | 1;
  size_t bytes = 0;
  long n = ceil_div(m, cap);
  bytes += (size_t)n * (b * b + b + b * b) * sizeof(double) + 256 * 3;  // leaf R, tau, E (+ alignment)
  bytes += (size_t)cap * ld * sizeof(double) + 256;                      // single-leaf reflectors
  while (n > 1) {
    n = ceil_div(n, TQ_FAN);
    bytes += (size_t)n * ((size_t)b * b * 2 + b + (size_t)cap * ld) * sizeof(double) + n * sizeof(int) + 5 * 256;
  }
  return bytes + 4096;
}

int tsqr_launch(hipStream_t st, float* W, long m, int b, float* Q, double* Rfinal, void* ws) {
  if (b < 1 || b > TQ_B || m < 1) return 8;
  char* p = (char*)ws;
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 255) & ~(size_t)255;
    return r;
  };
  const long cap = (long)TQ_FANROWS * b, ld = b | 1;
  const size_t lds = tq_lds_bytes(b);
  static bool attr_set = false;  // the widest block needs more dynamic LDS than the default
  if (!attr_set) {
    const int mx = (int)tq_lds_bytes(TQ_B);
    if (hipFuncSetAttribute((const void*)tsqr_leaf_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, mx) ||
        hipFuncSetAttribute((const void*)tsqr_tree_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, mx) ||
        hipFuncSetAttribute((const void*)tsqr_down_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, mx) ||
        hipFuncSetAttribute((const void*)tsqr_leafq_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, mx))
      return 5;
    attr_set = true;
  }
  const long n_leaf = ceil_div(m, cap);
  double* leafR = (double*)take((size_t)n_leaf * b * b * sizeof(double));
  double* leafTau = (double*)take((size_t)n_leaf * b * sizeof(double));
  double* leafE = (double*)take((size_t)n_leaf * b * b * sizeof(double));
  double* V1 = (double*)take((size_t)cap * ld * sizeof(double));
  TsqrLevel lv[24];
  int nl = 0;
  long n = n_leaf;
  while (n > 1) {
    const long nn = ceil_div(n, TQ_FAN);
    TsqrLevel& L = lv[nl++];
    L.n_nodes = nn;
    L.R = (double*)take((size_t)nn * b * b * sizeof(double));
    L.V = (double*)take((size_t)nn * cap * ld * sizeof(double));
    L.tau = (double*)take((size_t)nn * b * sizeof(double));
    L.E = (double*)take((size_t)nn * b * b * sizeof(double));
    L.rows = (int*)take((size_t)nn * sizeof(int));
    n = nn;
    if (nl >= 24) return 8;
  }
  hipLaunchKernelGGL(tsqr_leaf_kernel, dim3((unsigned)n_leaf), dim3(256), lds, st, W, m, b, n_leaf, leafR, leafTau,
                     V1);
  const double* Rc = leafR;
  long nc = n_leaf;
  for (int i = 0; i < nl; ++i) {
    hipLaunchKernelGGL(tsqr_tree_kernel, dim3((unsigned)lv[i].n_nodes), dim3(256), lds, st, Rc, nc, b, lv[i].R,
                       lv[i].V, lv[i].tau, lv[i].rows);
    Rc = lv[i].R;
    nc = lv[i].n_nodes;
  }
  // Root R.
  if (Rfinal) hipMemcpyAsync(Rfinal, Rc, (size_t)b * b * sizeof(double), hipMemcpyDeviceToDevice, st);
  // Down the tree.
  for (int i = nl - 1; i >= 0; --i) {
    double* Echild = (i == 0) ? leafE : lv[i - 1].E;
    hipLaunchKernelGGL(tsqr_down_kernel, dim3((unsigned)lv[i].n_nodes), dim3(256), lds, st, lv[i].V, lv[i].tau,
                       lv[i].rows, lv[i].E, b, (i == nl - 1) ? 1 : 0, Echild);
  }
  hipLaunchKernelGGL(tsqr_leafq_kernel, dim3((unsigned)n_leaf), dim3(256), lds, st, W, m, b, n_leaf, leafTau, leafE,
                     nl == 0 ? 1 : 0, V1, Q);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

}  // namespace ht
