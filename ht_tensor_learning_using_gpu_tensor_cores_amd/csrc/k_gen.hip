// Synthetic-input kernels: Philox Gaussian noise and the Gaussian-bump family.
// The Tucker part X0 = core x_n U_n is a low-rank product built with the lowrank kernel.
// These replace genHtensor (third_order_tensor_decompsition/CUDA/Baseline/kernel.cu:94-202),
// whose srand(time(NULL)) input changes from run to run.
#include "ht_common.h"

namespace ht {

__global__ void noise_sumsq_kernel(long n, uint64_t seed, uint32_t stream, double* __restrict__ part) {
  __shared__ double red[4];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double z = philox_normal(seed, stream, (uint64_t)i);
    s += z * z;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void add_noise_kernel(float* __restrict__ x, long n, uint64_t seed, uint32_t stream, double scale) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = (float)((double)x[i] + scale * philox_normal(seed, stream, (uint64_t)i));
}

// x[i1 + n i2 + ...] = sum_j amp[j] prod_m factors[(j*order + m)*n + i_m]
// Generic form (any order, any n): one thread per element.
__global__ void bumps_kernel(float* __restrict__ x, int order, long n, const float* __restrict__ factors,
                             const float* __restrict__ amp, int n_bumps) {
  long total = 1;
  for (int m = 0; m < order; ++m) total *= n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    long idx[8];
    long r = e;
    for (int m = 0; m < order; ++m) { idx[m] = r % n; r /= n; }
    double s = 0.0;
    for (int j = 0; j < n_bumps; ++j) {
      double t = amp[j];
      for (int m = 0; m < order; ++m) t *= factors[((long)j * order + m) * n + idx[m]];
      s += t;
    }
    x[e] = (float)s;
  }
}

// Plane form (order >= 2, n <= BP_NMAX, n_bumps <= BP_JMAX): one work-group per plane
// (i_2, ..., i_{d-1}) of the two leading modes. The plane's outer factor
// amp_j prod_{m >= 2} g_jm(i_m) is formed once, the two leading factors sit in LDS, and
// the n x n plane is written with 16-byte stores (n % 4 == 0) or 4-byte stores. When part
// is not null, part[plane] = sum of the squares of the written (FP32) values.
constexpr int BP_NMAX = 1024, BP_JMAX = 8;
__global__ __launch_bounds__(256) void bumps_plane_kernel(float* __restrict__ x, int order, int n,
                                                          const float* __restrict__ factors,
                                                          const float* __restrict__ amp, int n_bumps, long planes,
                                                          double* __restrict__ part) {
  __shared__ float g0[BP_JMAX * BP_NMAX];
  __shared__ float g1[BP_JMAX * BP_NMAX];
  __shared__ double outer[BP_JMAX];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  for (int e = tid; e < n_bumps * n; e += 256) {
    const int j = e / n, i = e % n;
    g0[j * n + i] = factors[((long)j * order + 0) * n + i];
    g1[j * n + i] = factors[((long)j * order + 1) * n + i];
  }
  const long nn = (long)n * n;
  for (long plane = blockIdx.x; plane < planes; plane += gridDim.x) {
    __syncthreads();
    if (tid < n_bumps) {
      double t = amp[tid];
      long r = plane;
      for (int m = 2; m < order; ++m) {
        const int im = (int)(r % n);
        r /= n;
        t *= factors[((long)tid * order + m) * n + im];
      }
      outer[tid] = t;
    }
    __syncthreads();
    float* xp = x + plane * nn;
    double ss = 0.0;
    if ((n & 3) == 0) {
      const int n4 = n >> 2;
      for (int e = tid; e < n4 * n; e += 256) {
        const int i1 = e / n4, i0 = 4 * (e % n4);
        float4 v;
        float* vp = reinterpret_cast<float*>(&v);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double s = 0.0;
          for (int j = 0; j < n_bumps; ++j) s += outer[j] * (double)g1[j * n + i1] * (double)g0[j * n + i0 + q];
          vp[q] = (float)s;
          ss += (double)vp[q] * (double)vp[q];
        }
        reinterpret_cast<float4*>(xp)[e] = v;
      }
    } else {
      for (int e = tid; e < n * n; e += 256) {
        const int i1 = e / n, i0 = e % n;
        double s = 0.0;
        for (int j = 0; j < n_bumps; ++j) s += outer[j] * (double)g1[j * n + i1] * (double)g0[j * n + i0];
        const float v = (float)s;
        xp[e] = v;
        ss += (double)v * (double)v;
      }
    }
    if (part) {
      ss = wave_sum(ss);
      if ((tid & 63) == 0) red[tid >> 6] = ss;
      __syncthreads();
      if (tid == 0) part[plane] = red[0] + red[1] + red[2] + red[3];
    }
  }
}

// out[0] = sum of part[0 .. n) (one work-group).
__global__ __launch_bounds__(256) void sum_part_kernel(const double* __restrict__ part, long n, double* __restrict__ out) {
  __shared__ double red[4];
  double s = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) s += part[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}

int gen_noise_sumsq_launch(hipStream_t st, long n, uint64_t seed, uint32_t stream, double* acc, int* nb) {
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  *nb = (int)blocks;
  hipLaunchKernelGGL(noise_sumsq_kernel, dim3((unsigned)blocks), dim3(256), 0, st, n, seed, stream, acc);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
int gen_add_noise_launch(hipStream_t st, float* x, long n, uint64_t seed, uint32_t stream, double scale) {
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(add_noise_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x, n, seed, stream, scale);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}
long gen_bumps_planes(int order, long n) {
  long planes = 1;
  for (int m = 2; m < order; ++m) planes *= n;
  return planes;
}

int gen_bumps_launch(hipStream_t st, float* x, int order, long n, const float* factors, const float* amp,
                     int n_bumps, double* sumsq, double* part) {
  if (order < 1 || order > 8) return 8;
  const bool plane = order >= 2 && n <= BP_NMAX && n_bumps <= BP_JMAX && ((uintptr_t)x & 15) == 0;
  if (plane) {
    const long planes = gen_bumps_planes(order, n);
    const long grid = planes < 65536 ? planes : 65536;
    hipLaunchKernelGGL(bumps_plane_kernel, dim3((unsigned)grid), dim3(256), 0, st, x, order, (int)n, factors, amp,
                       n_bumps, planes, sumsq ? part : nullptr);
    if (sumsq) hipLaunchKernelGGL(sum_part_kernel, dim3(1), dim3(256), 0, st, part, planes, sumsq);
  } else {
    if (sumsq) return 8;  // the fused sum of squares needs the plane form
    long total = 1;
    for (int m = 0; m < order; ++m) total *= n;
    long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(bumps_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x, order, n, factors, amp, n_bumps);
  }
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// Plain streaming copy (float4 when both pointers are 16-byte aligned): the reference
// kernel that calibrates hardware byte counters against a known traffic of 2 n floats.
__global__ void copy_f32_kernel(const float* __restrict__ src, float* __restrict__ dst, long n, int vec) {
  const long stride = (long)gridDim.x * blockDim.x;
  const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (vec) {
    const long n4 = n / 4;
    for (long i = t0; i < n4; i += stride)
      reinterpret_cast<float4*>(dst)[i] = reinterpret_cast<const float4*>(src)[i];
    for (long i = 4 * n4 + t0; i < n; i += stride) dst[i] = src[i];
  } else {
    for (long i = t0; i < n; i += stride) dst[i] = src[i];
  }
}
int copy_f32_launch(hipStream_t st, const float* src, float* dst, long n) {
  if (n <= 0) return 0;
  const int vec = ((((uintptr_t)src) | ((uintptr_t)dst)) & 15) == 0;
  const long work = vec ? (n + 3) / 4 : n;
  long blocks = (work + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(copy_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, n, vec);
  return hipGetLastError() == hipSuccess ? 0 : 5;
}

// An empty launch (one work-group, no element: no memory is touched) and a wait for it:
// the runtime binds a stream to a hardware queue when the stream is first used, so new
// streams touched one after another get distinct queues where streams first used at the
// same moment from several threads can end up sharing one.
int touch_stream(hipStream_t st) {
  hipLaunchKernelGGL(copy_f32_kernel, dim3(1), dim3(64), 0, st, nullptr, nullptr, 0L, 0);
  if (hipGetLastError() != hipSuccess) return 5;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : 5;
}

}  // namespace ht
