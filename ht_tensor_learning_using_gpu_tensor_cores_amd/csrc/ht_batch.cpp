// Host side of the batched entry point: validation, per-tensor descriptors and output
// layout, work-sorted order, launch-shape autotuning and the hipGraph of the launch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "ht_engine.h"

using namespace ht;

namespace {

// Tree {1},{2,3} in the heap numbering: 0 = [0,3), 1 = [0,1), 2 = [1,3), 3 = [1,2), 4 = [2,3).
const int32_t kLo[5] = {0, 0, 1, 1, 2}, kHi[5] = {3, 1, 3, 2, 3};
const int32_t kLeft[5] = {1, -1, 3, -1, -1}, kRight[5] = {2, -1, 4, -1, -1};

int describe(int count, const int64_t* dims, const int32_t* ranks, std::vector<BatchItem>* items, int64_t* total) {
  int64_t off = 0;
  if (items) items->resize(count);
  for (int b = 0; b < count; ++b) {
    const int64_t* d = dims + 3 * b;
    const int32_t* k = ranks + 5 * b;
    for (int m = 0; m < 3; ++m)
      if (d[m] < 1 || d[m] > 64) return HT_ERR_UNSUPPORTED;
    for (int t = 1; t < 5; ++t)
      if (k[t] > 8) return HT_ERR_UNSUPPORTED;
    Problem p;
    int s = p.setup(3, d, 5, kLo, kHi, kLeft, kRight, k, 0);
    if (s) return s;
    if (items) {
      BatchItem& it = (*items)[b];
      std::memset(&it, 0, sizeof(it));
      it.out_off = off;
      for (int m = 0; m < 3; ++m) it.n[m] = (int)d[m];
      for (int t = 0; t < 5; ++t) {
        it.k[t] = k[t];
        it.u_off[t] = (int)p.u_off[t];
        it.b_off[t] = (int)p.b_off[t];
        it.s_off[t] = (int)p.s_off[t];
      }
    }
    off += p.total;
  }
  if (total) *total = off;
  return 0;
}

}  // namespace

ht_status ht_batch_output_layout(int32_t count, const int64_t* dims, const int32_t* ranks, int64_t* out_offsets,
                                 int64_t* total) {
  if (count < 1 || !dims || !ranks) return HT_ERR_INVALID_ARG;
  std::vector<BatchItem> items;
  int s = describe(count, dims, ranks, &items, total);
  if (s) return (ht_status)s;
  if (out_offsets)
    for (int b = 0; b < count; ++b) out_offsets[b] = items[b].out_off;
  return HT_OK;
}

ht_status ht_batch_shapes(int32_t count, uint64_t seed, int32_t lo, int32_t hi, int32_t rlo, int32_t rhi,
                          int64_t* dims, int32_t* core) {
  if (count < 1 || !dims || !core || lo < 1 || hi < lo || rlo < 1 || rhi < rlo) return HT_ERR_INVALID_ARG;
  // uniforms of stream 4095, six per tensor: three mode sizes then three Tucker ranks
  for (int b = 0; b < count; ++b)
    for (int m = 0; m < 3; ++m) {
      const double ud = philox_uniform(seed, 4095, (uint64_t)6 * b + m);
      const double ur = philox_uniform(seed, 4095, (uint64_t)6 * b + 3 + m);
      dims[3 * b + m] = std::min<int64_t>(lo + (int64_t)std::floor(ud * (hi - lo + 1)), hi);
      core[3 * b + m] = std::min<int32_t>(rlo + (int32_t)std::floor(ur * (rhi - rlo + 1)), rhi);
    }
  return HT_OK;
}

ht_status ht_generate_batch(ht_handle_t h, int32_t count, const int64_t* dims, const int32_t* core, uint64_t seed,
                            double noise_rel, float* x_dev, const int64_t* x_offsets) {
  if (!h || count < 1 || !dims || !core || !x_dev || !x_offsets || noise_rel < 0) return HT_ERR_INVALID_ARG;
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  for (int b = 0; b < count; ++b) {
    const int64_t* d = dims + 3 * b;
    const int32_t* c = core + 3 * b;
    const uint32_t base = 4096u + 16u * (uint32_t)b;
    long nc = 1;
    for (int m = 0; m < 3; ++m) {
      if (d[m] < 1 || c[m] < 1 || c[m] > d[m]) return HT_ERR_INVALID_ARG;
      nc *= c[m];
    }
    std::vector<double> cv(nc);
    for (long i = 0; i < nc; ++i) cv[i] = philox_normal(seed, base, (uint64_t)i);
    std::vector<std::vector<double>> U(3);
    for (int m = 0; m < 3; ++m) U[m] = orthonormal_factor(seed, base + 1 + m, d[m], c[m]);
    ht_status s = generate_from_core(h, 3, d, c, std::move(cv), std::move(U), seed, noise_rel, x_dev + x_offsets[b],
                                     base + 15);
    if (s) return s;
  }
  return HT_OK;
}

ht_status ht_decompose_batched(ht_handle_t h, const float* x, int32_t count, const int64_t* dims,
                               const int32_t* ranks, const int64_t* x_offsets, float* out,
                               const int64_t* out_offsets, double* rel_err, ht_stats* stats) {
  if (!h || !x || !out || count < 1 || !dims || !ranks || !x_offsets || !out_offsets) return HT_ERR_INVALID_ARG;
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  std::vector<BatchItem> items;
  int64_t total = 0;
  int s = describe(count, dims, ranks, &items, &total);
  if (s) return (ht_status)s;
  int nmax = 1;
  std::vector<double> work(count);
  for (int b = 0; b < count; ++b) {
    items[b].x_off = x_offsets[b];
    items[b].out_off = out_offsets[b];
    const double n1 = items[b].n[0], n2 = items[b].n[1], n3 = items[b].n[2];
    work[b] = n1 * n2 * n3 * (n1 + n2 + n3) + 30.0 * (n1 * n1 + n2 * n2 + n3 * n3) * 64.0;
    nmax = std::max(nmax, std::max(items[b].n[0], std::max(items[b].n[1], items[b].n[2])));
  }
  // largest work first, so that the strided persistent blocks finish together
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
  int dev_cus = 0;
  if (hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess) return HT_ERR_HIP;
  h->arena.reset();
  BatchItem* itd = (BatchItem*)h->arena.get(sizeof(BatchItem) * count);
  int* ord = (int*)h->arena.get(sizeof(int) * count);
  double* errd = (double*)h->arena.get(sizeof(double) * count);
  const int max_blocks = 2 * dev_cus;
  double* scratch = (double*)h->arena.get(batch_scratch_bytes(max_blocks));
  if (!itd || !ord || !errd || !scratch) return HT_ERR_ALLOC;
  if (hipMemcpyAsync(itd, items.data(), sizeof(BatchItem) * count, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemcpyAsync(ord, order.data(), sizeof(int) * count, hipMemcpyHostToDevice, h->stream) != hipSuccess)
    return HT_ERR_HIP;
  // Launch shape: chunk length of the mode-3 Gram (32 or 64 positions) and persistent
  // grid size (1 or 2 blocks per CU), timed once per (count, largest mode size).
  const std::pair<int, int> key(count, nmax);
  auto tuned = h->batch_tune.find(key);
  if (tuned == h->batch_tune.end()) {
    float best = 1e30f;
    std::pair<int, int> pick(64, max_blocks);
    for (int kc : {32, 64})
      for (int blocks : {dev_cus, max_blocks}) {
        const int nb = std::min(blocks, count);
        if (hipEventRecord(h->ev[0], h->stream) != hipSuccess) return HT_ERR_HIP;
        if (batch_ht_launch(h->stream, x, itd, ord, count, out, errd, scratch, nb, kc)) return HT_ERR_HIP;
        if (hipEventRecord(h->ev[1], h->stream) != hipSuccess || hipEventSynchronize(h->ev[1]) != hipSuccess)
          return HT_ERR_HIP;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
        if (ms < best) {
          best = ms;
          pick = {kc, nb};
        }
      }
    tuned = h->batch_tune.emplace(key, pick).first;
  }
  const int kc = tuned->second.first, blocks = std::min(tuned->second.second, count);
  // The launch runs from a graph captured for these exact arguments.
  std::vector<const void*> bkey = {x, itd, ord, out, errd, scratch, (const void*)(intptr_t)count,
                                   (const void*)(intptr_t)blocks, (const void*)(intptr_t)kc};
  if (!h->bexec || h->bkey != bkey) {
    if (h->bexec) (void)hipGraphExecDestroy(h->bexec);
    if (h->bgraph) (void)hipGraphDestroy(h->bgraph);
    h->bexec = nullptr;
    h->bgraph = nullptr;
    if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return HT_ERR_HIP;
    const int ls = batch_ht_launch(h->stream, x, itd, ord, count, out, errd, scratch, blocks, kc);
    const hipError_t ce = hipStreamEndCapture(h->stream, &h->bgraph);
    if (ls || ce != hipSuccess) return HT_ERR_HIP;
    if (hipGraphInstantiate(&h->bexec, h->bgraph, nullptr, nullptr, 0) != hipSuccess) return HT_ERR_HIP;
    h->bkey = bkey;
  }
  if (hipEventRecord(h->ev[0], h->stream) != hipSuccess) return HT_ERR_HIP;
  if (hipGraphLaunch(h->bexec, h->stream) != hipSuccess) return HT_ERR_HIP;
  if (hipEventRecord(h->ev[1], h->stream) != hipSuccess) return HT_ERR_HIP;
  if (rel_err &&
      hipMemcpyAsync(rel_err, errd, sizeof(double) * count, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    return HT_ERR_HIP;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return HT_ERR_HIP;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->total_ms = ms;
    stats->bytes_x_passes = 4.0;
    stats->eig_calls = 3 * count;
  }
  return HT_OK;
}
