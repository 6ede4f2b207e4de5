// C ABI of the HT engine (declarations in include/ht_amd.h and include/ht_amd_debug.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <memory>
#include <vector>

#include "../../include/ht_amd_debug.h"
#include "ht_engine.h"

using namespace ht;

#define CKS(expr)        \
  do {                   \
    int _s = (expr);     \
    if (_s) return (ht_status)_s; \
  } while (0)

// ------------------------------------------------------------------ generators
// Q of the QR factorisation of the n x r Philox Gaussian (seed, stream), diag(R) > 0,
// by Gram-Schmidt with reorthogonalisation in FP64 (the Q factor with positive diag(R)
// is unique, so it matches any stable QR up to rounding).
std::vector<double> ht::orthonormal_factor(uint64_t seed, uint32_t stream, long n, int r) {
  std::vector<double> A((size_t)n * r);
  for (long i = 0; i < n * r; ++i) A[i] = philox_normal(seed, stream, (uint64_t)i);
  for (int j = 0; j < r; ++j) {
    double* a = &A[(size_t)n * j];
    for (int pass = 0; pass < 2; ++pass)
      for (int i = 0; i < j; ++i) {
        const double* q = &A[(size_t)n * i];
        double d = 0.0;
        for (long t = 0; t < n; ++t) d += q[t] * a[t];
        for (long t = 0; t < n; ++t) a[t] -= d * q[t];
      }
    double nr = 0.0;
    for (long t = 0; t < n; ++t) nr += a[t] * a[t];
    nr = std::sqrt(nr);
    for (long t = 0; t < n; ++t) a[t] /= nr;
  }
  return A;
}

// x = core x_n U_n (+ noise) with the modes split into [0, m) and [m, d): X_(left) = L R^T,
// L = U_{m-1} (x) ... (x) U_0, R = (core x_{n>=m} U_n)_(left)^T.
ht_status ht::generate_from_core(ht_handle_t h, int order, const int64_t* dims, const int32_t* core_dims,
                             std::vector<double> core, std::vector<std::vector<double>> U, uint64_t seed,
                             double noise_rel, float* x_dev, uint32_t noise_stream) {
  const int m = (order + 1) / 2;
  long nl = 1, kl = 1, nr = 1;
  for (int k = 0; k < m; ++k) { nl *= dims[k]; kl *= core_dims[k]; }
  for (int k = m; k < order; ++k) nr *= dims[k];
  if (kl > 4096) return HT_ERR_UNSUPPORTED;
  // L (nl x kl)
  std::vector<float> L((size_t)nl * kl);
  for (long col = 0; col < kl; ++col)
    for (long row = 0; row < nl; ++row) {
      long rr = row, cc = col;
      double v = 1.0;
      for (int k = 0; k < m; ++k) {
        const long i = rr % dims[k], j = cc % core_dims[k];
        rr /= dims[k];
        cc /= core_dims[k];
        v *= U[k][i + dims[k] * j];
      }
      L[row + nl * col] = (float)v;
    }
  // T = core x_{k>=m} U_k, modes (core_dims[0..m), dims[m..d)).
  std::vector<long> td(order);
  for (int k = 0; k < order; ++k) td[k] = core_dims[k];
  std::vector<double> T = std::move(core);
  for (int k = m; k < order; ++k) {
    long P = 1, S = 1;
    for (int q = 0; q < k; ++q) P *= td[q];
    for (int q = k + 1; q < order; ++q) S *= td[q];
    const long rk = td[k], nk = dims[k];
    std::vector<double> T2((size_t)P * nk * S, 0.0);
    for (long s = 0; s < S; ++s)
      for (long j = 0; j < rk; ++j)
        for (long i = 0; i < nk; ++i) {
          const double u = U[k][i + nk * j];
          const double* src = &T[(size_t)P * (j + rk * s)];
          double* dst = &T2[(size_t)P * (i + nk * s)];
          for (long p = 0; p < P; ++p) dst[p] += u * src[p];
        }
    T.swap(T2);
    td[k] = nk;
  }
  // R[i_right, j_left] = T[j_left + kl * i_right]
  std::vector<float> R((size_t)nr * kl);
  for (long i = 0; i < nr; ++i)
    for (long j = 0; j < kl; ++j) R[i + nr * j] = (float)T[j + kl * i];
  float *Ld = nullptr, *Rd = nullptr;
  double* acc = nullptr;
  const size_t accn = 2048;
  if (hipMalloc((void**)&Ld, L.size() * 4) != hipSuccess) return HT_ERR_ALLOC;
  if (hipMalloc((void**)&Rd, R.size() * 4) != hipSuccess) { (void)hipFree(Ld); return HT_ERR_ALLOC; }
  if (hipMalloc((void**)&acc, (accn + 8192) * 8) != hipSuccess) { (void)hipFree(Ld); (void)hipFree(Rd); return HT_ERR_ALLOC; }
  ht_status st = HT_OK;
  do {
    if (hipMemcpy(Ld, L.data(), L.size() * 4, hipMemcpyHostToDevice) != hipSuccess) { st = HT_ERR_HIP; break; }
    if (hipMemcpy(Rd, R.data(), R.size() * 4, hipMemcpyHostToDevice) != hipSuccess) { st = HT_ERR_HIP; break; }
    if (lowrank_launch(h->stream, 0, Ld, Rd, nl, nr, (int)kl, nullptr, x_dev, nullptr, nullptr)) { st = HT_ERR_HIP; break; }
    if (noise_rel > 0.0) {
      if (sumsq_f32_launch(h->stream, x_dev, nl * nr, acc)) { st = HT_ERR_HIP; break; }
      int nb = 0;
      if (gen_noise_sumsq_launch(h->stream, nl * nr, seed, noise_stream, acc + 2048, &nb)) { st = HT_ERR_HIP; break; }
      if (gram_reduce_norms_launch(h->stream, acc + 2048, nb, acc + 2047)) { st = HT_ERR_HIP; break; }
      double hv[2];
      if (hipMemcpyAsync(&hv[0], acc, 8, hipMemcpyDeviceToHost, h->stream) != hipSuccess) { st = HT_ERR_HIP; break; }
      if (hipMemcpyAsync(&hv[1], acc + 2047, 8, hipMemcpyDeviceToHost, h->stream) != hipSuccess) { st = HT_ERR_HIP; break; }
      if (hipStreamSynchronize(h->stream) != hipSuccess) { st = HT_ERR_HIP; break; }
      const double scale = noise_rel * std::sqrt(hv[0]) / std::sqrt(hv[1]);
      if (gen_add_noise_launch(h->stream, x_dev, nl * nr, seed, noise_stream, scale)) { st = HT_ERR_HIP; break; }
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) st = HT_ERR_HIP;
  } while (0);
  (void)hipFree(Ld);
  (void)hipFree(Rd);
  (void)hipFree(acc);
  return st;
}


// ------------------------------------------------------------------ no-allocation scratch
// The device workspace of ht_decompose_noalloc: one array of HT_NOALLOC_SCRATCH_BYTES per
// GPU the library runs on, held until the process ends. Its host-side name is not a device
// address: each device's address is taken once with hipGetSymbolAddress while that device
// is current (ht_create) and kept in g_scratch[device]; host code only ever passes that.
// Every ht_decompose_noalloc call on one GPU shares the array (the engine writes every
// workspace word it later reads, so nothing depends on what an earlier call left there);
// calls that overlap on one GPU race.
#define HT_SCRATCH_DEVICES 64
static __device__ __align__(256) char scratch[HT_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch[HT_SCRATCH_DEVICES] = {};

// Address of the current device's copy of `scratch` (device `device` must be current).
static ht_status scratch_setup(int device) {
  if (device < 0 || device >= HT_SCRATCH_DEVICES) return HT_OK;  // no-allocation entry unavailable there
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return HT_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HT_ERR_HIP;
  g_scratch[device] = p;
  return HT_OK;
}

extern "C" {

const char* ht_version(void) { return "ht_amd 0.1 (gfx950)"; }

void ht_default_options(ht_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->method = HT_METHOD_LTR;
  o->precision = HT_PREC_BF16;
  o->oversample = 8;
  o->power_iters = 1;
  o->err_mode = HT_ERR_MODE_RESIDUAL;
  o->max_eig_iters = 200;
  o->eig_tol = 1e-10;
  o->seed = 1234;
  o->want_interior_u = 0;
  o->check_finite = 1;
  o->sketch = HT_SKETCH_GAUSSIAN;
  o->sketch_width = 20;
  o->sketch_weights = nullptr;
  o->sketch_tol = 1e-3;
}

ht_status ht_create(ht_handle_t* handle, int32_t device) {
  if (!handle) return HT_ERR_INVALID_ARG;
  *handle = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return HT_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return HT_ERR_INVALID_ARG;
  if (hipSetDevice(device) != hipSuccess) return HT_ERR_HIP;
  if (scratch_setup(device) != HT_OK) return HT_ERR_HIP;
  ht_handle_s* h = new ht_handle_s();
  h->device = device;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return HT_ERR_HIP;
  }
  for (auto& e : h->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete h;
      return HT_ERR_HIP;
    }
  if (hipHostMalloc((void**)&h->pinned, 64 << 10, hipHostMallocDefault) == hipSuccess) h->pinned_bytes = 64 << 10;
  else h->pinned = nullptr;
  *handle = h;
  return HT_OK;
}

ht_status ht_destroy(ht_handle_t h) {
  if (!h) return HT_ERR_INVALID_ARG;
  for (ht_handle_s* w : h->workers) (void)ht_destroy(w);
  h->workers.clear();
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  if (h->bexec) (void)hipGraphExecDestroy(h->bexec);
  if (h->bgraph) (void)hipGraphDestroy(h->bgraph);
  h->arena.release();
  h->prof.destroy();
  if (h->pinned) (void)hipHostFree(h->pinned);
  if (h->xbuf) (void)hipFree(h->xbuf);
  if (h->obuf) (void)hipFree(h->obuf);
  for (auto& e : h->ev) (void)hipEventDestroy(e);
  for (auto& ew : h->stagger_ev)
    for (auto& e : ew)
      if (e) (void)hipEventDestroy(e);
  if (h->side[0]) {
    for (auto& s : h->side) (void)hipStreamSynchronize(s);
    for (auto& e : h->side_ev) (void)hipEventDestroy(e);
    for (auto& s : h->side) (void)hipStreamDestroy(s);
  }
  (void)hipStreamDestroy(h->stream);
  delete h;
  return HT_OK;
}

void* ht_get_stream(ht_handle_t h) { return h ? (void*)h->stream : nullptr; }

ht_status ht_validate_tree(int32_t order, int32_t n_nodes, const int32_t* lo, const int32_t* hi,
                           const int32_t* left, const int32_t* right) {
  return (ht_status)validate_tree(order, n_nodes, lo, hi, left, right);
}

ht_status ht_output_layout(int32_t order, const int64_t* dims, int32_t n_nodes, const int32_t* lo,
                           const int32_t* hi, const int32_t* left, const int32_t* right, const int32_t* ranks,
                           int32_t want_interior_u, int64_t* u_off, int64_t* b_off, int64_t* s_off,
                           int64_t* total) {
  Problem p;
  CKS(p.setup(order, dims, n_nodes, lo, hi, left, right, ranks, want_interior_u));
  for (int t = 0; t < n_nodes; ++t) {
    if (u_off) u_off[t] = p.u_off[t];
    if (b_off) b_off[t] = p.b_off[t];
    if (s_off) s_off[t] = p.s_off[t];
  }
  if (total) *total = p.total;
  return HT_OK;
}

static int ensure(float** buf, size_t* have, size_t bytes) {
  if (*have >= bytes) return 0;
  if (*buf) (void)hipFree(*buf);
  *buf = nullptr;
  *have = 0;
  if (hipMalloc((void**)buf, bytes) != hipSuccess) return HT_ERR_ALLOC;
  *have = bytes;
  return 0;
}

// The decomposition behind ht_decompose and ht_decompose_noalloc: `arena` is the handle's
// growable workspace or a fixed block that never allocates.
static ht_status decompose_impl(ht_handle_t h, const float* x, int32_t x_on_device, int32_t order,
                                const int64_t* dims, int32_t n_nodes, const int32_t* lo, const int32_t* hi,
                                const int32_t* left, const int32_t* right, const int32_t* ranks,
                                const ht_options* opt_in, float* out, int32_t out_on_device, double* rel_err,
                                ht_stats* stats, Arena* arena, double rel_eps = 0.0, int32_t max_rank = 0,
                                int32_t* ranks_out = nullptr, hipEvent_t pass_ev = nullptr,
                                std::atomic<int>* pass_flag = nullptr, int pass_mark = 1,
                                bool shared_gpu = false) {
  if (!h || !x) return HT_ERR_INVALID_ARG;
  if (rel_err) *rel_err = NAN;
  ht_options opt;
  if (opt_in) opt = *opt_in;
  else ht_default_options(&opt);
  if (opt.precision == HT_PREC_FP8 && opt.method != HT_METHOD_RSVD) return HT_ERR_UNSUPPORTED;
  if (opt.method < 0 || opt.method > 2 || opt.precision < 0 || opt.precision > 3 || opt.oversample < 0 ||
      opt.power_iters < 0 || opt.max_eig_iters < 1 || !(opt.eig_tol > 0) || opt.sketch < 0 || opt.sketch > 1)
    return HT_ERR_INVALID_ARG;
  if (opt.sketch == HT_SKETCH_LEARNED &&
      (opt.method != HT_METHOD_RSVD || !opt.sketch_weights || opt.sketch_width < 1 || opt.sketch_width > 64 ||
       !(opt.sketch_tol >= 0)))
    return HT_ERR_INVALID_ARG;
  Problem p;
  CKS(p.setup(order, dims, n_nodes, lo, hi, left, right, ranks, opt.want_interior_u));
  const bool adaptive = rel_eps > 0.0;
  if (adaptive) {
    if (!std::isfinite(rel_eps) || max_rank < 0) return HT_ERR_INVALID_ARG;
    // the ranks are caps: max_rank on every node, then k_t <= k_l k_r, deepest first
    std::vector<int> inner;
    for (int t = 1; t < p.n_nodes; ++t) {
      if (max_rank > 0) p.rank[t] = std::min(p.rank[t], (int)max_rank);
      if (p.left[t] >= 0) inner.push_back(t);
    }
    std::sort(inner.begin(), inner.end(), [&](int a, int b) { return p.depth(a) > p.depth(b); });
    for (int t : inner) p.rank[t] = std::min(p.rank[t], p.rank[p.left[t]] * p.rank[p.right[t]]);
  }
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  const size_t xbytes = (size_t)p.N * sizeof(float);
  float h2d_ms = 0.f;
  const float* xd = x;
  if (!x_on_device) {
    CKS(ensure(&h->xbuf, &h->xbuf_bytes, xbytes));
    if (hipEventRecord(h->ev[2], h->stream) != hipSuccess) return HT_ERR_HIP;
    if (hipMemcpyAsync(h->xbuf, x, xbytes, hipMemcpyHostToDevice, h->stream) != hipSuccess) return HT_ERR_HIP;
    if (hipEventRecord(h->ev[3], h->stream) != hipSuccess) return HT_ERR_HIP;
    xd = h->xbuf;
  }
  float* od = out;
  if (out && !out_on_device) {
    CKS(ensure(&h->obuf, &h->obuf_bytes, (size_t)p.total * sizeof(float)));
    od = h->obuf;
  }
  arena->reset();
  Engine e;
  e.st = h->stream;
  e.arena = arena;
  e.prob = &p;
  e.opt = opt;
  e.prof = &h->prof;
  e.pin = h->pinned;
  e.pin_bytes = h->pinned_bytes;
  e.side = h->side;
  e.side_ev = h->side_ev;
  e.plan = &h->plan;
  e.done_ev = h->ev[1];  // recorded by the run before its final wait
  e.adaptive = adaptive;
  e.adapt_eps = rel_eps;
  e.pass_ev = pass_ev;
  e.pass_flag = pass_flag;
  e.pass_mark = pass_mark;
  e.shared_gpu = shared_gpu;
  if (hipEventRecord(h->ev[0], h->stream) != hipSuccess) return HT_ERR_HIP;
  int s = e.run(xd, od);
  if (s && s != HT_ERR_NOCONV) {
    (void)hipStreamSynchronize(h->stream);
    if (arena->fixed && arena->overflow) return HT_ERR_WORKSPACE;
    return (ht_status)s;
  }
  if (!e.done_recorded && hipEventRecord(h->ev[1], h->stream) != hipSuccess) return HT_ERR_HIP;
  if (hipEventSynchronize(h->ev[1]) != hipSuccess) return HT_ERR_HIP;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
  if (!x_on_device) (void)hipEventElapsedTime(&h2d_ms, h->ev[2], h->ev[3]);
  if (out && !out_on_device) {
    if (hipMemcpy(out, od, (size_t)p.total * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return HT_ERR_HIP;
  }
  if (rel_err) *rel_err = e.rel_err;
  if (ranks_out)
    for (int t = 0; t < p.n_nodes; ++t) ranks_out[t] = p.rank[t];
  if (stats) {
    stats->total_ms = ms;
    stats->h2d_ms = h2d_ms;
    stats->error_ms = 0.0;
    stats->bytes_x_passes = e.passes;
    stats->eig_iters_max = e.eig_iters_max;
    stats->eig_calls = e.eig_calls;
    stats->sketch_fallbacks = e.sketch_fallbacks;
    stats->sketch_nodes = e.sketch_nodes;
    stats->sketch_resid = e.sketch_resid;
    stats->workspace_bytes = (double)arena->peak;
    stats->replayed = e.replayed;
  }
  if (!std::isfinite(e.x_sq)) return HT_ERR_NONFINITE;
  return (ht_status)s;
}

ht_status ht_decompose(ht_handle_t h, const float* x, int32_t x_on_device, int32_t order, const int64_t* dims,
                       int32_t n_nodes, const int32_t* lo, const int32_t* hi, const int32_t* left,
                       const int32_t* right, const int32_t* ranks, const ht_options* opt, float* out,
                       int32_t out_on_device, double* rel_err, ht_stats* stats) {
  if (!h) return HT_ERR_INVALID_ARG;
  return decompose_impl(h, x, x_on_device, order, dims, n_nodes, lo, hi, left, right, ranks, opt, out,
                        out_on_device, rel_err, stats, &h->arena);
}

ht_status ht_decompose_adaptive(ht_handle_t h, const float* x, int32_t x_on_device, int32_t order, const int64_t* dims,
                                int32_t n_nodes, const int32_t* lo, const int32_t* hi, const int32_t* left,
                                const int32_t* right, const int32_t* max_ranks, double rel_eps, int32_t max_rank,
                                const ht_options* opt, float* out, int32_t out_on_device, int32_t* ranks_out,
                                double* rel_err, ht_stats* stats) {
  if (!h || !ranks_out || !(rel_eps > 0.0)) return HT_ERR_INVALID_ARG;
  return decompose_impl(h, x, x_on_device, order, dims, n_nodes, lo, hi, left, right, max_ranks, opt, out,
                        out_on_device, rel_err, stats, &h->arena, rel_eps, max_rank, ranks_out);
}

ht_status ht_decompose_many(ht_handle_t h, int32_t count, const float* const* xs, int32_t order, const int64_t* dims,
                            int32_t n_nodes, const int32_t* lo, const int32_t* hi, const int32_t* left,
                            const int32_t* right, const int32_t* ranks, const ht_options* opt, float* const* outs,
                            double* rel_errs, int32_t inflight, ht_stats* stats) {
  if (!h || count < 0 || inflight < 1 || inflight > HT_MAX_INFLIGHT) return HT_ERR_INVALID_ARG;
  if (count > 0 && (!xs || !outs)) return HT_ERR_INVALID_ARG;
  for (int i = 0; i < count; ++i)
    if (!xs[i] || !outs[i]) return HT_ERR_INVALID_ARG;
  {
    ht_options o;
    if (opt) o = *opt;
    else ht_default_options(&o);
    Problem p;  // shape and tree errors before any work
    CKS(p.setup(order, dims, n_nodes, lo, hi, left, right, ranks, o.want_interior_u));
  }
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  // decomposition stream 0 is the handle's own; streams 1.. are worker handles
  while ((int)h->workers.size() < inflight - 1) {
    ht_handle_t w = nullptr;
    CKS(ht_create(&w, h->device));
    w->prof.enabled = h->prof.enabled;
    w->prof.mask = h->prof.mask;
    h->workers.push_back(w);
    if (w->prof.enabled && w->prof.create()) return HT_ERR_HIP;
    // first use of the worker's stream here, one worker after another (the stream then
    // holds its own hardware queue while queues are free)
    if (touch_stream(w->stream)) return HT_ERR_HIP;
  }
  // the workers start after what the caller's stream holds so far
  if (hipEventRecord(h->ev[0], h->stream) != hipSuccess) return HT_ERR_HIP;
  for (int w = 0; w + 1 < inflight; ++w)
    if (hipStreamWaitEvent(h->workers[w]->stream, h->ev[0], 0) != hipSuccess) return HT_ERR_HIP;
  for (int w = 0; w + 1 < inflight; ++w)
    for (auto& ev : h->stagger_ev[w])
      if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return HT_ERR_HIP;
  // Staggered workers: worker w's decomposition of round r (its r-th) waits for the first
  // pass over X of worker w - 1's decomposition of round r, so that one worker's passes
  // run beside another's solver phases. Workers left to themselves fall into step (every
  // pass at once, then every solver with HBM idle). passed[w] = rounds of worker w whose
  // first pass is enqueued (or that ended).
  std::unique_ptr<std::atomic<int>[]> passed(new std::atomic<int>[inflight]);
  for (int w = 0; w < inflight; ++w) passed[w].store(0);
  std::atomic<int> first_err{0};
  std::vector<ht_stats> wst(inflight);
  const auto t0 = std::chrono::steady_clock::now();
  auto work = [&](int w) {
    ht_handle_s* wh = w == 0 ? h : h->workers[w - 1];
    if (hipSetDevice(h->device) != hipSuccess) {
      int z = 0;
      first_err.compare_exchange_strong(z, (int)HT_ERR_HIP);
      return;
    }
    std::memset(&wst[w], 0, sizeof(ht_stats));
    for (int i = w; i < count && first_err.load() == 0; i += inflight) {
      double e = NAN;
      ht_stats st;
      std::memset(&st, 0, sizeof(st));
      const int r = i / inflight;
      if (w > 0) {
        while (passed[w - 1].load() <= r && first_err.load() == 0) std::this_thread::yield();
        (void)hipStreamWaitEvent(wh->stream, h->stagger_ev[w - 1][r % 4], 0);
      }
      hipEvent_t pev = w + 1 < inflight ? h->stagger_ev[w][r % 4] : nullptr;
      const ht_status s = decompose_impl(wh, xs[i], 1, order, dims, n_nodes, lo, hi, left, right, ranks, opt, outs[i],
                                         1, &e, &st, &wh->arena, 0.0, 0, nullptr, pev, pev ? &passed[w] : nullptr,
                                         r + 1, inflight > 1);
      if (passed[w].load() < r + 1) passed[w].store(r + 1);
      if (rel_errs) rel_errs[i] = e;
      if (s != HT_OK && s != HT_ERR_NOCONV) {
        int z = 0;
        first_err.compare_exchange_strong(z, (int)s);
      }
      wst[w].bytes_x_passes = st.bytes_x_passes;
      wst[w].eig_iters_max = std::max(wst[w].eig_iters_max, st.eig_iters_max);
      wst[w].eig_calls += st.eig_calls;
      wst[w].workspace_bytes = std::max(wst[w].workspace_bytes, st.workspace_bytes);
      wst[w].replayed += st.replayed;
    }
    passed[w].store(1 << 30);  // a worker that is done never holds the next one back
  };
  std::vector<std::thread> threads;
  for (int w = 1; w < inflight; ++w) threads.emplace_back(work, w);
  work(0);
  for (auto& t : threads) t.join();
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->total_ms = ms;
    for (const ht_stats& w : wst) {
      stats->bytes_x_passes = std::max(stats->bytes_x_passes, w.bytes_x_passes);
      stats->eig_iters_max = std::max(stats->eig_iters_max, w.eig_iters_max);
      stats->eig_calls += w.eig_calls;
      stats->workspace_bytes += w.workspace_bytes;
      stats->replayed += w.replayed;
    }
  }
  return (ht_status)first_err.load();
}

ht_status ht_decompose_noalloc(ht_handle_t h, const float* x, int32_t order, const int64_t* dims, int32_t n_nodes,
                               const int32_t* lo, const int32_t* hi, const int32_t* left, const int32_t* right,
                               const int32_t* ranks, const ht_options* opt, float* out, double* rel_err,
                               ht_stats* stats) {
  if (!h || !x || !out) return HT_ERR_INVALID_ARG;
  if (h->device < 0 || h->device >= HT_SCRATCH_DEVICES || !g_scratch[h->device]) return HT_ERR_HIP;
  Arena fa;
  fa.fixed = true;
  fa.fbase = static_cast<char*>(g_scratch[h->device]);
  fa.fsize = HT_NOALLOC_SCRATCH_BYTES;
  return decompose_impl(h, x, 1, order, dims, n_nodes, lo, hi, left, right, ranks, opt, out, 1, rel_err, stats,
                        &fa);
}

// Load an HT block from the packed layout into an engine (FP64 copies on the device).
static int load_packed(Engine& e, const Problem& p, const float* packed) {
  for (int t = 0; t < HT_MAX_NODES; ++t) {
    e.Ufull[t] = nullptr;
    e.Bd[t] = nullptr;
    e.SVd[t] = nullptr;
  }
  for (int t = 0; t < p.n_nodes; ++t) {
    if (p.left[t] < 0) {
      const long n = p.rows[t] * p.rank[t];
      e.Ufull[t] = (double*)e.arena->get(n * sizeof(double));
      if (!e.Ufull[t]) return HT_ERR_ALLOC;
      CKS(cast_f32_f64_launch(e.st, packed + p.u_off[t], e.Ufull[t], n));
    } else {
      const long n = (long)p.rank[p.left[t]] * p.rank[p.right[t]] * p.rank[t];
      e.Bd[t] = (double*)e.arena->get(n * sizeof(double));
      if (!e.Bd[t]) return HT_ERR_ALLOC;
      CKS(cast_f32_f64_launch(e.st, packed + p.b_off[t], e.Bd[t], n));
    }
  }
  return 0;
}

ht_status ht_residual(ht_handle_t h, const float* x, int32_t order, const int64_t* dims, int32_t n_nodes,
                      const int32_t* lo, const int32_t* hi, const int32_t* left, const int32_t* right,
                      const int32_t* ranks, int32_t want_interior_u, const float* packed, double* rel_err) {
  if (!h || !x || !packed || !rel_err) return HT_ERR_INVALID_ARG;
  Problem p;
  CKS(p.setup(order, dims, n_nodes, lo, hi, left, right, ranks, want_interior_u));
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  h->arena.reset();
  Engine e;
  e.st = h->stream;
  e.arena = &h->arena;
  e.prob = &p;
  ht_default_options(&e.opt);
  CKS(load_packed(e, p, packed));
  double esq = 0, xsq = 0;
  CKS(e.residual(x, &esq, &xsq));
  *rel_err = xsq > 0 ? std::sqrt(esq / xsq) : 0.0;
  return HT_OK;
}

ht_status ht_reconstruct(ht_handle_t h, int32_t order, const int64_t* dims, int32_t n_nodes, const int32_t* lo,
                         const int32_t* hi, const int32_t* left, const int32_t* right, const int32_t* ranks,
                         int32_t want_interior_u, const float* packed, float* x_out) {
  if (!h || !packed || !x_out) return HT_ERR_INVALID_ARG;
  Problem p;
  CKS(p.setup(order, dims, n_nodes, lo, hi, left, right, ranks, want_interior_u));
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  h->arena.reset();
  Engine e;
  e.st = h->stream;
  e.arena = &h->arena;
  e.prob = &p;
  ht_default_options(&e.opt);
  CKS(load_packed(e, p, packed));
  const int l = p.left[0], r = p.right[0];
  const long nl = p.rows[l], nr = p.rows[r];
  const int kl = p.rank[l], kr = p.rank[r];
  double *Ul, *Ur;
  CKS(e.expand(l, &Ul));
  CKS(e.expand(r, &Ur));
  double* Rr = (double*)h->arena.get(sizeof(double) * nr * kl);
  float* Lf = (float*)h->arena.get(sizeof(float) * nl * kl);
  float* Rf = (float*)h->arena.get(sizeof(float) * nr * kl);
  if (!Rr || !Lf || !Rf) return HT_ERR_ALLOC;
  CKS(dgemm_small_launch(e.st, 0, 1, (int)nr, kl, kr, 1.0, Ur, (int)nr, e.Bd[0], kl, 0.0, Rr, (int)nr));
  CKS(cast_f64_f32_launch(e.st, Ul, Lf, nl * kl));
  CKS(cast_f64_f32_launch(e.st, Rr, Rf, nr * kl));
  CKS(lowrank_launch(e.st, 0, Lf, Rf, nl, nr, kl, nullptr, x_out, nullptr, nullptr));
  if (hipStreamSynchronize(e.st) != hipSuccess) return HT_ERR_HIP;
  return HT_OK;
}

ht_status ht_generate_tucker(ht_handle_t h, int32_t order, const int64_t* dims, const int32_t* core_dims,
                             uint64_t seed, double noise_rel, float* x_dev) {
  if (!h || !dims || !core_dims || !x_dev || order < 2 || order > HT_MAX_ORDER || noise_rel < 0) return HT_ERR_INVALID_ARG;
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  long nc = 1;
  for (int k = 0; k < order; ++k) {
    if (dims[k] < 1 || core_dims[k] < 1 || core_dims[k] > dims[k]) return HT_ERR_INVALID_ARG;
    nc *= core_dims[k];
  }
  std::vector<double> core(nc);
  for (long i = 0; i < nc; ++i) core[i] = philox_normal(seed, 1, (uint64_t)i);
  std::vector<std::vector<double>> U(order);
  for (int k = 0; k < order; ++k) U[k] = orthonormal_factor(seed, 16 + k, dims[k], core_dims[k]);
  return generate_from_core(h, order, dims, core_dims, std::move(core), std::move(U), seed, noise_rel, x_dev, 64);
}

ht_status ht_generate_cp_decay(ht_handle_t h, int32_t order, const int64_t* dims, int32_t rank, double exponent,
                               uint64_t seed, double noise_rel, float* x_dev) {
  if (!h || !dims || !x_dev || order < 2 || order > HT_MAX_ORDER || rank < 1 || noise_rel < 0) return HT_ERR_INVALID_ARG;
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  std::vector<int32_t> cd(order, rank);
  long nc = 1, stride = 0, pk = 1;
  for (int k = 0; k < order; ++k) {
    if (dims[k] < rank) return HT_ERR_INVALID_ARG;
    nc *= rank;
    stride += pk;
    pk *= rank;
  }
  std::vector<double> core(nc, 0.0);
  for (int k = 0; k < rank; ++k) core[(size_t)k * stride] = std::pow(k + 1.0, -exponent);
  std::vector<std::vector<double>> U(order);
  for (int k = 0; k < order; ++k) U[k] = orthonormal_factor(seed, 16 + k, dims[k], rank);
  return generate_from_core(h, order, dims, cd.data(), std::move(core), std::move(U), seed, noise_rel, x_dev, 64);
}

ht_status ht_generate_bumps(ht_handle_t h, int32_t order, int64_t n, int32_t n_bumps, uint64_t seed, int64_t sample,
                            float* x_dev) {
  if (!h || !x_dev || order < 1 || order > HT_MAX_ORDER || n < 2 || n_bumps < 1 || sample < 0) return HT_ERR_INVALID_ARG;
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  std::vector<float> amp(n_bumps), fac((size_t)n_bumps * order * n);
  bump_factors_host(order, n, n_bumps, seed, sample, amp.data(), fac.data());
  float *fd = nullptr, *ad = nullptr;
  if (hipMalloc((void**)&fd, fac.size() * 4) != hipSuccess) return HT_ERR_ALLOC;
  if (hipMalloc((void**)&ad, amp.size() * 4) != hipSuccess) { (void)hipFree(fd); return HT_ERR_ALLOC; }
  ht_status st = HT_OK;
  if (hipMemcpy(fd, fac.data(), fac.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(ad, amp.data(), amp.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      gen_bumps_launch(h->stream, x_dev, order, n, fd, ad, n_bumps, nullptr, nullptr) ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    st = HT_ERR_HIP;
  (void)hipFree(fd);
  (void)hipFree(ad);
  return st;
}

// ------------------------------------------------------------------ compat wrappers
static ht_handle_t default_handle() {
  static std::once_flag once;
  static ht_handle_t h = nullptr;
  std::call_once(once, [] {
    if (ht_create(&h, 0) != HT_OK) h = nullptr;
  });
  return h;
}

static double compat_run(float* x, int order, const int64_t* dims, const int32_t* lo, const int32_t* hi,
                         const int32_t* left, const int32_t* right, const int* k, int method, double* ms) {
  ht_handle_t h = default_handle();
  if (!h || !x || !k) return -1.0;
  const int nn = 2 * order - 1;
  std::vector<int32_t> ranks(k, k + nn);
  ht_options opt;
  ht_default_options(&opt);
  opt.method = method;
  opt.precision = method == HT_METHOD_PLAIN ? HT_PREC_FP32 : HT_PREC_BF16;
  double err = NAN;
  ht_stats stats;
  const ht_status s = ht_decompose(h, x, 0, order, dims, nn, lo, hi, left, right, ranks.data(), &opt, nullptr, 0,
                                   &err, &stats);
  if (s != HT_OK) return -1.0 - (double)s;
  *ms = stats.total_ms + stats.h2d_ms;
  return err;
}

// Heap-numbered reference trees.
static const int32_t T3_LO[5] = {0, 0, 2, 0, 1}, T3_HI[5] = {3, 2, 3, 1, 2}, T3_L[5] = {1, 3, -1, -1, -1},
                     T3_R[5] = {2, 4, -1, -1, -1};
static const int32_t T4_LO[7] = {0, 0, 2, 0, 1, 2, 3}, T4_HI[7] = {4, 2, 4, 1, 2, 3, 4},
                     T4_L[7] = {1, 3, 5, -1, -1, -1, -1}, T4_R[7] = {2, 4, 6, -1, -1, -1, -1};

// rel_eps > 0: the ranks k[1..4] become caps (with max_rank > 0 as a further cap) and each
// node's rank is chosen for ||X - X_HT|| <= rel_eps ||X|| (ht_decompose_adaptive); the ranks
// chosen are written back into k. rel_eps <= 0: the fixed ranks k, as in the reference.
float htd(float* x, long a, long b, long c, int* k, float rel_eps, float max_rank) {
  const int64_t dims[3] = {a, b, c};
  double ms = 0;
  double err = -1.0;
  if (rel_eps > 0.0f) {
    ht_handle_t h = default_handle();
    if (!h || !x || !k) return -1.0f;
    int32_t caps[5], chosen[5];
    for (int t = 0; t < 5; ++t) caps[t] = k[t];
    ht_options opt;
    ht_default_options(&opt);
    opt.method = HT_METHOD_PLAIN;
    opt.precision = HT_PREC_FP32;
    ht_stats stats;
    const ht_status s = ht_decompose_adaptive(h, x, 0, 3, dims, 5, T3_LO, T3_HI, T3_L, T3_R, caps, rel_eps,
                                              max_rank > 0.0f ? (int32_t)max_rank : 0, &opt, nullptr, 0, chosen,
                                              &err, &stats);
    if (s != HT_OK) return -1.0f - (float)s;
    for (int t = 0; t < 5; ++t) k[t] = chosen[t];
    ms = stats.total_ms + stats.h2d_ms;
  } else {
    err = compat_run(x, 3, dims, T3_LO, T3_HI, T3_L, T3_R, k, HT_METHOD_PLAIN, &ms);
  }
  if (err >= 0) {
    printf("cost time :%gs\n", ms / 1000.0);
    printf("error rate %g\n", err);
  }
  return (float)err;
}

float htd_tc(float* x, long a, long b, long c, int* k) {
  const int64_t dims[3] = {a, b, c};
  double ms = 0;
  const double err = compat_run(x, 3, dims, T3_LO, T3_HI, T3_L, T3_R, k, HT_METHOD_LTR, &ms);
  if (err >= 0) {
    printf("cost time :%gs\n", ms / 1000.0);
    printf("error rate %g\n", err);
  }
  return (float)err;
}

void htd4(float* x, long a, long b, long c, long d, int* k) {
  const int64_t dims[4] = {a, b, c, d};
  double ms = 0;
  const double err = compat_run(x, 4, dims, T4_LO, T4_HI, T4_L, T4_R, k, HT_METHOD_PLAIN, &ms);
  if (err >= 0) printf("error rate %g\n", err);
  else printf("htd4 failed (status %d)\n", (int)(-1.0 - err));
}

void htd4_tc(float* x, long a, long b, long c, long d, int* k, void* blas_handle, void* solver_handle) {
  (void)blas_handle;
  (void)solver_handle;
  const int64_t dims[4] = {a, b, c, d};
  double ms = 0;
  const double err = compat_run(x, 4, dims, T4_LO, T4_HI, T4_L, T4_R, k, HT_METHOD_LTR, &ms);
  if (err >= 0) printf("error rate %g\n", err);
  else printf("htd4_tc failed (status %d)\n", (int)(-1.0 - err));
}

// ------------------------------------------------------------------ profiling
// The profiling calls act on the handle and on the worker handles of ht_decompose_many;
// reads sum over them.
ht_status ht_profile_enable(ht_handle_t h, int32_t on) {
  if (!h) return HT_ERR_INVALID_ARG;
  h->prof.enabled = on != 0;
  for (ht_handle_s* w : h->workers) w->prof.enabled = on != 0;
  if (on) {
    // the events are made here, not inside the work that is timed
    if (hipSetDevice(h->device) != hipSuccess || h->prof.create()) return HT_ERR_HIP;
    for (ht_handle_s* w : h->workers)
      if (w->prof.create()) return HT_ERR_HIP;
  }
  return HT_OK;
}

ht_status ht_profile_select(ht_handle_t h, int32_t mask) {
  if (!h) return HT_ERR_INVALID_ARG;
  h->prof.mask = mask;
  for (ht_handle_s* w : h->workers) w->prof.mask = mask;
  return HT_OK;
}

ht_status ht_profile_read(ht_handle_t h, double* ms, int64_t* launches, int32_t n, int32_t reset) {
  if (!h || n < 0) return HT_ERR_INVALID_ARG;
  for (int i = 0; i < n && i < PK_COUNT; ++i) {
    double m = h->prof.ms[i];
    int64_t c = h->prof.cnt[i];
    for (ht_handle_s* w : h->workers) {
      m += w->prof.ms[i];
      c += w->prof.cnt[i];
    }
    if (ms) ms[i] = m;
    if (launches) launches[i] = c;
  }
  if (reset) {
    for (int i = 0; i < PK_COUNT; ++i) {
      h->prof.ms[i] = 0.0;
      h->prof.cnt[i] = 0;
      for (ht_handle_s* w : h->workers) {
        w->prof.ms[i] = 0.0;
        w->prof.cnt[i] = 0;
      }
    }
  }
  return HT_OK;
}

// ------------------------------------------------------------------ component entry points
ht_status ht_debug_qr_fallbacks(ht_handle_t h, int64_t* count) {
  if (!h || !count) return HT_ERR_INVALID_ARG;
  if (hipSetDevice(h->device) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) return HT_ERR_HIP;
  unsigned long long v = 0;
  if (qr_fallback_count(&v)) return HT_ERR_HIP;
  *count = (int64_t)v;
  return HT_OK;
}

ht_status ht_debug_copy(ht_handle_t h, const float* src, float* dst, int64_t n) {
  if (!h || n < 0 || (n > 0 && (!src || !dst))) return HT_ERR_INVALID_ARG;
  CKS(copy_f32_launch(h->stream, src, dst, n));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_debug_mfma_probe(ht_handle_t h, const float* a, const float* b, float* o_bf16_32, float* o_f32_32,
                              float* o_f32_16, double* o_f64_16, float* o_bf16_16) {
  if (!h) return HT_ERR_INVALID_ARG;
  CKS(mfma_probe_launch(h->stream, a, b, o_bf16_32, o_f32_32, o_f32_16, o_f64_16, o_bf16_16));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_mode_gram(ht_handle_t h, const float* x, int32_t order, const int64_t* dims, int32_t lo, int32_t hi,
                       int32_t precision, double* G, double* sumsq) {
  if (!h || !x || !dims || !G || order < 1 || order > HT_MAX_ORDER || lo < 0 || hi > order || lo >= hi)
    return HT_ERR_INVALID_ARG;
  PMS v{1, 1, 1};
  for (int m = 0; m < lo; ++m) v.P *= dims[m];
  for (int m = lo; m < hi; ++m) v.M *= dims[m];
  for (int m = hi; m < order; ++m) v.S *= dims[m];
  if (v.M > 1024) return HT_ERR_UNSUPPORTED;
  const int prec = precision == HT_PREC_FP32 ? PREC_F32 : (precision == HT_PREC_FP16 ? PREC_F16 : PREC_BF16);
  h->arena.reset();
  Engine e;
  e.st = h->stream;
  e.arena = &h->arena;
  double* s = sumsq ? (double*)h->arena.get(8) : nullptr;
  CKS(e.gram(x, v, prec, G, s));
  if (sumsq) CKS(e.d2h(sumsq, s, 8));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_mode_gram16(ht_handle_t h, const float* x, int32_t order, const int64_t* dims, int32_t lo,
                         int32_t hi, int32_t precision, double* G, double* sumsq) {
  if (!h || !x || !dims || !G || order < 1 || order > HT_MAX_ORDER || lo < 0 || hi > order || lo >= hi)
    return HT_ERR_INVALID_ARG;
  PMS v{1, 1, 1};
  for (int m = 0; m < lo; ++m) v.P *= dims[m];
  for (int m = lo; m < hi; ++m) v.M *= dims[m];
  for (int m = hi; m < order; ++m) v.S *= dims[m];
  if (precision != HT_PREC_BF16 && precision != HT_PREC_FP16) return HT_ERR_UNSUPPORTED;
  if (!gram16_ok(v) || ((uintptr_t)x & 15)) return HT_ERR_UNSUPPORTED;
  const int prec = precision == HT_PREC_FP16 ? PREC_F16 : PREC_BF16;
  const long n = v.P * v.M * v.S;
  h->arena.reset();
  Engine e;
  e.st = h->stream;
  e.arena = &h->arena;
  uint16_t* y = (uint16_t*)h->arena.get((size_t)n * sizeof(uint16_t));
  double* part = (double*)h->arena.get((size_t)cast16_part_size() * sizeof(double));
  double* s = (double*)h->arena.get(8);
  if (!y || !part || !s) return HT_ERR_ALLOC;
  CKS(cast16_sumsq_launch(h->stream, x, n, prec, y, part, s));
  e.x16 = y;
  e.x16_prec = prec;
  CKS(e.gram(x, v, prec, G, nullptr));
  if (sumsq) CKS(e.d2h(sumsq, s, 8));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_slab_gram(ht_handle_t h, const float* x, int32_t n1, int32_t n2, int64_t S, int32_t precision,
                       double* G1, double* G2, double* sumsq) {
  if (!h || !x || !G1 || !G2 || n1 < 1 || n2 < 1 || n1 > 128 || n2 > 128 || S < 1) return HT_ERR_INVALID_ARG;
  const int prec = precision == HT_PREC_FP16 ? PREC_F16 : PREC_BF16;
  h->arena.reset();
  const int nb = (int)std::min<int64_t>(S, 256);
  float* p1 = (float*)h->arena.get((size_t)nb * 128 * 128 * 4);
  float* p2 = (float*)h->arena.get((size_t)nb * 128 * 128 * 4);
  double* norms = (double*)h->arena.get((size_t)nb * 8 + 64);
  if (!p1 || !p2 || !norms) return HT_ERR_ALLOC;
  CKS(slab_gram_launch(h->stream, x, n1, n2, S, prec, nb, p1, p2, norms));
  CKS(gram_reduce_launch(h->stream, p1, n1, nb, G1));
  CKS(gram_reduce_launch(h->stream, p2, n2, nb, G2));
  if (sumsq) {
    CKS(gram_reduce_norms_launch(h->stream, norms, nb, norms + nb));
    if (hipMemcpyAsync(sumsq, norms + nb, 8, hipMemcpyDeviceToHost, h->stream) != hipSuccess) return HT_ERR_HIP;
  }
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_slab_project(ht_handle_t h, const float* x, int32_t n1, int32_t n2, int64_t S, const float* U1, int32_t r1,
                          const float* U2, int32_t r2, float* z) {
  if (!h || !x || !U1 || !U2 || !z) return HT_ERR_INVALID_ARG;
  CKS(slab_proj_launch(h->stream, x, n1, n2, S, U1, r1, U2, r2, z));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_sym_eig(ht_handle_t h, const double* A, int32_t n, double* w, double* V) {
  if (!h || !A || !w || !V || n < 1 || n > 64) return HT_ERR_INVALID_ARG;
  h->arena.reset();
  int* info = (int*)h->arena.get(64);
  CKS(jacobi_eig_launch(h->stream, A, n, n, w, V, n, info, 60, 1e-15));
  int hinfo = 0;
  if (hipMemcpyAsync(&hinfo, info, 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess) return HT_ERR_HIP;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return HT_ERR_HIP;
  return hinfo ? HT_ERR_NOCONV : HT_OK;
}

ht_status ht_sym_topk(ht_handle_t h, const double* G, int32_t n, int32_t k, double tol, double* U, double* w,
                      int32_t* iters) {
  if (!h || !G || !U || !w || n < 1 || k < 1 || k > n || k > 64) return HT_ERR_INVALID_ARG;
  h->arena.reset();
  Engine e;
  e.st = h->stream;
  e.arena = &h->arena;
  ht_default_options(&e.opt);
  e.eig_iters_max = 0;
  e.noconv = 0;
  CKS(e.sym_topk(G, n, k, U, w, tol, 500, 77, 0.0));
  if (iters) *iters = e.eig_iters_max;
  if (hipStreamSynchronize(h->stream) != hipSuccess) return HT_ERR_HIP;
  return e.noconv ? HT_ERR_NOCONV : HT_OK;
}

ht_status ht_tsqr(ht_handle_t h, float* W, int64_t m, int32_t b, float* Q, double* R) {
  if (!h || !W || !Q || !R || m < 1 || b < 1 || b > 64) return HT_ERR_INVALID_ARG;
  h->arena.reset();
  void* ws = h->arena.get(tsqr_ws_bytes(m, b));
  if (!ws) return HT_ERR_ALLOC;
  CKS(tsqr_launch(h->stream, W, m, b, Q, R, ws));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_tall_orth(ht_handle_t h, float* Y, int64_t m, int32_t b, float* Q, int32_t* fallback) {
  if (!h || !Y || !Q || m < 1 || b < 1 || b > 64 || m < b) return HT_ERR_INVALID_ARG;
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  h->arena.reset();
  Engine e;
  e.st = h->stream;
  e.arena = &h->arena;
  ht_default_options(&e.opt);
  CKS(e.tall_orth(Y, m, b, Q));
  if (fallback) *fallback = e.orth_fallbacks;
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_skinny_gemm(ht_handle_t h, const float* A, int64_t M, int64_t K, int32_t transA, const float* B, int32_t b,
                         int32_t precision, float* Y) {
  if (!h || !A || !B || !Y || M < 1 || K < 1 || b < 1 || b > 64) return HT_ERR_INVALID_ARG;
  const int prec = precision == HT_PREC_FP32 ? PREC_F32
                   : precision == HT_PREC_FP16 ? PREC_F16
                   : precision == HT_PREC_FP8  ? PREC_FP8
                                               : PREC_BF16;
  h->arena.reset();
  const size_t wsf = gemm_ws_floats(M, K, b);
  float* ws = (float*)h->arena.get(wsf * 4);
  if (!ws) return HT_ERR_ALLOC;
  if (transA) CKS(gemm_atb_launch(h->stream, A, M, K, B, b, prec, Y, ws, wsf));
  else CKS(gemm_ab_launch(h->stream, A, M, K, B, b, prec, Y, ws, wsf));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_skinny_gemm16(ht_handle_t h, const float* A, int64_t M, int64_t K, int32_t transA, const float* B,
                           int32_t b, int32_t precision, float* Y) {
  if (!h || !A || !B || !Y || M < 1 || K < 1 || b < 1 || b > 64) return HT_ERR_INVALID_ARG;
  if (precision != HT_PREC_BF16 && precision != HT_PREC_FP16) return HT_ERR_UNSUPPORTED;
  const int prec = precision == HT_PREC_FP16 ? PREC_F16 : PREC_BF16;
  if (((uintptr_t)A & 15) || ((uintptr_t)B & 15) || M % (transA ? 8 : 4)) return HT_ERR_UNSUPPORTED;
  h->arena.reset();
  const size_t wsf = gemm_ws_floats(M, K, b);
  float* ws = (float*)h->arena.get(wsf * 4);
  uint16_t* A16 = (uint16_t*)h->arena.get((size_t)M * K * sizeof(uint16_t));
  double* part = (double*)h->arena.get((size_t)cast16_part_size() * sizeof(double) + 8);
  if (!ws || !A16 || !part) return HT_ERR_ALLOC;
  CKS(cast16_sumsq_launch(h->stream, A, M * K, prec, A16, part, part + cast16_part_size()));
  if (transA) CKS(gemm_atb16_launch(h->stream, A16, M, K, B, b, prec, Y, ws, wsf));
  else CKS(gemm_ab16_launch(h->stream, A16, M, K, B, b, prec, Y, ws, wsf));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

ht_status ht_philox_normals(ht_handle_t h, uint64_t seed, uint32_t stream, int64_t n, double* out_dev) {
  if (!h || !out_dev || n < 0) return HT_ERR_INVALID_ARG;
  CKS(fill_normals_f64_launch(h->stream, out_dev, n, seed, stream));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

void ht_philox_normals_host(uint64_t seed, uint32_t stream, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = philox_normal(seed, stream, (uint64_t)i);
}

}  // extern "C"
