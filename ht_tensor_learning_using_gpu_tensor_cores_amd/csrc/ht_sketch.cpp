// Learned per-mode sketch: layout, the nodes it serves, and the training / evaluation
// entry points over Gaussian-bump tensors generated on the device (declarations in
// include/ht_amd.h). No counterpart in the reference; it replaces the dense Gaussian test
// matrix of fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:419-444.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ht_engine.h"

using namespace ht;

#define CKS(expr)                 \
  do {                            \
    int _s = (expr);              \
    if (_s) return (ht_status)_s; \
  } while (0)

int ht::sketch_offsets(int order, const int64_t* dims, int w, int64_t* off, int64_t* total) {
  if (order < 1 || order > HT_MAX_ORDER || !dims || w < 1 || w > 64) return HT_ERR_INVALID_ARG;
  int64_t o = 0;
  for (int m = 0; m < order; ++m) {
    if (dims[m] < 1) return HT_ERR_INVALID_ARG;
    if (off) off[m] = o;
    o += (int64_t)w * dims[m];
  }
  if (total) *total = o;
  return 0;
}

// Same traversal as the randomised method: a non-root interior node gets a range finder
// unless its sibling (a child of the root) already produced its basis from the right
// singular vectors.
int ht::sketched_nodes(const Problem& p, std::vector<int>& nodes) {
  nodes.clear();
  std::vector<int> done(p.n_nodes, 0);
  for (int t = 1; t < p.n_nodes; ++t) {
    if (p.left[t] < 0 || done[t]) continue;
    if (!(p.lo[t] == 0 || p.hi[t] == p.order)) return HT_ERR_UNSUPPORTED;
    nodes.push_back(t);
    if (p.parent[t] == 0) {
      const int sib = (p.left[0] == t) ? p.right[0] : p.left[0];
      if (p.left[sib] >= 0) done[sib] = 1;
    }
  }
  return 0;
}

int ht::node_col_modes(const Problem& p, int t, int* m0, int* m1) {
  if (p.lo[t] == 0 && p.hi[t] < p.order) {
    *m0 = p.hi[t];
    *m1 = p.order;
  } else if (p.hi[t] == p.order && p.lo[t] > 0) {
    *m0 = 0;
    *m1 = p.lo[t];
  } else {
    return HT_ERR_UNSUPPORTED;
  }
  return *m1 - *m0 <= HT_KR_MAX ? 0 : HT_ERR_UNSUPPORTED;
}

// Parameters of sample `sample`: a = 0.5 + 1.5 u0, w = 0.1 + 0.3 u1, c_m = u_{2+m}
// (Philox uniforms of stream 96, n_bumps * (2 + order) per sample); factor
// g_{j,m}(i) = exp(-(i/(n-1) - c_jm)^2 / (2 w_j^2)).
void ht::bump_factors_host(int order, long n, int n_bumps, uint64_t seed, int64_t sample, float* amp, float* fac) {
  const int per = n_bumps * (2 + order);
  for (int j = 0; j < n_bumps; ++j) {
    auto u = [&](int c) { return philox_uniform(seed, 96, (uint64_t)sample * per + (uint64_t)j * (2 + order) + c); };
    const double a = 0.5 + 1.5 * u(0), w = 0.1 + 0.3 * u(1);
    amp[j] = (float)a;
    for (int m = 0; m < order; ++m) {
      const double c = u(2 + m);
      for (long i = 0; i < n; ++i) {
        const double g = (double)i / (double)(n - 1);
        fac[((size_t)j * order + m) * n + i] = (float)std::exp(-(g - c) * (g - c) / (2.0 * w * w));
      }
    }
  }
}

void ht_sketch_default_train_options(ht_sketch_train_options* o) {
  if (!o) return;
  o->width = 20;
  o->epochs = 2;
  o->batch = 8;
  o->lr = 0.05f;
  o->beta1 = 0.9f;
  o->beta2 = 0.999f;
  o->eps = 1e-8f;
  o->min_loss = 1e-5f;
}

ht_status ht_sketch_layout(int32_t order, const int64_t* dims, int32_t width, int64_t* offsets, int64_t* total) {
  return (ht_status)sketch_offsets(order, dims, width, offsets, total);
}

ht_status ht_sketch_init_gaussian(ht_handle_t h, int32_t order, const int64_t* dims, int32_t width, uint64_t seed,
                                  float* sketch_dev) {
  if (!h || !sketch_dev) return HT_ERR_INVALID_ARG;
  int64_t total = 0;
  CKS(sketch_offsets(order, dims, width, nullptr, &total));
  if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
  CKS(fill_normals_f32_launch(h->stream, sketch_dev, total, seed, 200));
  return hipStreamSynchronize(h->stream) == hipSuccess ? HT_OK : HT_ERR_HIP;
}

namespace {

// Shared set-up of the sketch entry points: problem (ranks 1: the loss needs none), the
// sketched nodes, offsets, and an engine on the handle's stream and arena.
struct SketchCtx {
  Problem p;
  std::vector<int> nodes;
  int64_t off[HT_MAX_ORDER] = {}, total = 0;
  Engine e;
  int init(ht_handle_t h, int order, const int64_t* dims, int n_nodes, const int32_t* lo, const int32_t* hi,
           const int32_t* left, const int32_t* right, int w) {
    if (!h || !dims) return HT_ERR_INVALID_ARG;
    std::vector<int32_t> ranks(std::max(n_nodes, 1), 1);
    int s = p.setup(order, dims, n_nodes, lo, hi, left, right, ranks.data(), 0);
    if (s) return s;
    s = sketched_nodes(p, nodes);
    if (s) return s;
    if (nodes.empty()) return HT_ERR_UNSUPPORTED;
    s = sketch_offsets(order, dims, w, off, &total);
    if (s) return s;
    if (hipSetDevice(h->device) != hipSuccess) return HT_ERR_HIP;
    h->arena.reset();
    e.st = h->stream;
    e.arena = &h->arena;
    e.prob = &p;
    ht_default_options(&e.opt);
    e.prof = nullptr;
    for (int t = 0; t < HT_MAX_NODES; ++t) {
      e.Ufull[t] = nullptr;
      e.Bd[t] = nullptr;
      e.SVd[t] = nullptr;
    }
    return 0;
  }
};

// Device factors of samples [first, first + count) (amplitudes, factors).
int upload_bumps(Engine& e, int order, long n, int n_bumps, uint64_t seed, int64_t first, int64_t count, float** amp_d,
                 float** fac_d) {
  const size_t per_f = (size_t)n_bumps * order * n;
  std::vector<float> amp((size_t)count * n_bumps), fac((size_t)count * per_f);
  for (int64_t s = 0; s < count; ++s)
    bump_factors_host(order, n, n_bumps, seed, first + s, &amp[(size_t)s * n_bumps], &fac[(size_t)s * per_f]);
  *amp_d = e.alloc<float>(amp.size());
  *fac_d = e.alloc<float>(fac.size());
  if (!*amp_d || !*fac_d) return HT_ERR_ALLOC;
  HT_HIP_CHECK(hipMemcpyAsync(*amp_d, amp.data(), amp.size() * 4, hipMemcpyHostToDevice, e.st));
  HT_HIP_CHECK(hipMemcpyAsync(*fac_d, fac.data(), fac.size() * 4, hipMemcpyHostToDevice, e.st));
  return e.sync();  // the host vectors go out of scope
}

// The training and evaluation loops use the plane generator with its fused ||X||^2.
int bump_args_ok(int order, long n, int n_bumps, int64_t first, int64_t count) {
  return order >= 2 && order <= HT_MAX_ORDER && n >= 2 && n <= 1024 && n_bumps >= 1 && n_bumps <= 8 && first >= 0 &&
         count >= 1;
}

}  // namespace

ht_status ht_sketch_loss_grad(ht_handle_t h, const float* x_dev, int32_t order, const int64_t* dims, int32_t n_nodes,
                              const int32_t* lo, const int32_t* hi, const int32_t* left, const int32_t* right,
                              int32_t width, const float* sketch_dev, float* grad_dev, double* loss) {
  if (!x_dev || !sketch_dev || !loss) return HT_ERR_INVALID_ARG;
  SketchCtx c;
  CKS(c.init(h, order, dims, n_nodes, lo, hi, left, right, width));
  Engine& e = c.e;
  double* xsq = e.alloc<double>(1 + 1024);
  double* lacc = e.alloc<double>(1);
  int* info = e.alloc<int>(1);
  if (!xsq || !lacc || !info) return HT_ERR_ALLOC;
  CKS(fill_f64_launch(e.st, lacc, 1, 0.0));
  if (hipMemsetAsync(info, 0, sizeof(int), e.st) != hipSuccess) return HT_ERR_HIP;
  if (grad_dev) CKS(fill_f32_launch(e.st, grad_dev, c.total, 0.f));
  CKS(sumsq_f32_launch(e.st, x_dev, c.p.N, xsq));
  const float scale = 1.0f / (float)c.nodes.size();
  for (int t : c.nodes) CKS(e.sketch_node(x_dev, t, sketch_dev, c.off, width, grad_dev, scale, lacc, xsq, info));
  double lv = 0.0;
  int iv = 0;
  CKS(e.d2h(&lv, lacc, sizeof(double)));
  CKS(e.d2h(&iv, info, sizeof(int)));
  *loss = lv / (double)c.nodes.size();
  if (!std::isfinite(*loss)) return HT_ERR_NONFINITE;
  return iv ? HT_ERR_NOCONV : HT_OK;
}

ht_status ht_sketch_train_bumps(ht_handle_t h, int32_t order, int64_t n, int32_t n_bumps, uint64_t seed,
                                int64_t first_sample, int64_t count, int32_t n_nodes, const int32_t* lo,
                                const int32_t* hi, const int32_t* left, const int32_t* right,
                                const ht_sketch_train_options* opt_in, float* sketch_dev, double* epoch_loss) {
  ht_sketch_train_options opt;
  if (opt_in) opt = *opt_in;
  else ht_sketch_default_train_options(&opt);
  if (!sketch_dev || !epoch_loss || !bump_args_ok(order, n, n_bumps, first_sample, count) || opt.epochs < 1 ||
      opt.batch < 1 || !(opt.lr > 0) || !(opt.beta1 >= 0 && opt.beta1 < 1) || !(opt.beta2 >= 0 && opt.beta2 < 1) ||
      !(opt.eps > 0) || !(opt.min_loss >= 0))
    return HT_ERR_INVALID_ARG;
  std::vector<int64_t> dims(order, n);
  SketchCtx c;
  CKS(c.init(h, order, dims.data(), n_nodes, lo, hi, left, right, opt.width));
  Engine& e = c.e;
  const int w = opt.width;
  float* X = e.alloc<float>((size_t)c.p.N);
  float* grad = e.alloc<float>(c.total);
  float* m1 = e.alloc<float>(c.total);
  float* m2 = e.alloc<float>(c.total);
  double* xsq = e.alloc<double>(1);
  double* part = e.alloc<double>((size_t)gen_bumps_planes(order, n));
  double* lacc = e.alloc<double>(1);
  double* sacc = e.alloc<double>(1);
  int* info = e.alloc<int>(1);
  float *amp_d = nullptr, *fac_d = nullptr;
  if (!X || !grad || !m1 || !m2 || !xsq || !part || !lacc || !sacc || !info) return HT_ERR_ALLOC;
  CKS(upload_bumps(e, order, n, n_bumps, seed, first_sample, count, &amp_d, &fac_d));
  CKS(fill_f32_launch(e.st, grad, c.total, 0.f));
  CKS(fill_f32_launch(e.st, m1, c.total, 0.f));
  CKS(fill_f32_launch(e.st, m2, c.total, 0.f));
  if (hipMemsetAsync(info, 0, sizeof(int), e.st) != hipSuccess) return HT_ERR_HIP;
  const size_t per_f = (size_t)n_bumps * order * n;
  const float node_scale = 1.0f / (float)c.nodes.size();
  long step = 0;
  CKS(fill_f64_launch(e.st, sacc, 1, 0.0));
  for (int ep = 0; ep < opt.epochs; ++ep) {
    CKS(fill_f64_launch(e.st, lacc, 1, 0.0));
    int in_step = 0;
    for (int64_t s = 0; s < count; ++s) {
      CKS(gen_bumps_launch(e.st, X, order, n, fac_d + (size_t)s * per_f, amp_d + (size_t)s * n_bumps, n_bumps, xsq,
                           part));
      const auto mk = h->arena.mark();
      for (int t : c.nodes) CKS(e.sketch_node(X, t, sketch_dev, c.off, w, grad, node_scale, sacc, xsq, info));
      h->arena.rewind(mk);
      ++in_step;
      if (in_step == opt.batch || s == count - 1) {
        ++step;
        const float bc1 = 1.0f - std::pow(opt.beta1, (float)step), bc2 = 1.0f - std::pow(opt.beta2, (float)step);
        CKS(adam_launch(e.st, sketch_dev, grad, m1, m2, c.total, opt.lr, opt.beta1, opt.beta2, opt.eps,
                        1.0f / (float)in_step, bc1, bc2, sacc,
                        (double)opt.min_loss * in_step * (double)c.nodes.size()));
        CKS(fold_loss_launch(e.st, sacc, lacc));
        in_step = 0;
      }
    }
    double lv = 0.0;
    CKS(e.d2h(&lv, lacc, sizeof(double)));
    epoch_loss[ep] = lv / ((double)count * (double)c.nodes.size());
  }
  int iv = 0;
  CKS(e.d2h(&iv, info, sizeof(int)));
  for (int ep = 0; ep < opt.epochs; ++ep)
    if (!std::isfinite(epoch_loss[ep])) return HT_ERR_NONFINITE;
  return iv ? HT_ERR_NOCONV : HT_OK;
}

ht_status ht_sketch_eval_bumps(ht_handle_t h, int32_t order, int64_t n, int32_t n_bumps, uint64_t seed,
                               int64_t first_sample, int64_t count, int32_t n_nodes, const int32_t* lo,
                               const int32_t* hi, const int32_t* left, const int32_t* right, int32_t width,
                               const float* sketch_dev, double* loss) {
  if (!sketch_dev || !loss || !bump_args_ok(order, n, n_bumps, first_sample, count)) return HT_ERR_INVALID_ARG;
  std::vector<int64_t> dims(order, n);
  SketchCtx c;
  CKS(c.init(h, order, dims.data(), n_nodes, lo, hi, left, right, width));
  Engine& e = c.e;
  float* X = e.alloc<float>((size_t)c.p.N);
  double* xsq = e.alloc<double>(1);
  double* part = e.alloc<double>((size_t)gen_bumps_planes(order, n));
  double* lacc = e.alloc<double>(count);
  int* info = e.alloc<int>(1);
  float *amp_d = nullptr, *fac_d = nullptr;
  if (!X || !xsq || !part || !lacc || !info) return HT_ERR_ALLOC;
  CKS(upload_bumps(e, order, n, n_bumps, seed, first_sample, count, &amp_d, &fac_d));
  CKS(fill_f64_launch(e.st, lacc, count, 0.0));
  if (hipMemsetAsync(info, 0, sizeof(int), e.st) != hipSuccess) return HT_ERR_HIP;
  const size_t per_f = (size_t)n_bumps * order * n;
  for (int64_t s = 0; s < count; ++s) {
    CKS(gen_bumps_launch(e.st, X, order, n, fac_d + (size_t)s * per_f, amp_d + (size_t)s * n_bumps, n_bumps, xsq,
                         part));
    const auto mk = h->arena.mark();
    for (int t : c.nodes) CKS(e.sketch_node(X, t, sketch_dev, c.off, width, nullptr, 0.f, lacc + s, xsq, info));
    h->arena.rewind(mk);
  }
  CKS(e.d2h(loss, lacc, sizeof(double) * count));
  int iv = 0;
  CKS(e.d2h(&iv, info, sizeof(int)));
  for (int64_t s = 0; s < count; ++s) {
    loss[s] /= (double)c.nodes.size();
    if (!std::isfinite(loss[s])) return HT_ERR_NONFINITE;
  }
  return iv ? HT_ERR_NOCONV : HT_OK;
}
