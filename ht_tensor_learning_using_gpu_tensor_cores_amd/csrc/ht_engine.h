// Internal host-side types of the HT engine.
#pragma once
#include <atomic>
#include <map>
#include <utility>
#include <vector>

#include "../../include/ht_amd.h"
#include "ht_common.h"

namespace ht {

// Bump allocator over persistent device chunks; reset() rewinds without freeing, so a
// repeated call of the same shape allocates nothing after the first.
//
// A fixed arena serves one caller-provided block (fbase, fsize) and never allocates: a
// request beyond it returns nullptr and sets `overflow`.
struct Arena {
  std::vector<std::pair<void*, size_t>> chunks;
  size_t cur = 0, off = 0;
  // bytes handed out since reset() as if from one contiguous block, and their maximum:
  // the size a fixed arena needs for the same sequence of requests
  size_t lin = 0, peak = 0;
  char* fbase = nullptr;
  size_t fsize = 0;
  bool fixed = false, overflow = false;
  void* get(size_t bytes);
  void reset();
  void release();
  // Rewind to an earlier position: what was handed out after `mark()` is reused.
  struct Mark {
    size_t cur, off, lin;
  };
  Mark mark() const { return {cur, off, lin}; }
  void rewind(Mark m) {
    cur = m.cur;
    off = m.off;
    lin = m.lin;
  }
};

int validate_tree(int order, int n_nodes, const int32_t* lo, const int32_t* hi, const int32_t* left,
                  const int32_t* right);

// Validated problem description and packed output layout.
struct Problem {
  int order = 0, n_nodes = 0, want_interior_u = 0;
  int64_t dims[HT_MAX_ORDER] = {};
  int64_t N = 0;
  int lo[HT_MAX_NODES], hi[HT_MAX_NODES], left[HT_MAX_NODES], right[HT_MAX_NODES], parent[HT_MAX_NODES];
  int rank[HT_MAX_NODES];
  int64_t rows[HT_MAX_NODES];
  int64_t u_off[HT_MAX_NODES], b_off[HT_MAX_NODES], s_off[HT_MAX_NODES];
  int64_t total = 0;
  int setup(int order, const int64_t* dims, int n_nodes, const int32_t* lo, const int32_t* hi,
            const int32_t* left, const int32_t* right, const int32_t* ranks, int want_interior_u);
  void layout();  // u_off, b_off, s_off and total of the current ranks
  int depth(int t) const;
  int leaf_of_mode(int m) const;
};

// Per-kernel device timing with HIP events recorded on the launching stream.
enum ProfKernel {
  PK_SLAB_GRAM = 0, PK_SLAB_PROJ, PK_GRAM, PK_RESIDUAL, PK_GEMM, PK_TSQR, PK_EIG, PK_SMALL, PK_GEN, PK_COUNT
};
struct Profiler {
  static constexpr int NREC = 1024;
  hipEvent_t ev[2 * NREC];
  int kid[NREC];
  int n = 0;
  bool created = false, enabled = false;
  int mask = -1;  // kernel ids timed (bit 1 << id)
  double ms[PK_COUNT] = {};
  long cnt[PK_COUNT] = {};
  int create();  // the events, made once (ht_profile_enable makes them before timed work)
  int begin(hipStream_t st, int k);
  void end(hipStream_t st, int slot);
  void collect();
  void destroy();
};

// Host decisions of the leaves-to-root method, recorded by a run that reads every
// eigensolver round's results and replayed by later runs of the same problem and options,
// which then read nothing back until the end of the run. Each replayed eigensolver call
// writes the convergence flags of its last round (and each small-side SVD its rank test)
// into a device verdict array that is read with the error at the end: any zero there
// means the replay is discarded and the run is repeated with reads (and re-recorded).
struct EigRound {
  int q, mode, b, try_chol;
  int widen;  // block width of the next round when it grows (0: the Ritz block is reused)
};
struct EigCall {
  int n, k, batch, b0, q0;
  double tol;
  std::vector<EigRound> rounds;
};
struct RunPlan {
  std::vector<int64_t> key;  // problem and options the plan was recorded for
  std::vector<EigCall> calls;
  std::vector<int> small;    // outcome of each small-side SVD (1: used, 0: fell back)
  bool valid = false;
};

// Side streams for the large leaves' eigen problems (leaf i on stream i % NSIDE) and
// their events: one per leaf whose Gram a solve waits for, and one join per stream.
constexpr int NSIDE = 3;
constexpr int SIDE_EVENTS = HT_MAX_ORDER + NSIDE;

// One large-leaf subspace iteration in progress (sym_topk_large_begin / _step): its
// stream, buffers and the host state of its rounds.
struct LargeEig {
  hipStream_t st = nullptr;
  const double* G = nullptr;
  double *U = nullptr, *w = nullptr;
  int n = 0, k = 0, b = 0, maxit = 0;
  uint32_t stream = 0;
  double tol = 0.0, floor = 0.0;
  double *Q = nullptr, *Y = nullptr, *T = nullptr, *S = nullptr, *th = nullptr, *W = nullptr, *res = nullptr,
         *Zt = nullptr, *Ri = nullptr, *ws = nullptr, *QS = nullptr, *YS = nullptr;
  int* info = nullptr;
  int it = 0, steps = 2;
  bool done = false;
};

struct Engine {
  Profiler* prof = nullptr;
  hipStream_t st;
  Arena* arena;
  Problem* prob;  // ranks change during a rank-adaptive run
  ht_options opt;
  // Rank-adaptive truncation: a non-root node's rank on entry is its cap; once the node's
  // spectrum is known its rank becomes the smallest r whose truncation removes at most
  // tau2 of energy (trace of the current Gram minus its r largest eigenvalues).
  bool adaptive = false;
  double adapt_eps = 0.0;  // rel_eps
  double tau2 = 0.0;       // (rel_eps ||X||_F)^2 / (number of truncations the error adds up over)
  double adapt_xsq = 0.0;  // ||X||_F^2 (host)
  int adapt_tau(const double* xsq_dev, int truncations);
  // w_dev: k eigenvalues (sq = 1) or singular values (sq = 0), descending; trace: energy of
  // the tensor the node is truncated in
  int choose_rank(int t, const double* w_dev, int k, double trace, int sq);
  int gram_trace(const double* G, long n, double* tr);
  // per-node device data (FP64): full bases (leaves; exact interior bases for plain),
  // transfer tensors, singular values
  double* Ufull[HT_MAX_NODES];
  double* Bd[HT_MAX_NODES];
  double* SVd[HT_MAX_NODES];
  int SVsq[HT_MAX_NODES] = {};  // 1: SVd[t] holds eigenvalues (squared singular values)
  double passes = 0.0;
  // recorded on st after the run's first pass over X when set, then *pass_flag =
  // pass_mark (ht_decompose_many starts each decomposition of a worker after the first
  // pass of the previous worker's decomposition of the same round)
  hipEvent_t pass_ev = nullptr;
  std::atomic<int>* pass_flag = nullptr;
  int pass_mark = 1;
  // other decompositions run beside this one (ht_decompose_many): the slab passes, one
  // work-group per CU for their whole length, leave an eighth of the CUs to the others'
  // single-work-group solver kernels, which cannot share a CU with them
  bool shared_gpu = false;
  void note_pass(double amount);
  int eig_iters_max = 0, eig_calls = 0, noconv = 0;
  int orth_fallbacks = 0;  // tall_orth calls that took the TSQR basis
  // 16-bit copy of X (values rounded to precision x16_prec) that the Grams of the views
  // gram16_ok accepts read instead of X; nullptr when the run made none
  const uint16_t* x16 = nullptr;
  int x16_prec = 0;
  // Large-leaf eigen problems whose Grams are enqueued and whose side-stream solves wait
  // for solve_pending_leaves (leaves_from_x with defer = true), so that the host enqueues
  // the work after them on st first.
  struct LeafSolves {
    std::vector<int> grp;
    long n = 0;
    int k = 0;
    double *G = nullptr, *U = nullptr, *w = nullptr;
    std::vector<double*> Q0;
    int ev0 = 0;  // side_ev[ev0 + i]: recorded after leaf i's Gram
  };
  std::vector<LeafSolves> pending_leaves;
  int solve_pending_leaves();
  int solve_large_leaves(const LeafSolves& j);
  // orth_defer: tall_orth leaves its pivot check to the end of the run (one device flag,
  // orth_flag, for every call; a run that finds it set is repeated with the checks)
  bool orth_defer = false;
  int* orth_flag = nullptr;
  int replayed = 0;  // 1: the run replayed the recorded plan and passed its verdicts
  int sketch_fallbacks = 0, sketch_nodes = 0;
  double sketch_resid = 0.0;
  double rel_err = 0.0, x_sq = 0.0;
  // pinned host staging for the small device-to-host reads (flags, Ritz values, scalars):
  // copies into pageable memory complete synchronously one by one
  char* pin = nullptr;
  size_t pin_bytes = 0;
  // The handle's NSIDE side streams and their events (SIDE_EVENTS), made on first use:
  // the eigen problems of large leaves run there beside the work that follows their
  // Grams on st. Null: no side streams (they run on st).
  hipStream_t* side = nullptr;
  hipEvent_t* side_ev = nullptr;
  // record / replay of the host decisions (leaves-to-root only)
  RunPlan* plan = nullptr;
  bool recording = false, replay = false, plan_mismatch = false;
  size_t call_idx = 0, small_idx = 0;
  int* verdict = nullptr;  // device verdict slots of a replayed run
  int n_verdict = 0;
  // reads deferred to the end of a replayed run: the error terms (into the upper half of
  // the pinned staging) and the verdicts after them
  bool deferred = false;
  int deferred_mode = 0;
  // event recorded after the last work of the run, before its final wait (optional)
  hipEvent_t done_ev = nullptr;
  bool done_recorded = false;

  template <typename T>
  T* alloc(size_t n);
  int sync();
  int d2h(void* dst, const void* src, size_t bytes);
  int gram(const float* x, PMS v, int prec, double* G, double* sumsq);
  int gram2(const float* x, PMS v1, PMS v2, int prec, double* G1, double* G2);
  int make_x16(const float* x, int prec, double* xsq);
  // Top-k eigenpairs of symmetric G (FP64). A Ritz pair counts as converged when its
  // residual is <= tol * th_1 or its Ritz value is <= floor * th_1 (below what the Gram's
  // own accumulation error resolves, where any basis of the cluster serves).
  int sym_topk(const double* G, int n, int k, double* U, double* w, double tol, int maxit, uint32_t stream,
               double floor);
  // G32in (optional): an FP32 copy of G for the filter steps (used when tol >= 1e-5).
  int sym_topk_batch(const double* G, int n, int k, int batch, double* U, double* w, double tol, int q0, int b,
                     uint32_t stream, double floor, const float* G32in = nullptr, double tol_rel = 0.0);
  int canon_signs(double* U, int m, int k, int ldu);
  int sym_topk_large(const double* G, int n, int k, double* U, double* w, double tol, int maxit, uint32_t stream,
                     double floor, double* Q0 = nullptr);
  // The same solve in pieces, on stream L.st: _begin allocates and enqueues the first
  // round; _step reads a round's residuals (waits for L.st) and enqueues the next round,
  // or the result copies once converged (L.done).
  int sym_topk_large_begin(LargeEig& L, const double* G, int n, int k, double* U, double* w, double tol, int maxit,
                           uint32_t stream, double floor, double* Q0);
  int sym_topk_large_step(LargeEig& L);
  int sym_topk_large_round(LargeEig& L);
  int tsqr(float* Y, long m, int b, float* Q, double* R);
  int tall_orth(float* Y, long m, int b, float* Q, int passes = 3);
  bool side_ready();
  int view_gemm(const float* x, PMS v, int op, const float* B, int b, int prec, float* Y);
  int exact_svd_subspace(const float* x, PMS v, int kU, int kW, double* U, double* Wr, double* sv, uint32_t stream);
  int contract_transfer(const double* Ut, const double* Ul, const double* Ur, long nl, long nr, int kl, int kr,
                        int kt, double* B);
  int expand(int t, double** out);
  int expand2(int ta, int tb, double** oa, double** ob);
  int residual(const float* x, double* err_sq, double* x_sq);
  int leaves_from_x_plain(const float* x, double* xsq_dev);
  int leaves_from_x(const float* x, int prec, double* xsq_dev, bool defer = false);
  int eig_to_sv(const double* w, int t);
  int early_finite_check(const double* xsq_dev);
  int root_transfer(const float* x);
  // Randomised method in FP32: the root child rz_t whose range finder ran keeps its
  // co-range product Z = X_(rz_t)^T Q (rz_rows x rz_l) and S' = Q^T U_rz_t (rz_l x k);
  // the root transfer then needs no pass over X.
  int rz_t = -1, rz_l = 0;
  long rz_rows = 0;
  const float* rz_Z = nullptr;
  const double* rz_Sv = nullptr;
  int root_transfer_from_z();
  int finish_error(const float* x, const double* xsq_dev);
  int run_plain(const float* x);
  int run_ltr(const float* x);
  int ltr_interior(double* C, int64_t* cd, const double* xsq, const float* x);
  int small_side_svd(const double* C, PMS v, int k, double* U, const double** w, int* done);
  int run_rsvd(const float* x);
  int rsvd_pass(const float* x);
  void reset_run_state();
  int range_finder(const float* x, PMS v, const float* Om, int l, int prec, float* Y, float* Q, float* Z, float* Qz);
  int range_residual(const float* x, PMS v, int t, const float* Q, int l, const double* xsq, double* est);
  // One sketched node's loss (accumulated into loss_acc[0]) and, when grad != nullptr,
  // scale times its gradient (accumulated into grad).
  int sketch_node(const float* x, int t, const float* S, const int64_t* off, int w, float* grad, float scale,
                  double* loss_acc, const double* xsq, int* info);
  int write_outputs(float* out);
  int run_once(const float* x, float* out);
  int finish_deferred(bool* verified);
  int run(const float* x, float* out);
};

}  // namespace ht

struct ht_handle_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {};
  // side streams for large leaves' eigen problems (Engine::side), made on first use
  hipStream_t side[ht::NSIDE] = {};
  // ht_decompose_many: worker w records stagger_ev[w][r % 4] after the first pass over X
  // of its decomposition of round r
  hipEvent_t stagger_ev[HT_MAX_INFLIGHT][4] = {};
  hipEvent_t side_ev[ht::SIDE_EVENTS] = {};
  ht::Arena arena;
  float* xbuf = nullptr;   // device copy of host inputs
  size_t xbuf_bytes = 0;
  float* obuf = nullptr;   // device output staging for host outputs
  size_t obuf_bytes = 0;
  ht::Profiler prof;
  // batched entry point: autotuned launch shape per (count, largest mode size) and the
  // graph of the last launch with the arguments it was captured for
  std::map<std::pair<int, int>, std::pair<int, int>> batch_tune;
  hipGraph_t bgraph = nullptr;
  hipGraphExec_t bexec = nullptr;
  std::vector<const void*> bkey;
  // host decisions of the last leaves-to-root run, replayed by the next of the same shape
  ht::RunPlan plan;
  // pinned staging for small device-to-host reads
  char* pinned = nullptr;
  size_t pinned_bytes = 0;
  // ht_decompose_many: worker handles (own stream, workspace, plan), made on first use
  std::vector<ht_handle_s*> workers;
};

namespace ht {
// Learned-sketch helpers (ht_sketch.cpp).
int sketch_offsets(int order, const int64_t* dims, int w, int64_t* off, int64_t* total);
// Nodes the randomised method runs a range finder on, in its order.
int sketched_nodes(const Problem& p, std::vector<int>& nodes);
// Column modes [m0, m1) of node t's matricisation.
int node_col_modes(const Problem& p, int t, int* m0, int* m1);
// Amplitudes (n_bumps) and per-mode factors (n_bumps x order x n) of one bump sample.
void bump_factors_host(int order, long n, int n_bumps, uint64_t seed, int64_t sample, float* amp, float* fac);
// Generators shared by the entry points (ht_api.cpp).
std::vector<double> orthonormal_factor(uint64_t seed, uint32_t stream, long n, int r);
ht_status generate_from_core(ht_handle_t h, int order, const int64_t* dims, const int32_t* core_dims,
                             std::vector<double> core, std::vector<std::vector<double>> U, uint64_t seed,
                             double noise_rel, float* x_dev, uint32_t noise_stream);
}  // namespace ht
