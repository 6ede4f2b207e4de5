/*
 * ht_amd.h -- C ABI of the MI355X hierarchical Tucker (HT) decomposition engine.
 *
 * A dense FP32 tensor (column-major, mode 1 fastest), a binary dimension tree and one
 * rank per tree node go in; leaf bases U_t, transfer tensors B_t and the relative
 * reconstruction error come out. Plain pointers and sizes only: no HIP, torch or
 * library types cross this boundary.
 *
 * Reference surface this replaces (paths relative to the reference repository):
 *   float htd(dt*,long,long,long,int*,dt,dt)         third_order_tensor_decompsition/CUDA/Baseline/function.h:6
 *   float htd(dt*,long,long,long,int*)               third_order_tensor_decompsition/CUDA/Optimized/function.h:6
 *   void  htd(dt*,long,long,long,int*)               third_order_tensor_decompsition/CUDA/TSQR/function.h:6
 *   void  htd4(dt*,long,long,long,long,int*)         fourth_order_tensor_decomposition/single_GPU/CUDA/Baseline/function.h:6
 *   void  htd4(dt*,long,long,long,long,int*,cublasHandle_t,cusolverDnHandle_t)
 *                                                    fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/function.h:6
 * The compatibility wrappers at the end keep those names and argument meanings; the
 * general entry point ht_decompose adds what the reference hard-codes (the tree), what
 * it throws away (U_t, B_t are returned) and status codes (the reference reports none).
 *
 * Conventions shared by every entry point:
 *   - Tree nodes are contiguous mode ranges [lo, hi); node 0 is the root; a leaf holds
 *     one mode; an interior node's left child holds its lower modes. Nodes are numbered
 *     breadth-first, which is the reference's heap numbering for its complete trees
 *     (third_order_tensor_decompsition/Python/HT_TensorLy.py:23).
 *   - rank[0] must be 1 (k[0]=1, third_order_tensor_decompsition/CUDA/Baseline/ht_init.cu:16-19).
 *   - U_t is (rows_t x k_t) column-major, columns in descending singular-value order,
 *     each column signed so that its largest-magnitude entry is positive.
 *   - B_t is (k_l x k_r x k_t) column-major (left child's rank index fastest), the layout
 *     of the reference's d_B buffers (third_order_tensor_decompsition/CUDA/Baseline/htd_unopt.cu:105-108).
 */
#ifndef HT_AMD_H
#define HT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HT_MAX_ORDER 8
#define HT_MAX_NODES (2 * HT_MAX_ORDER - 1)
#define HT_MAX_RANK 64

typedef enum {
  HT_OK = 0,
  HT_ERR_INVALID_ARG = 1,   /* null pointer, bad order, non-positive size            */
  HT_ERR_INVALID_TREE = 2,  /* not a binary contiguous tree covering every mode       */
  HT_ERR_RANK = 3,          /* rank < 1, rank > rows, rank > k_l*k_r, rank > HT_MAX_RANK */
  HT_ERR_ALLOC = 4,         /* device or host allocation failed                       */
  HT_ERR_HIP = 5,           /* a HIP runtime call failed                              */
  HT_ERR_NONFINITE = 6,     /* the input holds NaN or Inf                             */
  HT_ERR_NOCONV = 7,        /* an eigensolver did not reach its tolerance             */
  HT_ERR_UNSUPPORTED = 8,   /* a shape this build does not handle                     */
  HT_ERR_NO_DEVICE = 9,     /* no usable GPU                                          */
  HT_ERR_WORKSPACE = 10     /* ht_decompose_noalloc: the problem needs more workspace than
                               HT_NOALLOC_SCRATCH_BYTES                               */
} ht_status;

/* Method selection. */
typedef enum {
  /* Reference semantics: every non-root node's basis is the exact truncated SVD of its
   * own matricisation of X, computed in FP32 on the GPU (FP64 for the small eigen
   * problems). third_order_tensor_decompsition/CUDA/Baseline/htd_unopt.cu:2-297. */
  HT_METHOD_PLAIN = 0,
  /* Leaves-to-root: leaf Grams on MFMA (BF16/FP16 in, FP32 accumulate), then a single
   * projection pass to the leaf core and exact FP64 interior problems on that core. */
  HT_METHOD_LTR = 1,
  /* Root-to-leaves randomised: every non-root node by a range finder (oversampling p,
   * q power iterations) on MFMA, shifted CholeskyQR orthonormalisation (three passes for
   * the final block, one for the intermediate ones), FP64 projected SVD.
   * fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:419-504. */
  HT_METHOD_RSVD = 2
} ht_method;

/* HT_PREC_FP8 (HT_METHOD_RSVD only): the range finder's products with X run on FP8 MFMA
 * (OCP E4M3, a power-of-two scale per operand tile of every 64-deep chunk, FP32
 * accumulation); everything else runs as in HT_PREC_BF16. After each range finder the
 * a-posteriori estimate ||(I - QQ^T) X_(t)||_F / ||X||_F (four FP32 Gaussian probes) must be
 * at most opt.sketch_tol, or the range finder is redone in BF16; ht_stats reports the
 * nodes checked, the largest estimate and the redone nodes. */
typedef enum { HT_PREC_FP32 = 0, HT_PREC_BF16 = 1, HT_PREC_FP16 = 2, HT_PREC_FP8 = 3 } ht_precision;

typedef enum {
  HT_ERR_MODE_NONE = 0,      /* do not compute the error                                   */
  HT_ERR_MODE_IDENTITY = 1,  /* sqrt(||X||^2 - ||B_root||^2)/||X|| (orthonormal bases)     */
  HT_ERR_MODE_RESIDUAL = 2   /* one fused pass: ||X - X_HT||_F / ||X||_F                   */
} ht_err_mode;

typedef struct {
  int32_t method;          /* ht_method                                                 */
  int32_t precision;       /* ht_precision of the large GEMMs (plain always FP32)        */
  int32_t oversample;      /* p for HT_METHOD_RSVD (default 8)                           */
  int32_t power_iters;     /* q for HT_METHOD_RSVD (default 1)                           */
  int32_t err_mode;        /* ht_err_mode (default HT_ERR_MODE_RESIDUAL)                 */
  int32_t max_eig_iters;   /* cap on subspace iterations of the exact solvers (200)      */
  double eig_tol;          /* relative Ritz-value tolerance of the exact solvers (1e-10) */
  uint64_t seed;           /* seed of the Philox test matrices (1234)                    */
  int32_t want_interior_u; /* also return U_t of interior non-root nodes                 */
  /* NaN/Inf in X make ||X||^2 (formed by the first pass over X) non-finite, and the call
   * returns HT_ERR_NONFINITE. check_finite = 1 (default): a run that reads the device
   * between its steps checks ||X||^2 right after that first pass, before any eigen problem;
   * a replayed leaves-to-root run (it reads the device once, at its end) checks it there,
   * before anything else. 0: only the end of the run checks. */
  int32_t check_finite;
  /* Test matrix of HT_METHOD_RSVD (ht_sketch_kind). With HT_SKETCH_LEARNED the range
   * finder of node t uses Omega_t = Khatri-Rao product of the per-mode sketches of t's
   * column modes (width sketch_width, which replaces k + oversample), and an FP32
   * a-posteriori estimate of the range residual ||(I - QQ^T) X_(t)||_F / ||X||_F from four
   * Gaussian probe vectors decides: above sketch_tol the node is redone with the dense
   * Gaussian test matrix. */
  int32_t sketch;          /* ht_sketch_kind (default HT_SKETCH_GAUSSIAN)                */
  int32_t sketch_width;    /* width w of the learned sketch                              */
  const float* sketch_weights; /* device: per-mode sketches at ht_sketch_layout offsets  */
  double sketch_tol;       /* residual estimate above which the Gaussian is used (1e-3)  */
} ht_options;

typedef enum {
  HT_SKETCH_GAUSSIAN = 0,  /* dense Philox Gaussian test matrix (the reference's choice)  */
  HT_SKETCH_LEARNED = 1    /* learned per-mode Khatri-Rao sketch (see ht_sketch_*)        */
} ht_sketch_kind;

/* Timing and accounting of the last call (filled when stats != NULL). */
typedef struct {
  double total_ms;         /* device time of the decomposition (HIP events)              */
  double h2d_ms;           /* host-to-device copy of X when x_on_device == 0             */
  double error_ms;         /* residual pass when err_mode == HT_ERR_MODE_RESIDUAL        */
  double bytes_x_passes;   /* number of full passes over X made by the decomposition     */
  int32_t eig_iters_max;   /* largest subspace-iteration count of any node               */
  int32_t eig_calls;       /* number of node eigen problems solved                       */
  int32_t sketch_fallbacks;/* nodes redone: Gaussian after a learned sketch, BF16 after FP8 */
  int32_t sketch_nodes;    /* nodes whose range finder was checked (learned sketch or FP8) */
  double sketch_resid;     /* largest range-residual estimate of those checks            */
  double workspace_bytes;  /* device workspace the call used (one contiguous block)      */
  int32_t replayed;        /* decompositions that replayed recorded solver rounds (read
                              the device once, at their end)                          */
} ht_stats;

typedef struct ht_handle_s* ht_handle_t;

/* Fill *opt with the defaults listed above. */
void ht_default_options(ht_options* opt);

/* Create a handle bound to `device`: it owns a HIP stream, a growable device workspace
 * and HIP events. Handles are independent; one handle must not be used concurrently. */
ht_status ht_create(ht_handle_t* handle, int32_t device);
ht_status ht_destroy(ht_handle_t handle);
/* The HIP stream (hipStream_t) the handle launches on, as an opaque pointer. */
void* ht_get_stream(ht_handle_t handle);

/* Validate a tree (per-node lo/hi/left/right arrays, node 0 = root) for `order` modes. */
ht_status ht_validate_tree(int32_t order, int32_t n_nodes, const int32_t* lo, const int32_t* hi,
                           const int32_t* left, const int32_t* right);

/* Packed output layout. For node t: U_t (leaves, and interior non-root nodes when
 * want_interior_u) starts at u_off[t] (else -1), B_t (interior nodes) at b_off[t]
 * (else -1), the node's singular values (k_t doubles stored as floats) at s_off[t]
 * (non-root nodes, else -1). *total receives the number of floats of the whole block. */
ht_status ht_output_layout(int32_t order, const int64_t* dims, int32_t n_nodes, const int32_t* lo,
                           const int32_t* hi, const int32_t* left, const int32_t* right,
                           const int32_t* ranks, int32_t want_interior_u, int64_t* u_off,
                           int64_t* b_off, int64_t* s_off, int64_t* total);

/* General entry point.
 *   x            FP32 tensor, column-major; a device pointer when x_on_device != 0,
 *                otherwise host memory that is copied in (timed separately as h2d_ms).
 *   dims         `order` mode sizes (order 2..HT_MAX_ORDER).
 *   tree, ranks  as above (ranks[0] == 1).
 *   out          packed output block of ht_output_layout's size; device memory when
 *                out_on_device != 0, otherwise host memory.
 *   rel_err      receives the relative error (NaN when err_mode == NONE); may be NULL.
 * The call is synchronous with respect to the host. */
ht_status ht_decompose(ht_handle_t handle, const float* x, int32_t x_on_device, int32_t order,
                       const int64_t* dims, int32_t n_nodes, const int32_t* lo, const int32_t* hi,
                       const int32_t* left, const int32_t* right, const int32_t* ranks,
                       const ht_options* opt, float* out, int32_t out_on_device, double* rel_err,
                       ht_stats* stats);

/* Rank-adaptive decomposition: what the reference's rel_eps and max_rank arguments
 * (float htd(dt*,long,long,long,int*,dt rel_eps,dt max_rank),
 * third_order_tensor_decompsition/CUDA/Baseline/function.h:6) stand for; the reference
 * accepts and ignores them. max_ranks[t] caps node t (max_ranks[0] = 1), max_rank > 0 caps
 * every node, and an interior node keeps at most k_l k_r columns. Once a node's spectrum is
 * known its rank becomes the smallest r within its cap whose truncation removes at most
 * (rel_eps ||X||_F)^2 / m of energy, m the number of truncations the method's error adds up
 * over: 2d - 3 for HT_METHOD_PLAIN (the two root children truncate one spectrum and get one
 * rank), 2d - 2 for HT_METHOD_LTR. Then ||X - X_HT||_F <= rel_eps ||X||_F whenever the caps
 * leave room (the reported rel_err is measured, not the bound). ranks_out receives the
 * ranks chosen; `out` must hold ht_output_layout(max_ranks) floats and receives the packed
 * layout of ranks_out (a shorter block). HT_METHOD_RSVD returns HT_ERR_UNSUPPORTED. */
ht_status ht_decompose_adaptive(ht_handle_t handle, const float* x, int32_t x_on_device, int32_t order,
                                const int64_t* dims, int32_t n_nodes, const int32_t* lo, const int32_t* hi,
                                const int32_t* left, const int32_t* right, const int32_t* max_ranks,
                                double rel_eps, int32_t max_rank, const ht_options* opt, float* out,
                                int32_t out_on_device, int32_t* ranks_out, double* rel_err, ht_stats* stats);

/* Decompose `count` device tensors of one shape and tree (xs[i] -> outs[i], rel_errs[i]),
 * with up to `inflight` decompositions running at once (1..HT_MAX_INFLIGHT), each on its
 * own stream, workspace and recorded solver plan (worker handles that `handle` creates on
 * first use and keeps): the latency-bound small solver kernels of one decomposition run
 * beside the passes over X and solver kernels of the others. Each stream needs a hardware
 * queue of its own for that (HIP's GPU_MAX_HW_QUEUES, 4 by default, counts every stream
 * of the process, torch's included); streams that share a queue run one after the other,
 * with the same results. Outputs equal those of
 * ht_decompose one tensor at a time (to rounding where, with several in flight, the slab
 * passes of the leaves-to-root method leave an eighth of the CUs to the other decompositions
 * and split their partial sums differently; bit for bit otherwise). Worker w starts each
 * decomposition after the first pass over X of worker w-1's of the same round, so the
 * workers' passes and solver phases interleave. The work starts after what the handle's stream holds
 * and the call returns when every decomposition is done; stats (optional) receives the
 * host wall time of the whole call in total_ms. No counterpart in the reference, whose
 * drivers decompose one tensor per call (third_order_tensor_decompsition/CUDA/Optimized/
 * ht_init.cu:9-67 loops over sizes). Returns the first failing status. */
#define HT_MAX_INFLIGHT 8
ht_status ht_decompose_many(ht_handle_t handle, int32_t count, const float* const* xs, int32_t order,
                            const int64_t* dims, int32_t n_nodes, const int32_t* lo, const int32_t* hi,
                            const int32_t* left, const int32_t* right, const int32_t* ranks, const ht_options* opt,
                            float* const* outs, double* rel_errs, int32_t inflight, ht_stats* stats);

/* ht_decompose without any device allocation. No counterpart in the reference (whose
 * drivers cudaMalloc and cudaFree on every call): it is for callers that cannot allocate
 * device memory, such as code under HIP graph capture, tight serving loops and harnesses
 * that forbid hipMalloc. x and out are device pointers (as x_on_device = out_on_device = 1
 * of ht_decompose); the workspace is a fixed library array of HT_NOALLOC_SCRATCH_BYTES
 * per GPU, held until the process ends (its address per device is taken by ht_create).
 * A problem whose workspace does not fit returns HT_ERR_WORKSPACE (ht_stats.
 * workspace_bytes of an ht_decompose call of the same problem gives its need, or more: for
 * randomised runs in 16-bit precision on tensors of 2^22 elements and up, ht_decompose also
 * holds a 16-bit copy of X, which this entry does without; the 4th-order 128^4 rank-16
 * leaves-to-root decomposition fits). Every call on one GPU shares that
 * array: calls that overlap on one GPU (through different handles) race. Results equal
 * ht_decompose's. */
#define HT_NOALLOC_SCRATCH_BYTES ((size_t)160 << 20)
ht_status ht_decompose_noalloc(ht_handle_t handle, const float* x, int32_t order, const int64_t* dims,
                               int32_t n_nodes, const int32_t* lo, const int32_t* hi, const int32_t* left,
                               const int32_t* right, const int32_t* ranks, const ht_options* opt, float* out,
                               double* rel_err, ht_stats* stats);

/* ||X - X_HT||_F / ||X||_F in one fused device pass for an HT block in the packed layout
 * (the reconstruction + error of third_order_tensor_decompsition/CUDA/Baseline/htd_unopt.cu:202-266
 * and fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/htd4_tensorcore.cu:232-291,
 * without materialising X_HT). x and packed are device pointers. */
ht_status ht_residual(ht_handle_t handle, const float* x, int32_t order, const int64_t* dims,
                      int32_t n_nodes, const int32_t* lo, const int32_t* hi, const int32_t* left,
                      const int32_t* right, const int32_t* ranks, int32_t want_interior_u,
                      const float* packed, double* rel_err);

/* Explicit reconstruction X_HT (device pointers). */
ht_status ht_reconstruct(ht_handle_t handle, int32_t order, const int64_t* dims, int32_t n_nodes,
                         const int32_t* lo, const int32_t* hi, const int32_t* left,
                         const int32_t* right, const int32_t* ranks, int32_t want_interior_u,
                         const float* packed, float* x_out);

/* ---- synthetic inputs (seeded Philox4x32-10, generated on the device) ---------------
 * Tucker: core (prod core_dims, i.i.d. N(0,1)) x_n U_n, U_n orthonormal (Q of QR of a
 * Gaussian, diag R > 0), plus Gaussian noise with ||E||_F = noise_rel * ||X0||_F.
 * Replaces genHtensor (third_order_tensor_decompsition/CUDA/Baseline/kernel.cu:94-202,
 * fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:106-248), whose
 * srand(time(NULL)) input is not reproducible. x_dev is a device pointer. */
ht_status ht_generate_tucker(ht_handle_t handle, int32_t order, const int64_t* dims,
                             const int32_t* core_dims, uint64_t seed, double noise_rel,
                             float* x_dev);
/* Diagonal core sigma_k = k^-exponent (every mode spectrum decays as k^-exponent). */
ht_status ht_generate_cp_decay(ht_handle_t handle, int32_t order, const int64_t* dims, int32_t rank,
                               double exponent, uint64_t seed, double noise_rel, float* x_dev);
/* Gaussian-bump family f = sum_j a_j exp(-||x-c_j||^2/(2 w_j^2)) on the [0,1]^order grid. */
ht_status ht_generate_bumps(ht_handle_t handle, int32_t order, int64_t n, int32_t n_bumps,
                            uint64_t seed, int64_t sample, float* x_dev);

/* ---- batched HT ------------------------------------------------------------------
 * Many small, unevenly sized 3rd-order tensors decomposed in one persistent launch
 * (captured in a hipGraph and replayed while the arguments stay the same), with the
 * tree {1},{2,3} (heap nodes: 0 = {1,2,3}, 1 = {1}, 2 = {2,3}, 3 = {2}, 4 = {3}) and
 * the plain semantics. No counterpart in the reference: its third_order_tensor_decompsition/
 * CUDA/Optimized/htd_batch.cu is a single decomposition and batch_tensorcore.cu:46-95 a
 * standalone batched-GEMM test. Limits: mode sizes <= 64, ranks <= 8.
 *   dims       count x 3 mode sizes (host);  ranks  count x 5 heap ranks, k[0] = 1 (host)
 *   x          device; tensor b (column-major) at x + x_offsets[b] (host offsets)
 *   out        device; tensor b's block, in ht_output_layout's order for that tree, at
 *              out + out_offsets[b] (ht_batch_output_layout gives a packed choice)
 *   rel_err    host, count relative errors (may be NULL)
 * The launch shape (fibre-chunk length, blocks per CU) is timed once per (count,
 * largest mode size) on the first call and cached in the handle. */
ht_status ht_batch_output_layout(int32_t count, const int64_t* dims, const int32_t* ranks,
                                 int64_t* out_offsets, int64_t* total);
ht_status ht_decompose_batched(ht_handle_t handle, const float* x, int32_t count, const int64_t* dims,
                               const int32_t* ranks, const int64_t* x_offsets, float* out,
                               const int64_t* out_offsets, double* rel_err, ht_stats* stats);
/* Shapes of the batched workload: mode sizes uniform in [lo, hi] and Tucker ranks
 * uniform in [rlo, rhi] (Philox uniforms of stream 4095, six per tensor; host). */
ht_status ht_batch_shapes(int32_t count, uint64_t seed, int32_t lo, int32_t hi, int32_t rlo, int32_t rhi,
                          int64_t* dims, int32_t* core);
/* Tensor b = Tucker(core[b] i.i.d. N(0,1), orthonormal factors) + noise, Philox streams
 * 4096 + 16 b (core), + 1 + m (factor of mode m), + 15 (noise); written at x_dev + x_offsets[b]. */
ht_status ht_generate_batch(ht_handle_t handle, int32_t count, const int64_t* dims, const int32_t* core,
                            uint64_t seed, double noise_rel, float* x_dev, const int64_t* x_offsets);

/* ---- learned per-mode sketch ------------------------------------------------------
 * No counterpart in the reference, whose range finder draws a dense Gaussian test
 * matrix (fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:419-444).
 * Mode m owns a sketch S_m (width x dims[m], row-major) at offsets[m] of one flat FP32
 * array; the nodes sketched are those HT_METHOD_RSVD runs a range finder on (a non-root
 * interior node whose sibling is not handled through it), and node t's test matrix is
 *   Omega_t[c, j] = prod_{m in cols(t)} S_m[j, i_m(c)],
 * cols(t) being the modes of t's matricisation columns (column-major index c).
 * The loss of one tensor is the mean over the sketched nodes of the energy fraction the
 * sketched range misses, 1 - ||P_Y X_(t)||_F^2 / ||X||_F^2 with Y = X_(t) Omega_t, computed
 * with FP32 MFMA products and an FP64 solve (relative ridge 1e-6 on Y^T Y). */
typedef struct {
  int32_t width;   /* w, 1..64 (default 20)                                   */
  int32_t epochs;  /* passes over the training tensors (default 2)            */
  int32_t batch;   /* tensors per Adam step (default 8)                       */
  float lr;        /* Adam step size (default 0.05)                           */
  float beta1;     /* Adam moments (0.9, 0.999) and epsilon (1e-8)            */
  float beta2;
  float eps;
  float min_loss;  /* a step whose mean loss is below this is skipped: there the
                      gradient is rounding noise of the FP32 products (1e-5)   */
} ht_sketch_train_options;
void ht_sketch_default_train_options(ht_sketch_train_options* opt);
/* offsets[order] and *total (floats) of the flat sketch array. */
ht_status ht_sketch_layout(int32_t order, const int64_t* dims, int32_t width, int64_t* offsets, int64_t* total);
/* i.i.d. N(0,1) sketches (Philox stream 200 of `seed`), the untrained starting point. */
ht_status ht_sketch_init_gaussian(ht_handle_t handle, int32_t order, const int64_t* dims, int32_t width,
                                  uint64_t seed, float* sketch_dev);
/* Loss of one device tensor and, when grad_dev != NULL, its gradient (written, same
 * layout as the sketch; modes outside every sketched node get zeros). loss: host. */
ht_status ht_sketch_loss_grad(ht_handle_t handle, const float* x_dev, int32_t order, const int64_t* dims,
                              int32_t n_nodes, const int32_t* lo, const int32_t* hi, const int32_t* left,
                              const int32_t* right, int32_t width, const float* sketch_dev, float* grad_dev,
                              double* loss);
/* Train on `count` Gaussian-bump tensors (ht_generate_bumps samples first_sample.. of
 * `seed`, each generated on the device when it is used) with Adam; sketch_dev holds the
 * starting point and receives the result; epoch_loss (host, opt->epochs) the mean loss
 * of each epoch, taken before each step's update. */
ht_status ht_sketch_train_bumps(ht_handle_t handle, int32_t order, int64_t n, int32_t n_bumps, uint64_t seed,
                                int64_t first_sample, int64_t count, int32_t n_nodes, const int32_t* lo,
                                const int32_t* hi, const int32_t* left, const int32_t* right,
                                const ht_sketch_train_options* opt, float* sketch_dev, double* epoch_loss);
/* Per-tensor loss (host, count) of a sketch on held-out bump tensors. */
ht_status ht_sketch_eval_bumps(ht_handle_t handle, int32_t order, int64_t n, int32_t n_bumps, uint64_t seed,
                               int64_t first_sample, int64_t count, int32_t n_nodes, const int32_t* lo,
                               const int32_t* hi, const int32_t* left, const int32_t* right, int32_t width,
                               const float* sketch_dev, double* loss);

/* ---- compatibility wrappers keeping the reference names --------------------------- */

/* third_order_tensor_decompsition/CUDA/Baseline/function.h:6: plain FP32 HT of the
 * reference's tree {1,2},{3}; x is host memory (a*b*c floats), k[5] heap-indexed ranks
 * with k[0]=1. rel_eps <= 0 (what the reference's driver effectively gets: it ignores
 * both): the ranks k. rel_eps > 0: k[1..4] are caps, max_rank > 0 a further cap, and the
 * ranks are chosen as ht_decompose_adaptive chooses them, then written back into k.
 * Returns the relative error, or a negative value on failure. */
float htd(float* x, long a, long b, long c, int* k, float rel_eps, float max_rank);
/* third_order_tensor_decompsition/CUDA/Optimized/function.h:6: the mixed-precision
 * variant (LTR, BF16 MFMA Grams) of the same tree. */
float htd_tc(float* x, long a, long b, long c, int* k);
/* fourth_order_tensor_decomposition/single_GPU/CUDA/Baseline/function.h:6: plain FP32
 * HT of the balanced tree {1,2},{3,4}; k[7] heap-indexed. The error is printed as in the
 * reference ("error rate ..."). */
void htd4(float* x, long a, long b, long c, long d, int* k);
/* fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/function.h:6: the
 * mixed-precision variant; the two BLAS/solver handles are accepted and ignored. */
void htd4_tc(float* x, long a, long b, long c, long d, int* k, void* blas_handle,
             void* solver_handle);

/* Library version string. */
const char* ht_version(void);

#ifdef __cplusplus
}

/* C++ overloads with the exact signatures the reference's optimized drivers declare and
 * call, so that their call lines compile unchanged against this header:
 *   error = htd(X,a,b,c,k);              third_order_tensor_decompsition/CUDA/Optimized/ht_init.cu:47
 *   htd(X,a,b,c,k);                      third_order_tensor_decompsition/CUDA/TSQR/ht_init.cu:35
 *   htd4(X,a,b,c,d,k,handle,cusolverH);  fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/ht_init.cu:53
 * (the C names above are unchanged: these are inline and add no exported symbol). */

/* float htd(dt*,long,long,long,int*): third_order_tensor_decompsition/CUDA/Optimized/function.h:6
 * (and the void variant of third_order_tensor_decompsition/CUDA/TSQR/function.h:6, whose
 * callers discard the result): the mixed-precision decomposition of htd_tc. */
inline float htd(float* x, long a, long b, long c, int* k) { return htd_tc(x, a, b, c, k); }

/* void htd4(dt*,long,long,long,long,int*,cublasHandle_t,cusolverDnHandle_t):
 * fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/function.h:6. The two
 * library handles are opaque pointers here (no cuBLAS/cuSOLVER on this platform) and are
 * ignored, as htd4_tc does. */
inline void htd4(float* x, long a, long b, long c, long d, int* k, void* blas_handle, void* solver_handle) {
  htd4_tc(x, a, b, c, d, k, blas_handle, solver_handle);
}
#endif

#endif /* HT_AMD_H */
