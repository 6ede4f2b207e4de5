"""ctypes binding of libht_amd.so (the C ABI declared in include/ht_amd.h and
include/ht_amd_debug.h). Loading fails loudly when the library is missing: there is no
CPU fallback behind these functions."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libht_amd.so")

STATUS_NAMES = {
    0: "HT_OK", 1: "HT_ERR_INVALID_ARG", 2: "HT_ERR_INVALID_TREE", 3: "HT_ERR_RANK", 4: "HT_ERR_ALLOC",
    5: "HT_ERR_HIP", 6: "HT_ERR_NONFINITE", 7: "HT_ERR_NOCONV", 8: "HT_ERR_UNSUPPORTED", 9: "HT_ERR_NO_DEVICE",
    10: "HT_ERR_WORKSPACE",
}

METHOD = {"plain": 0, "ltr": 1, "rsvd": 2}
PRECISION = {"fp32": 0, "bf16": 1, "fp16": 2, "fp8": 3}
ERR_MODE = {"none": 0, "identity": 1, "residual": 2}
SKETCH = {"gaussian": 0, "learned": 1}


class HTError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = int(status)
        super().__init__(f"{what}: {STATUS_NAMES.get(self.status, self.status)}")


class Options(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32), ("precision", ctypes.c_int32), ("oversample", ctypes.c_int32),
        ("power_iters", ctypes.c_int32), ("err_mode", ctypes.c_int32), ("max_eig_iters", ctypes.c_int32),
        ("eig_tol", ctypes.c_double), ("seed", ctypes.c_uint64), ("want_interior_u", ctypes.c_int32),
        ("check_finite", ctypes.c_int32), ("sketch", ctypes.c_int32), ("sketch_width", ctypes.c_int32),
        ("sketch_weights", ctypes.c_void_p), ("sketch_tol", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("total_ms", ctypes.c_double), ("h2d_ms", ctypes.c_double), ("error_ms", ctypes.c_double),
        ("bytes_x_passes", ctypes.c_double), ("eig_iters_max", ctypes.c_int32), ("eig_calls", ctypes.c_int32),
        ("sketch_fallbacks", ctypes.c_int32), ("sketch_nodes", ctypes.c_int32), ("sketch_resid", ctypes.c_double),
        ("workspace_bytes", ctypes.c_double), ("replayed", ctypes.c_int32),
    ]


class SketchTrainOptions(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("epochs", ctypes.c_int32), ("batch", ctypes.c_int32), ("lr", ctypes.c_float),
        ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("min_loss", ctypes.c_float),
    ]


P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
D = ctypes.c_double
PI32 = ctypes.POINTER(ctypes.c_int32)
PI64 = ctypes.POINTER(ctypes.c_int64)
PD = ctypes.POINTER(ctypes.c_double)
H = ctypes.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "ht_version": (ctypes.c_char_p, []),
    "ht_default_options": (None, [ctypes.POINTER(Options)]),
    "ht_create": (I32, [ctypes.POINTER(H), I32]),
    "ht_destroy": (I32, [H]),
    "ht_get_stream": (P, [H]),
    "ht_validate_tree": (I32, [I32, I32, PI32, PI32, PI32, PI32]),
    "ht_output_layout": (I32, [I32, PI64, I32, PI32, PI32, PI32, PI32, PI32, I32, PI64, PI64, PI64, PI64]),
    "ht_decompose": (I32, [H, P, I32, I32, PI64, I32, PI32, PI32, PI32, PI32, PI32, ctypes.POINTER(Options), P, I32,
                           PD, ctypes.POINTER(Stats)]),
    "ht_decompose_adaptive": (I32, [H, P, I32, I32, PI64, I32, PI32, PI32, PI32, PI32, PI32, D, I32,
                                    ctypes.POINTER(Options), P, I32, PI32, PD, ctypes.POINTER(Stats)]),
    "ht_decompose_many": (I32, [H, I32, P, I32, PI64, I32, PI32, PI32, PI32, PI32, PI32, ctypes.POINTER(Options), P,
                                PD, I32, ctypes.POINTER(Stats)]),
    "ht_decompose_noalloc": (I32, [H, P, I32, PI64, I32, PI32, PI32, PI32, PI32, PI32, ctypes.POINTER(Options), P,
                                   PD, ctypes.POINTER(Stats)]),
    "ht_residual": (I32, [H, P, I32, PI64, I32, PI32, PI32, PI32, PI32, PI32, I32, P, PD]),
    "ht_reconstruct": (I32, [H, I32, PI64, I32, PI32, PI32, PI32, PI32, PI32, I32, P, P]),
    "ht_generate_tucker": (I32, [H, I32, PI64, PI32, U64, D, P]),
    "ht_batch_output_layout": (I32, [I32, PI64, PI32, PI64, PI64]),
    "ht_decompose_batched": (I32, [H, P, I32, PI64, PI32, PI64, P, PI64, PD, ctypes.POINTER(Stats)]),
    "ht_batch_shapes": (I32, [I32, U64, I32, I32, I32, I32, PI64, PI32]),
    "ht_generate_batch": (I32, [H, I32, PI64, PI32, U64, D, P, PI64]),
    "ht_generate_cp_decay": (I32, [H, I32, PI64, I32, D, U64, D, P]),
    "ht_generate_bumps": (I32, [H, I32, I64, I32, U64, I64, P]),
    "htd": (ctypes.c_float, [P, ctypes.c_long, ctypes.c_long, ctypes.c_long, PI32, ctypes.c_float, ctypes.c_float]),
    "htd_tc": (ctypes.c_float, [P, ctypes.c_long, ctypes.c_long, ctypes.c_long, PI32]),
    "htd4": (None, [P, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_long, PI32]),
    "htd4_tc": (None, [P, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_long, PI32, P, P]),
    "ht_debug_mfma_probe": (I32, [H, P, P, P, P, P, P, P]),
    "ht_mode_gram": (I32, [H, P, I32, PI64, I32, I32, I32, P, PD]),
    "ht_mode_gram16": (I32, [H, P, I32, PI64, I32, I32, I32, P, PD]),
    "ht_slab_gram": (I32, [H, P, I32, I32, I64, I32, P, P, PD]),
    "ht_slab_project": (I32, [H, P, I32, I32, I64, P, I32, P, I32, P]),
    "ht_sym_eig": (I32, [H, P, I32, P, P]),
    "ht_sym_topk": (I32, [H, P, I32, I32, D, P, P, PI32]),
    "ht_tsqr": (I32, [H, P, I64, I32, P, P]),
    "ht_tall_orth": (I32, [H, P, I64, I32, P, P]),
    "ht_skinny_gemm": (I32, [H, P, I64, I64, I32, P, I32, I32, P]),
    "ht_skinny_gemm16": (I32, [H, P, I64, I64, I32, P, I32, I32, P]),
    "ht_philox_normals": (I32, [H, U64, U32, I64, P]),
    "ht_philox_normals_host": (None, [U64, U32, I64, PD]),
    "ht_profile_enable": (I32, [H, I32]),
    "ht_profile_select": (I32, [H, I32]),
    "ht_profile_read": (I32, [H, PD, PI64, I32, I32]),
    "ht_debug_copy": (I32, [H, P, P, I64]),
    "ht_debug_qr_fallbacks": (I32, [H, PI64]),
    "ht_sketch_default_train_options": (None, [ctypes.POINTER(SketchTrainOptions)]),
    "ht_sketch_layout": (I32, [I32, PI64, I32, PI64, PI64]),
    "ht_sketch_init_gaussian": (I32, [H, I32, PI64, I32, U64, P]),
    "ht_sketch_loss_grad": (I32, [H, P, I32, PI64, I32, PI32, PI32, PI32, PI32, I32, P, P, PD]),
    "ht_sketch_train_bumps": (I32, [H, I32, I64, I32, U64, I64, I64, I32, PI32, PI32, PI32, PI32,
                                    ctypes.POINTER(SketchTrainOptions), P, PD]),
    "ht_sketch_eval_bumps": (I32, [H, I32, I64, I32, U64, I64, I64, I32, PI32, PI32, PI32, PI32, I32, P, PD]),
}

PROFILE_KERNELS = ["slab_gram", "slab_proj", "mode_gram", "residual", "skinny_gemm", "tsqr", "small_eig",
                   "small_contract", "generator"]

_lib = None


def load():
    """Load libht_amd.so (raises ImportError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64 and opens it by path,
    # so if the library (linked against the system runtime) were loaded first, torch
    # would start a second runtime that cannot see the device. Importing torch first
    # makes the library bind to the runtime torch already loaded (same soname).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not found: build it with "
                          "`python -m ht_tensor_learning_using_gpu_tensor_cores_amd.csrc.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status, what):
    if status != 0:
        raise HTError(status, what)


def i32_array(vals):
    return (ctypes.c_int32 * len(vals))(*[int(v) for v in vals])


def i64_array(vals):
    return (ctypes.c_int64 * len(vals))(*[int(v) for v in vals])
