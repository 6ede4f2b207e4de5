"""MI355X-native hierarchical Tucker (HT) decomposition engine (gfx950 HIP kernels behind a C ABI).

    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    res = ht.decompose(x, dims, ht.tree.balanced_4d(), ranks, method="ltr")
"""
from . import sketch, tree
from ._lib import HTError, LIB_PATH, load as load_library
from .engine import (Batch, Handle, HTensor, batch_ranks, batch_shapes, decompose, decompose_adaptive, decompose_batched,
                     decompose_many,
                     default_handle, generate_batch, generate_bumps, generate_cp_decay, generate_tucker, htd, htd4,
                     htd4_tc, htd_tc, make_options, output_layout, reconstruct, residual)

__all__ = [
    "sketch", "tree", "HTError", "LIB_PATH", "load_library", "Handle", "HTensor", "decompose", "default_handle",
    "generate_bumps", "generate_cp_decay", "generate_tucker", "htd", "htd4", "htd4_tc", "htd_tc",
    "make_options", "output_layout", "reconstruct", "residual", "Batch", "batch_ranks", "batch_shapes",
    "decompose_adaptive", "decompose_batched", "decompose_many", "generate_batch",
]
