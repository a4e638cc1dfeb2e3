"""ttgpu: MI355X (gfx950) TT rounding and TT-matrix x TT-tensor contraction.

Drop-in for TT-Toolbox's ``round(X,eps,rmax)`` and ``A*x`` on the
``tt_tensor``/``tt_matrix`` core/ps/r/n layout.  The compute path is
libttgpu.so (hand-written HIP kernels behind a C ABI, include/ttgpu.h).
"""
from ._lib import (LIB_PATH, EXPORTS, Context, TTGError, TTGInfo, default_context, load,
                   PREC_FP32, PREC_FP64, STATUS_NAMES)
from .tt import (TTTensor, TTMatrix, capacity, dot, dot_dev, full_to_tt_dev, tt_tensor_from_full, mtimes, mtimes_dev, mtimes_round, mtimes_round_dev,
                 norm, norm_dev, plus, plus_dev, plus_norm_dev, plus_round, plus_round_dev, rel_diff, round,
                 round_batched_dev, round_dev, round_gram_graph, round_gram_launch, round_gram_prepare,
                 round_matrix, sdv_info, torch_context, tt_read, tt_read_dev,
                 tt_save, tt_save_dev)

__all__ = [
    "LIB_PATH", "EXPORTS", "Context", "TTGError", "TTGInfo", "default_context", "load",
    "PREC_FP32", "PREC_FP64", "STATUS_NAMES", "TTTensor", "TTMatrix", "capacity", "mtimes",
    "mtimes_dev", "mtimes_round", "mtimes_round_dev", "round", "round_batched_dev", "round_dev",
    "round_matrix", "torch_context", "dot", "dot_dev", "norm", "norm_dev", "plus", "plus_dev",
    "plus_norm_dev", "plus_round", "plus_round_dev", "rel_diff", "sdv_info", "tt_read", "tt_read_dev",
    "tt_save", "tt_save_dev", "full_to_tt_dev", "tt_tensor_from_full", "round_gram_prepare",
    "round_gram_launch", "round_gram_graph",
]
