"""ctypes binding of libttgpu.so (the C ABI declared in include/ttgpu.h).

The shared library is the product; this module only loads it and converts
status codes into exceptions.  There is no CPU fallback: if the library or a
gfx950 device is missing, every call raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libttgpu.so"

TTG_OK = 0
TTG_BAD_ARG = 1
TTG_NEED_SPACE = 2
TTG_NONFINITE = 3
TTG_NO_CONVERGENCE = 4
TTG_HIP_ERROR = 5
TTG_NO_DEVICE = 6
STATUS_NAMES = {
    0: "TTG_OK", 1: "TTG_BAD_ARG", 2: "TTG_NEED_SPACE", 3: "TTG_NONFINITE",
    4: "TTG_NO_CONVERGENCE", 5: "TTG_HIP_ERROR", 6: "TTG_NO_DEVICE", 7: "TTG_IO_ERROR",
}
PREC_FP64 = 0
PREC_FP32 = 1

# every symbol include/ttgpu.h declares
EXPORTS = [
    "ttg_ctx_create", "ttg_ctx_destroy", "ttg_last_error", "ttg_ctx_stream", "ttg_version",
    "ttg_round_capacity", "ttg_round", "ttg_round_dev", "ttg_round_matrix",
    "ttg_mtimes", "ttg_mtimes_dev", "ttg_mtimes_round", "ttg_mtimes_round_dev",
    "ttg_round_batched_dev",
    "ttg_adjoint_completion_f64", "ttg_adjoint_completion_f32",
    "ttg_adjoint_bilinear_f64", "ttg_adjoint_bilinear_f32", "ttg_adjoint_bilinear_f64_noalloc",
    "ttg_plus_dev", "ttg_plus_round_dev", "ttg_norm_dev", "ttg_plus_norm_dev", "ttg_dot_dev",
    "ttg_round_gram_prepare", "ttg_round_gram_launch", "ttg_round_gram_graph",
    "ttg_full_to_tt_dev", "ttg_sdv_info", "ttg_sdv_read", "ttg_sdv_read_dev", "ttg_sdv_write", "ttg_sdv_write_dev",
    "ttg_profile_begin", "ttg_profile_end",
]
# kernel-level test hooks (include/ttgpu_testing.h)
TEST_EXPORTS = ["ttg_test_gemm", "ttg_test_householder", "ttg_test_chol", "ttg_test_trsm", "ttg_test_jacobi", "ttg_test_jacobi_timing",
                "ttg_test_chol_timing", "ttg_test_trsm_timing", "ttg_test_chol_trsm"]
KERNEL_CLASSES = {"gemm": 1, "jacobi": 2, "cholesky": 3, "trtri": 4, "householder": 5, "mtimes": 6}


class TTGError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


class TTGInfo(C.Structure):
    _fields_ = [
        ("qr_fallbacks", C.c_int64),
        ("jacobi_sweeps", C.c_int64),
        ("jacobi_maxsweeps", C.c_int64),
        ("seconds", C.c_double),
        ("qr_retries", C.c_int64),
        ("svd_skipped", C.c_int64),
        ("path", C.c_int64),
        ("err_est", C.c_double),
        ("orth_est", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None

P = C.c_void_p
I64 = C.c_int64
I64P = C.POINTER(C.c_int64)
DP = C.POINTER(C.c_double)


def load(path: Path | str | None = None):
    """Load libttgpu.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(str(p))
    st = C.c_int
    sig = {
        "ttg_ctx_create": (st, [C.c_int, P, C.POINTER(P)]),
        "ttg_ctx_destroy": (st, [P]),
        "ttg_last_error": (C.c_char_p, [P]),
        "ttg_ctx_stream": (P, [P]),
        "ttg_version": (C.c_char_p, []),
        "ttg_round_capacity": (st, [I64, I64P, I64P, I64P]),
        "ttg_round": (st, [P, I64, I64P, I64P, P, C.c_double, I64P, C.c_int, I64P, P, I64, I64P, P]),
        "ttg_round_dev": (st, [P, I64, I64P, I64P, P, C.c_double, I64P, C.c_int, I64P, P, I64, I64P, P]),
        "ttg_round_matrix": (st, [P, I64, I64P, I64P, I64P, P, C.c_double, I64P, C.c_int, I64P, P, I64, I64P, P]),
        "ttg_mtimes": (st, [P, I64, I64P, I64P, I64P, P, I64P, P, I64P, P, I64, I64P]),
        "ttg_mtimes_dev": (st, [P, I64, I64P, I64P, I64P, P, I64P, P, I64P, P, I64, I64P]),
        "ttg_mtimes_round": (st, [P, I64, I64P, I64P, I64P, P, I64P, P, C.c_double, I64P, C.c_int, I64P, P, I64, I64P, P]),
        "ttg_mtimes_round_dev": (st, [P, I64, I64P, I64P, I64P, P, I64P, P, C.c_double, I64P, C.c_int, I64P, P, I64, I64P, P]),
        "ttg_round_batched_dev": (st, [P, I64, I64P, I64P, I64P, I64P, I64P, I64P, P, C.c_double, I64P, C.c_int,
                                       I64P, I64P, P, P]),
        "ttg_adjoint_completion_f64": (st, [P, I64, I64P, I64P, P, I64, P, P, DP, P]),
        "ttg_adjoint_completion_f32": (st, [P, I64, I64P, I64P, P, I64, P, P, DP, P]),
        "ttg_adjoint_bilinear_f64": (st, [P, I64, I64P, I64P, I64P, P, I64P, P, I64P, P, DP, P, P]),
        "ttg_adjoint_bilinear_f32": (st, [P, I64, I64P, I64P, I64P, P, I64P, P, I64P, P, DP, P, P]),
        "ttg_adjoint_bilinear_f64_noalloc": (st, [P, I64, I64P, I64P, I64P, P, I64P, P, I64P, P, DP, P, P]),
        "ttg_plus_dev": (st, [P, I64, I64P, I64P, P, C.c_double, I64P, P, C.c_double, I64P, P, I64, I64P]),
        "ttg_plus_round_dev": (st, [P, I64, I64P, I64P, P, C.c_double, I64P, P, C.c_double, C.c_double, I64P,
                                    C.c_int, I64P, P, I64, I64P, P]),
        "ttg_norm_dev": (st, [P, I64, I64P, I64P, P, DP]),
        "ttg_plus_norm_dev": (st, [P, I64, I64P, I64P, P, C.c_double, I64P, P, C.c_double, DP]),
        "ttg_dot_dev": (st, [P, I64, I64P, I64P, P, I64P, P, DP]),
        "ttg_full_to_tt_dev": (st, [P, I64, I64P, P, C.c_double, I64P, I64P, P, I64, I64P, P]),
        "ttg_round_gram_prepare": (st, [P, I64, I64P, I64P, P, C.c_double, C.c_int, I64P, P, I64, I64P,
                                        C.POINTER(C.c_int)]),
        "ttg_round_gram_launch": (st, [P, P]),
        "ttg_round_gram_graph": (st, [P, C.POINTER(P)]),
        "ttg_sdv_info": (st, [C.c_char_p, I64P, I64P, C.POINTER(C.c_int)]),
        "ttg_sdv_read": (st, [C.c_char_p, I64, I64P, I64P, I64P, P, I64, I64P]),
        "ttg_sdv_read_dev": (st, [P, C.c_char_p, I64, I64P, I64P, I64P, P, I64, I64P]),
        "ttg_sdv_write": (st, [C.c_char_p, I64, I64P, I64P, P, C.c_int]),
        "ttg_sdv_write_dev": (st, [P, C.c_char_p, I64, I64P, I64P, P, C.c_int]),
        "ttg_profile_begin": (st, [P, C.c_int]),
        "ttg_profile_end": (st, [P, I64P, DP, DP, DP]),
        "ttg_test_gemm": (st, [P, C.c_int, C.c_int, C.c_int, I64, I64, I64, P, I64, P, I64, P, I64]),
        "ttg_test_householder": (st, [P, I64, I64, P, P, P]),
        "ttg_test_chol": (st, [P, I64, P, C.c_double, P, P, P]),
        "ttg_test_jacobi": (st, [P, I64, P, C.c_int, P, I64, I64P, P, P]),
        "ttg_test_trsm": (st, [P, C.c_int, I64, I64, P, P, P]),
        "ttg_test_jacobi_timing": (st, [P, I64, P, P, I64]),
        "ttg_test_chol_timing": (st, [P, I64, P, P, I64]),
        "ttg_test_trsm_timing": (st, [P, I64, I64, P, P, P, I64]),
        "ttg_test_chol_trsm": (st, [P, I64, I64, P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def i64(seq):
    """int64 ctypes array from a sequence (None passes through)."""
    if seq is None:
        return None
    vals = [int(v) for v in seq]
    return (C.c_int64 * max(len(vals), 1))(*vals)


class Context:
    """One HIP stream + device workspace (``ttg_ctx``).  One per host thread."""

    def __init__(self, device: int = 0, stream: int | None = None):
        lib = load()
        h = P()
        s = lib.ttg_ctx_create(int(device), P(stream) if stream else None, C.byref(h))
        if s != TTG_OK:
            raise TTGError(s, "ttg_ctx_create failed (no gfx950 device?)")
        self.h = h
        self.lib = lib
        self.device = device
        _live.append(weakref.ref(self))

    def check(self, status, allow=(TTG_OK,)):
        if status in allow:
            return status
        msg = self.lib.ttg_last_error(self.h)
        raise TTGError(status, msg.decode() if msg else "")

    @property
    def stream(self) -> int:
        return int(self.lib.ttg_ctx_stream(self.h) or 0)

    def profile_begin(self, kernel_class):
        cls = KERNEL_CLASSES[kernel_class] if isinstance(kernel_class, str) else int(kernel_class)
        self.check(self.lib.ttg_profile_begin(self.h, cls))

    def profile_end(self):
        n = C.c_int64()
        ms, fl, by = C.c_double(), C.c_double(), C.c_double()
        self.check(self.lib.ttg_profile_end(self.h, C.byref(n), C.byref(ms), C.byref(fl), C.byref(by)))
        return {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}

    def close(self):
        if getattr(self, "h", None):
            self.lib.ttg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}
_live = []


def _close_all():
    """At interpreter exit, destroy every live context, newest first, before
    the HIP runtime tears itself down: each destroy waits for the context's
    streams and releases its worker contexts, captured graph, streams,
    events, pinned host memory and device workspace (ttg_ctx_destroy)."""
    for ref in reversed(list(_live)):
        ctx = ref()
        if ctx is not None:
            ctx.close()
    _live.clear()
    _default_ctx.clear()


atexit.register(_close_all)


def default_context(device: int = 0) -> Context:
    ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = Context(device)
        _default_ctx[device] = ctx
    return ctx
