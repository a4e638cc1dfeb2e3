"""Host-side mirror of TT-Toolbox's ``tt_tensor`` / ``tt_matrix`` for the GPU path.

Same fields and layout as the MATLAB classes (``@tt_tensor/tt_tensor.m:33-41``,
``@tt_matrix/tt_matrix.m:33-39``): ``d``, ``n``, ``r``, one flat column-major
``core`` vector and ``ps`` (0-based here).  ``round`` and ``mtimes`` keep the
reference argument meaning: ``round(tt, eps, rmax)`` (``@tt_tensor/round.m:1``,
``@tt_matrix/round.m:23-27``) and ``A*x`` (``@tt_matrix/mtimes.m:38-65``).
All arithmetic runs in libttgpu.so on the GPU; nothing here computes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import PREC_FP32, PREC_FP64, TTG_NO_CONVERGENCE, TTG_OK, TTGInfo, default_context, i64

_PREC = {"fp64": PREC_FP64, "plain": PREC_FP64, "fp32": PREC_FP32, "mixed": PREC_FP32}


def _prec(p):
    if isinstance(p, int):
        return p
    try:
        return _PREC[p]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(_PREC)}") from None


def _ps(n, r):
    n = np.asarray(n, dtype=np.int64)
    r = np.asarray(r, dtype=np.int64)
    return np.concatenate([[0], np.cumsum(n * r[:-1] * r[1:])]).astype(np.int64)


@dataclass
class TTTensor:
    """``tt_tensor``: mode sizes ``n[d]``, ranks ``r[d+1]``, flat FP64 ``core``."""
    n: np.ndarray
    r: np.ndarray
    core: np.ndarray

    def __post_init__(self):
        self.n = np.asarray(self.n, dtype=np.int64)
        self.r = np.asarray(self.r, dtype=np.int64)
        self.core = np.ascontiguousarray(self.core, dtype=np.float64).ravel()
        if self.r.size != self.n.size + 1 or self.core.size != int(self.ps[-1]):
            raise ValueError("inconsistent n, r and core length")

    @property
    def d(self):
        return int(self.n.size)

    @property
    def ps(self):
        return _ps(self.n, self.r)

    def core_k(self, k):
        ps = self.ps
        return self.core[ps[k]:ps[k + 1]].reshape((self.r[k], self.n[k], self.r[k + 1]), order="F")

    def round(self, eps, rmax=None, precision="fp64", ctx=None, info=None):
        return round(self, eps, rmax, precision, ctx, info)


@dataclass
class TTMatrix:
    """``tt_matrix``: a TT tensor with mode sizes ``n.*m``; core k is ``r x n x m x r'``."""
    n: np.ndarray
    m: np.ndarray
    r: np.ndarray
    core: np.ndarray

    def __post_init__(self):
        self.n = np.asarray(self.n, dtype=np.int64)
        self.m = np.asarray(self.m, dtype=np.int64)
        self.r = np.asarray(self.r, dtype=np.int64)
        self.core = np.ascontiguousarray(self.core, dtype=np.float64).ravel()
        if self.core.size != int(_ps(self.n * self.m, self.r)[-1]):
            raise ValueError("inconsistent n, m, r and core length")

    @property
    def d(self):
        return int(self.n.size)

    @property
    def tt(self):
        return TTTensor(self.n * self.m, self.r, self.core)

    def __matmul__(self, x):
        return mtimes(self, x)

    def round(self, eps, rmax=None, precision="fp64", ctx=None, info=None):
        return round_matrix(self, eps, rmax, precision, ctx, info)


def _rmax_arg(rmax, d):
    if rmax is None:
        return None
    v = np.asarray(rmax, dtype=np.int64).ravel()
    if v.size == 1:
        v = np.full(d + 1, int(v[0]), dtype=np.int64)
    if v.size != d + 1:
        raise ValueError("rmax must be a scalar or have d+1 entries")
    return i64(v)


def capacity(n, r):
    lib = _lib.load()
    cap = C.c_int64()
    d = len(n)
    s = lib.ttg_round_capacity(d, i64(n), i64(r), C.byref(cap))
    if s != TTG_OK:
        raise _lib.TTGError(s, "ttg_round_capacity")
    return int(cap.value)


def round(tt: TTTensor, eps: float, rmax=None, precision="fp64", ctx=None, info=None) -> TTTensor:
    """``Y = round(X, eps, rmax)`` on the GPU (host arrays in, host arrays out)."""
    ctx = ctx or default_context()
    d = tt.d
    cap = capacity(tt.n, tt.r)
    out = np.empty(max(cap, 1), dtype=np.float64)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    inf = TTGInfo()
    s = ctx.lib.ttg_round(ctx.h, d, i64(tt.n), i64(tt.r), tt.core.ctypes.data_as(C.c_void_p),
                          float(eps), _rmax_arg(rmax, d), _prec(precision), r_out,
                          out.ctypes.data_as(C.c_void_p), cap, C.byref(used), C.byref(inf))
    ctx.check(s, allow=(TTG_OK, TTG_NO_CONVERGENCE))
    if info is not None:
        info.update(inf.as_dict())
        info["status"] = s
    return TTTensor(tt.n.copy(), np.array(r_out[:], dtype=np.int64), out[:used.value].copy())


def round_matrix(a: TTMatrix, eps, rmax=None, precision="fp64", ctx=None, info=None) -> TTMatrix:
    """``round`` on a ``tt_matrix`` (``@tt_matrix/round.m:23-27``)."""
    y = round(a.tt, eps, rmax, precision, ctx, info)
    return TTMatrix(a.n.copy(), a.m.copy(), y.r, y.core)


def mtimes(a: TTMatrix, x: TTTensor, ctx=None) -> TTTensor:
    """``y = A*x`` (``@tt_matrix/mtimes.m:38-65``), no rounding."""
    ctx = ctx or default_context()
    d = a.d
    if x.d != d or not np.array_equal(a.m, x.n):
        raise ValueError("inconsistent sizes")
    ry = a.r * x.r
    cap = int(_ps(a.n, ry)[-1])
    out = np.empty(max(cap, 1), dtype=np.float64)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    s = ctx.lib.ttg_mtimes(ctx.h, d, i64(a.n), i64(a.m), i64(a.r), a.core.ctypes.data_as(C.c_void_p),
                           i64(x.r), x.core.ctypes.data_as(C.c_void_p), r_out,
                           out.ctypes.data_as(C.c_void_p), cap, C.byref(used))
    ctx.check(s)
    return TTTensor(a.n.copy(), np.array(r_out[:], dtype=np.int64), out[:used.value].copy())


def mtimes_round(a: TTMatrix, x: TTTensor, eps, rmax=None, precision="fp64", ctx=None, info=None) -> TTTensor:
    """``round(A*x, eps, rmax)`` with the product cores generated inside the sweep."""
    ctx = ctx or default_context()
    d = a.d
    if x.d != d or not np.array_equal(a.m, x.n):
        raise ValueError("inconsistent sizes")
    cap = capacity(a.n, a.r * x.r)
    out = np.empty(max(cap, 1), dtype=np.float64)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    inf = TTGInfo()
    s = ctx.lib.ttg_mtimes_round(ctx.h, d, i64(a.n), i64(a.m), i64(a.r),
                                 a.core.ctypes.data_as(C.c_void_p), i64(x.r),
                                 x.core.ctypes.data_as(C.c_void_p), float(eps), _rmax_arg(rmax, d),
                                 _prec(precision), r_out, out.ctypes.data_as(C.c_void_p), cap,
                                 C.byref(used), C.byref(inf))
    ctx.check(s, allow=(TTG_OK, TTG_NO_CONVERGENCE))
    if info is not None:
        info.update(inf.as_dict())
        info["status"] = s
    return TTTensor(a.n.copy(), np.array(r_out[:], dtype=np.int64), out[:used.value].copy())


# ---------------------------------------------------------------------------
# device-resident forms (torch tensors on the GPU; torch is plumbing only)
# ---------------------------------------------------------------------------
def torch_context(device=0):
    """Context bound to torch's current stream on ``device``."""
    import torch
    s = torch.cuda.current_stream(device).cuda_stream
    return _lib.Context(device, stream=s)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def round_dev(ctx, n, r, dcore, eps, rmax=None, precision="fp64", out=None, info=None):
    """Round a TT whose flat core is a CUDA float64 tensor; returns (r_out, core_out view)."""
    import torch
    d = len(n)
    cap = capacity(n, r)
    if out is None:
        out = torch.empty(max(cap, 1), dtype=torch.float64, device=dcore.device)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    inf = TTGInfo()
    s = ctx.lib.ttg_round_dev(ctx.h, d, i64(n), i64(r), _ptr(dcore), float(eps), _rmax_arg(rmax, d),
                              _prec(precision), r_out, _ptr(out), int(out.numel()), C.byref(used),
                              C.byref(inf))
    ctx.check(s, allow=(TTG_OK, TTG_NO_CONVERGENCE))
    if info is not None:
        info.update(inf.as_dict())
        info["status"] = s
    return np.array(r_out[:], dtype=np.int64), out[:used.value]


def mtimes_dev(ctx, n, m, ra, dacore, rx, dxcore, out=None):
    """``A*x`` (``@tt_matrix/mtimes.m:38-65``) on CUDA float64 cores; returns (ry, core view)."""
    import torch
    d = len(n)
    ry = np.asarray(ra, dtype=np.int64) * np.asarray(rx, dtype=np.int64)
    cap = int(_ps(n, ry)[-1])
    if out is None:
        out = torch.empty(max(cap, 1), dtype=torch.float64, device=dacore.device)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    s = ctx.lib.ttg_mtimes_dev(ctx.h, d, i64(n), i64(m), i64(ra), _ptr(dacore), i64(rx), _ptr(dxcore),
                               r_out, _ptr(out), int(out.numel()), C.byref(used))
    ctx.check(s)
    return np.array(r_out[:], dtype=np.int64), out[:used.value]


def mtimes_round_dev(ctx, n, m, ra, dacore, rx, dxcore, eps, rmax=None, precision="fp64", out=None,
                     info=None):
    import torch
    d = len(n)
    ry = np.asarray(ra, dtype=np.int64) * np.asarray(rx, dtype=np.int64)
    cap = capacity(n, ry)
    if out is None:
        out = torch.empty(max(cap, 1), dtype=torch.float64, device=dacore.device)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    inf = TTGInfo()
    s = ctx.lib.ttg_mtimes_round_dev(ctx.h, d, i64(n), i64(m), i64(ra), _ptr(dacore), i64(rx),
                                     _ptr(dxcore), float(eps), _rmax_arg(rmax, d), _prec(precision),
                                     r_out, _ptr(out), int(out.numel()), C.byref(used), C.byref(inf))
    ctx.check(s, allow=(TTG_OK, TTG_NO_CONVERGENCE))
    if info is not None:
        info.update(inf.as_dict())
        info["status"] = s
    return np.array(r_out[:], dtype=np.int64), out[:used.value]


def round_batched_dev(ctx, tensors_n, tensors_r, dcore, cofs, eps, precision="fp64", info=None):
    """Round many TT tensors packed in one CUDA buffer (``cofs[t]`` = start of tensor t)."""
    import torch
    count = len(tensors_n)
    ds = [len(nn) for nn in tensors_n]
    nofs = np.concatenate([[0], np.cumsum(ds)])[:-1]
    rofs = np.concatenate([[0], np.cumsum([dd + 1 for dd in ds])])[:-1]
    nflat = np.concatenate([np.asarray(nn, dtype=np.int64) for nn in tensors_n])
    rflat = np.concatenate([np.asarray(rr, dtype=np.int64) for rr in tensors_r])
    caps = [capacity(nn, rr) for nn, rr in zip(tensors_n, tensors_r)]
    cofs_out = np.concatenate([[0], np.cumsum(caps)])
    out = torch.empty(max(int(cofs_out[-1]), 1), dtype=torch.float64, device=dcore.device)
    r_out = (C.c_int64 * len(rflat))()
    inf = TTGInfo()
    s = ctx.lib.ttg_round_batched_dev(ctx.h, count, i64(ds), i64(nofs), i64(nflat), i64(rofs),
                                      i64(rflat), i64(cofs), _ptr(dcore), float(eps),
                                      _prec(precision), r_out, i64(cofs_out[:-1]), _ptr(out),
                                      C.byref(inf))
    ctx.check(s, allow=(TTG_OK, TTG_NO_CONVERGENCE))
    if info is not None:
        info.update(inf.as_dict())
    rr = np.array(r_out[:], dtype=np.int64)
    res = []
    for t in range(count):
        r_t = rr[rofs[t]:rofs[t] + ds[t] + 1]
        ln = int(_ps(tensors_n[t], r_t)[-1])
        res.append((r_t, out[int(cofs_out[t]):int(cofs_out[t]) + ln]))
    return res
