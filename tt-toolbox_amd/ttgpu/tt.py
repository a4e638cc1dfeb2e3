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


def _np_rmax(v, d):
    v = np.asarray(v, dtype=np.int64).ravel()
    if v.size == 1:
        v = np.full(d + 1, int(v[0]), dtype=np.int64)
    if v.size != d + 1:
        raise ValueError("rmax must be a scalar or have d+1 entries")
    return v


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
    n_arr, r_arr = i64(n), i64(r)  # marshalled once per call
    if out is None:  # the capacity is needed only to allocate the output
        cap = C.c_int64()
        sc = ctx.lib.ttg_round_capacity(d, n_arr, r_arr, C.byref(cap))
        if sc != TTG_OK:
            raise _lib.TTGError(sc, "ttg_round_capacity")
        out = torch.empty(max(int(cap.value), 1), dtype=torch.float64, device=dcore.device)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    inf = TTGInfo()
    s = ctx.lib.ttg_round_dev(ctx.h, d, n_arr, r_arr, _ptr(dcore), float(eps), _rmax_arg(rmax, d),
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


def round_batched_dev(ctx, tensors_n, tensors_r, dcore, cofs, eps, precision="fp64", info=None, rmax=None):
    """Round many TT tensors packed in one CUDA buffer (``cofs[t]`` = start of tensor t).
    ``rmax``: None, a scalar for every bond of every tensor, or one d+1 vector per tensor."""
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
    rm = None
    if rmax is not None:
        if np.isscalar(rmax):
            rm = i64(np.full(len(rflat), int(rmax), dtype=np.int64))
        else:
            rm = i64(np.concatenate([_np_rmax(v, dd) for v, dd in zip(rmax, ds)]))
    s = ctx.lib.ttg_round_batched_dev(ctx.h, count, i64(ds), i64(nofs), i64(nflat), i64(rofs),
                                      i64(rflat), i64(cofs), _ptr(dcore), float(eps), rm,
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


# ---------------------------------------------------------------------------
# sums, norms and inner products (device pointers; host forms stage through torch)
# ---------------------------------------------------------------------------
def _plus_ranks(rx, ry):
    rx = np.asarray(rx, dtype=np.int64)
    ry = np.asarray(ry, dtype=np.int64)
    rz = rx + ry
    rz[0], rz[-1] = rx[0], rx[-1]
    return rz


def plus_dev(ctx, n, rx, dx, ry, dy, a=1.0, b=1.0, out=None):
    """``a*X + b*Y`` in the rank concatenation of ``@tt_tensor/plus.m:41-86``
    (scalars on core 1, ``@tt_tensor/mtimes.m:10-27``); returns (rz, core view)."""
    import torch
    d = len(n)
    cap = int(_ps(n, _plus_ranks(rx, ry))[-1])
    if out is None:
        out = torch.empty(max(cap, 1), dtype=torch.float64, device=dx.device)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    s = ctx.lib.ttg_plus_dev(ctx.h, d, i64(n), i64(rx), _ptr(dx), float(a), i64(ry), _ptr(dy), float(b),
                             r_out, _ptr(out), int(out.numel()), C.byref(used))
    ctx.check(s)
    return np.array(r_out[:], dtype=np.int64), out[:used.value]


def plus_round_dev(ctx, n, rx, dx, ry, dy, eps, a=1.0, b=1.0, rmax=None, precision="fp64", out=None,
                   info=None):
    """``round(a*X + b*Y, eps, rmax)`` with the sum never stored as a TT by the
    sweeps; returns (r_out, core view)."""
    import torch
    d = len(n)
    cap = capacity(n, _plus_ranks(rx, ry))
    if out is None:
        out = torch.empty(max(cap, 1), dtype=torch.float64, device=dx.device)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    inf = TTGInfo()
    s = ctx.lib.ttg_plus_round_dev(ctx.h, d, i64(n), i64(rx), _ptr(dx), float(a), i64(ry), _ptr(dy), float(b),
                                   float(eps), _rmax_arg(rmax, d), _prec(precision), r_out, _ptr(out),
                                   int(out.numel()), C.byref(used), C.byref(inf))
    ctx.check(s, allow=(TTG_OK, TTG_NO_CONVERGENCE))
    if info is not None:
        info.update(inf.as_dict())
        info["status"] = s
    return np.array(r_out[:], dtype=np.int64), out[:used.value]


def norm_dev(ctx, n, r, dcore):
    """``norm(X)`` (``@tt_tensor/norm.m:25-41``, QR sweep) of a device-resident TT."""
    v = C.c_double()
    ctx.check(ctx.lib.ttg_norm_dev(ctx.h, len(n), i64(n), i64(r), _ptr(dcore), C.byref(v)))
    return float(v.value)


def plus_norm_dev(ctx, n, rx, dx, ry, dy, a=1.0, b=-1.0):
    """``norm(a*X + b*Y)`` without storing the sum (default: ``norm(X - Y)``)."""
    v = C.c_double()
    ctx.check(ctx.lib.ttg_plus_norm_dev(ctx.h, len(n), i64(n), i64(rx), _ptr(dx), float(a), i64(ry), _ptr(dy),
                                        float(b), C.byref(v)))
    return float(v.value)


def dot_dev(ctx, n, rx, dx, ry, dy):
    """``dot(X, Y)`` (``@tt_tensor/dot.m:92-111``) of two device-resident TTs."""
    v = C.c_double()
    ctx.check(ctx.lib.ttg_dot_dev(ctx.h, len(n), i64(n), i64(rx), _ptr(dx), i64(ry), _ptr(dy), C.byref(v)))
    return float(v.value)


def _staged(*tts):
    """Device copies (cuda:0) of host TT cores; torch's stream is drained so
    that a context on its own stream sees them."""
    import torch
    out = [torch.from_numpy(t.core).to("cuda:0") for t in tts]
    torch.cuda.synchronize()
    return out


def plus(x: TTTensor, y: TTTensor, a=1.0, b=1.0, ctx=None) -> TTTensor:
    """``a*x + b*y`` (``@tt_tensor/plus.m:41-86``) on the GPU."""
    ctx = ctx or default_context()
    dx, dy = _staged(x, y)
    rz, z = plus_dev(ctx, x.n, x.r, dx, y.r, dy, a, b)
    return TTTensor(x.n.copy(), rz, z.cpu().numpy().copy())


def plus_round(x: TTTensor, y: TTTensor, eps, a=1.0, b=1.0, rmax=None, precision="fp64", ctx=None,
               info=None) -> TTTensor:
    """``round(a*x + b*y, eps, rmax)`` on the GPU, the sum fused into the round."""
    ctx = ctx or default_context()
    dx, dy = _staged(x, y)
    r, z = plus_round_dev(ctx, x.n, x.r, dx, y.r, dy, eps, a, b, rmax, precision, info=info)
    return TTTensor(x.n.copy(), r, z.cpu().numpy().copy())


def norm(x: TTTensor, ctx=None) -> float:
    """``norm(x)`` (``@tt_tensor/norm.m:25-41``) on the GPU."""
    ctx = ctx or default_context()
    (dx,) = _staged(x)
    return norm_dev(ctx, x.n, x.r, dx)


def dot(x: TTTensor, y: TTTensor, ctx=None) -> float:
    """``dot(x, y)`` (``@tt_tensor/dot.m:92-111``) on the GPU."""
    ctx = ctx or default_context()
    dx, dy = _staged(x, y)
    return dot_dev(ctx, x.n, x.r, dx, y.r, dy)


def rel_diff(y: TTTensor, yref: TTTensor, ctx=None) -> float:
    """``norm(y - yref) / norm(yref)`` on the GPU (the sum is never stored)."""
    ctx = ctx or default_context()
    dy, dr = _staged(y, yref)
    den = norm_dev(ctx, yref.n, yref.r, dr)
    num = plus_norm_dev(ctx, y.n, y.r, dy, yref.r, dr, 1.0, -1.0)
    return num / den if den > 0 else num


# ---------------------------------------------------------------------------
# launch interface of the Gram-chain round (repeated rounds of one shape)
# ---------------------------------------------------------------------------
def round_gram_prepare(ctx, n, r, dcore, eps, precision="fp32", out=None):
    """Round once and keep the captured launch sequence; returns (r_out, core view, proven)."""
    import torch
    d = len(n)
    cap = capacity(n, r)
    if out is None:
        out = torch.empty(max(cap, 1), dtype=torch.float64, device=dcore.device)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    proven = C.c_int()
    ctx.check(ctx.lib.ttg_round_gram_prepare(ctx.h, d, i64(n), i64(r), _ptr(dcore), float(eps), _prec(precision),
                                             r_out, _ptr(out), int(out.numel()), C.byref(used), C.byref(proven)))
    return np.array(r_out[:], dtype=np.int64), out[:used.value], bool(proven.value)


def round_gram_launch(ctx, dstatus):
    """Re-round the prepared input buffer on the context's stream (no host
    synchronisation); ``dstatus`` is a one-element int32 device tensor."""
    ctx.check(ctx.lib.ttg_round_gram_launch(ctx.h, _ptr(dstatus)))


def round_gram_graph(ctx):
    """The prepared round as a hipGraph_t handle (for hipGraphAddChildGraphNode)."""
    g = C.c_void_p()
    ctx.check(ctx.lib.ttg_round_gram_graph(ctx.h, C.byref(g)))
    return g


# ---------------------------------------------------------------------------
# TT-SVD (tt_tensor(full, eps), @tt_tensor/tt_tensor.m:175-244)
# ---------------------------------------------------------------------------
def full_to_tt_dev(ctx, n, dfull, eps=1e-14, rmax=None, info=None):
    """TT-SVD of a column-major full tensor in device memory; returns (r, core view)."""
    import torch
    n = np.asarray(n, dtype=np.int64)
    d = int(n.size)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64()
    inf = TTGInfo()
    cap = max(int(dfull.numel()), 1)
    for _ in range(2):
        out = torch.empty(cap, dtype=torch.float64, device=dfull.device)
        s = ctx.lib.ttg_full_to_tt_dev(ctx.h, d, i64(n), _ptr(dfull), float(eps), _rmax_arg(rmax, d), r_out,
                                       _ptr(out), cap, C.byref(used), C.byref(inf))
        if s != _lib.TTG_NEED_SPACE:
            break
        cap = int(used.value)
    ctx.check(s, allow=(TTG_OK, TTG_NO_CONVERGENCE))
    if info is not None:
        info.update(inf.as_dict())
        info["status"] = s
    return np.array(r_out[:], dtype=np.int64), out[:used.value]


def tt_tensor_from_full(full, eps=1e-14, n=None, rmax=None, ctx=None, info=None) -> TTTensor:
    """``tt_tensor(full, eps, n)`` on the GPU (``full`` a numpy array, column-major
    order as MATLAB's ``b(:)``; ``n`` defaults to ``full.shape``)."""
    import torch
    ctx = ctx or default_context()
    a = np.asarray(full, dtype=np.float64)
    n = np.asarray(a.shape if n is None else n, dtype=np.int64).ravel()
    flat = np.ascontiguousarray(a.reshape(-1, order="F"))
    if flat.size != int(np.prod(n)):
        raise ValueError("prod(n) must equal the number of entries")
    d = torch.from_numpy(flat).to("cuda:0")
    torch.cuda.synchronize()
    r, core = full_to_tt_dev(ctx, n, d, eps, rmax, info)
    return TTTensor(n.copy(), r, core.cpu().numpy().copy())


# ---------------------------------------------------------------------------
# SDV TT files (fmex/tt_read.c, fmex/tt_write.c; tt_tensor(filename), tt_save)
# ---------------------------------------------------------------------------
def _path(p):
    import os
    return os.fsencode(p)


def sdv_info(path):
    """(d, core length, dtype) of an SDV TT file (dtype 0 double, 2 float)."""
    lib = _lib.load()
    d, ln, dt = C.c_int64(), C.c_int64(), C.c_int()
    s = lib.ttg_sdv_info(_path(path), C.byref(d), C.byref(ln), C.byref(dt))
    if s != TTG_OK:
        raise _lib.TTGError(s, f"ttg_sdv_info({path})")
    return int(d.value), int(ln.value), int(dt.value)


def tt_read(path) -> TTTensor:
    """``tt_tensor(filename)`` (``@tt_tensor/tt_tensor.m:149-165``): host arrays."""
    lib = _lib.load()
    d, ln, _ = sdv_info(path)
    n, r = (C.c_int64 * d)(), (C.c_int64 * (d + 1))()
    core = np.empty(max(ln, 1), dtype=np.float64)
    dd, used = C.c_int64(), C.c_int64()
    s = lib.ttg_sdv_read(_path(path), d, C.byref(dd), n, r, core.ctypes.data_as(C.c_void_p), ln, C.byref(used))
    if s != TTG_OK:
        raise _lib.TTGError(s, f"ttg_sdv_read({path})")
    return TTTensor(np.array(n[:], np.int64), np.array(r[:], np.int64), core[:used.value])


def tt_save(path, tt: TTTensor, precision=0):
    """``tt_save(filename, tt, precision)`` (``@tt_tensor/tt_save.m:22``): 0 double, 1 float."""
    lib = _lib.load()
    s = lib.ttg_sdv_write(_path(path), tt.d, i64(tt.n), i64(tt.r), tt.core.ctypes.data_as(C.c_void_p),
                          int(precision))
    if s != TTG_OK:
        raise _lib.TTGError(s, f"ttg_sdv_write({path})")


def tt_read_dev(ctx, path, device="cuda:0"):
    """Read an SDV TT file straight into device memory; returns (n, r, dcore)."""
    import torch
    d, ln, _ = sdv_info(path)
    n, r = (C.c_int64 * d)(), (C.c_int64 * (d + 1))()
    out = torch.empty(max(ln, 1), dtype=torch.float64, device=device)
    dd, used = C.c_int64(), C.c_int64()
    ctx.check(ctx.lib.ttg_sdv_read_dev(ctx.h, _path(path), d, C.byref(dd), n, r, _ptr(out), int(out.numel()),
                                       C.byref(used)))
    return np.array(n[:], np.int64), np.array(r[:], np.int64), out[:used.value]


def tt_save_dev(ctx, path, n, r, dcore, precision=0):
    """Write a device-resident TT to an SDV TT file (0 double, 1 float)."""
    ctx.check(ctx.lib.ttg_sdv_write_dev(ctx.h, _path(path), len(n), i64(n), i64(r), _ptr(dcore), int(precision)))

