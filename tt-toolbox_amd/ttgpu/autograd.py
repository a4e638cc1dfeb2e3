"""torch.autograd bindings of the hand-written adjoint kernels (test binding).

The forward value and every core gradient come from libttgpu.so in one call
(``ttg_adjoint_completion_*`` / ``ttg_adjoint_bilinear_*``); backward only
scales the stored gradient by the incoming scalar.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import default_context, i64


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class CompletionLoss(torch.autograd.Function):
    """``0.5*||P_Omega(X) - y||^2`` of a TT X given by its flat core vector."""

    @staticmethod
    def forward(ctx, core, n, r, idx, vals, tctx=None):
        tctx = tctx or default_context(core.device.index or 0)
        core = core.contiguous()
        idx = idx.to(torch.int32).contiguous()
        vals = vals.to(core.dtype).contiguous()
        grad = torch.empty_like(core)
        loss = C.c_double()
        fn = tctx.lib.ttg_adjoint_completion_f64 if core.dtype == torch.float64 else tctx.lib.ttg_adjoint_completion_f32
        if core.dtype not in (torch.float32, torch.float64):
            raise TypeError("core must be float32 or float64")
        torch.cuda.synchronize(core.device)
        s = fn(tctx.h, len(n), i64(n), i64(r), _ptr(core), int(idx.shape[0]), _ptr(idx), _ptr(vals),
               C.byref(loss), _ptr(grad))
        tctx.check(s)
        ctx.save_for_backward(grad)
        return torch.tensor(loss.value, dtype=core.dtype, device=core.device)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout * grad, None, None, None, None, None


class Bilinear(torch.autograd.Function):
    """``<y, A x>`` with gradients for the flat cores of x and y (A held fixed)."""

    @staticmethod
    def forward(ctx, ycore, xcore, acore, n, m, ry, ra, rx, tctx=None):
        tctx = tctx or default_context(xcore.device.index or 0)
        ycore, xcore, acore = ycore.contiguous(), xcore.contiguous(), acore.contiguous()
        gx = torch.empty_like(xcore)
        gy = torch.empty_like(ycore)
        val = C.c_double()
        fn = tctx.lib.ttg_adjoint_bilinear_f64 if xcore.dtype == torch.float64 else tctx.lib.ttg_adjoint_bilinear_f32
        torch.cuda.synchronize(xcore.device)
        s = fn(tctx.h, len(n), i64(n), i64(m), i64(ry), _ptr(ycore), i64(ra), _ptr(acore), i64(rx),
               _ptr(xcore), C.byref(val), _ptr(gx), _ptr(gy))
        tctx.check(s)
        ctx.save_for_backward(gy, gx)
        return torch.tensor(val.value, dtype=xcore.dtype, device=xcore.device)

    @staticmethod
    def backward(ctx, gout):
        gy, gx = ctx.saved_tensors
        return gout * gy, gout * gx, None, None, None, None, None, None, None


def completion_loss(core, n, r, idx, vals, tctx=None):
    return CompletionLoss.apply(core, list(map(int, n)), list(map(int, r)), idx, vals, tctx)


def bilinear(ycore, xcore, acore, n, m, ry, ra, rx, tctx=None):
    return Bilinear.apply(ycore, xcore, acore, list(map(int, n)), list(map(int, m)),
                          list(map(int, ry)), list(map(int, ra)), list(map(int, rx)), tctx)
