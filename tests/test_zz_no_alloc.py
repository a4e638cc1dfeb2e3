"""The no-allocation entry point ttg_adjoint_bilinear_f64_noalloc against
ttg_adjoint_bilinear_f64 through the C ABI: two calls in a row (the second
with a smaller input), value and both gradients bitwise equal (same kernels,
same order, scratch in the static device array instead of the context's
workspace), the value also against the oracle's <y, A x>; ranks whose scratch
exceeds the array are refused with TTG_BAD_ARG before any kernel runs."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import tt_oracle as O
from ttgpu import _lib
from ttgpu._lib import i64

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _case(seed, d, n, ry, ra, rx):
    rng = np.random.default_rng(seed)
    rs = lambda r: [1] + [r] * (d - 1) + [1]
    acores = [rng.standard_normal((a0, n, n, a1)) / np.sqrt(n * a0) for a0, a1 in zip(rs(ra)[:-1], rs(ra)[1:])]
    a = O.TTM.from_cores(acores)
    x = O.rand_tt(rng, [n] * d, rs(rx))
    y = O.rand_tt(rng, [n] * d, rs(ry))
    return a, x, y


def _call(ctx, fn, a, x, y):
    d = x.d
    dy = torch.tensor(y.core, device=DEV)
    da = torch.tensor(a.tt.core, device=DEV)
    dx = torch.tensor(x.core, device=DEV)
    gx = torch.zeros_like(dx)
    gy = torch.zeros_like(dy)
    val = C.c_double()
    p = lambda t: C.c_void_p(t.data_ptr())
    s = fn(ctx.h, d, i64(a.n), i64(a.m), i64(y.r), p(dy), i64(a.tt.r), p(da), i64(x.r), p(dx), C.byref(val), p(gx), p(gy))
    torch.cuda.synchronize()
    return s, val.value, gx.cpu().numpy(), gy.cpu().numpy()


def test_noalloc_matches_the_allocating_entry_point(gpu_ctx):
    lib = gpu_ctx.lib
    for case in ((71, 12, 4, 20, 5, 20), (72, 6, 3, 8, 3, 6)):  # the second call is smaller
        a, x, y = _case(*case)
        s1, v1, gx1, gy1 = _call(gpu_ctx, lib.ttg_adjoint_bilinear_f64, a, x, y)
        s2, v2, gx2, gy2 = _call(gpu_ctx, lib.ttg_adjoint_bilinear_f64_noalloc, a, x, y)
        assert s1 == s2 == _lib.TTG_OK
        assert v1 == v2
        assert gx1.tobytes() == gx2.tobytes() and gy1.tobytes() == gy2.tobytes()
        ref = O.bilinear_value_grad(y, a, x)[0]
        assert abs(v2 - ref) <= 1e-12 * max(1.0, abs(ref))


def test_noalloc_refuses_ranks_beyond_its_scratch(gpu_ctx):
    a, x, y = _case(73, 4, 4, 120, 8, 120)  # 2 * 120*8*120 * 5 doubles of interfaces > 8 MiB
    s, _, _, _ = _call(gpu_ctx, gpu_ctx.lib.ttg_adjoint_bilinear_f64_noalloc, a, x, y)
    assert s == _lib.TTG_BAD_ARG
    assert b"scratch" in gpu_ctx.lib.ttg_last_error(gpu_ctx.h)
