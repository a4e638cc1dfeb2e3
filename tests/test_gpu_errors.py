"""Status codes of the C ABI forced from the host, and run-to-run determinism (GPU).

* TTG_NEED_SPACE: output capacity one short; *len_out reports the length
  needed (round, A*x, fused round(A*x); host and device forms).
* TTG_NO_CONVERGENCE: the Jacobi sweep cap lowered to 1 through
  TTG_JACOBI_MAX_SWEEPS (analogue of round.m:64-74's svd -> svds warning);
  the call still returns its result.  Every kernel exits normally.
* TTG_BAD_ARG: A*x sizes are checked before any staging allocation;
  completion sample indices are checked before any gather.
* Bitwise determinism: fixed reduction orders, so two calls on the same input
  give identical bytes (multi-CU and one-workgroup paths, both forms, fused
  A*x, batched).
"""
import ctypes as C

import numpy as np
import pytest
import torch

import ttgpu
from ttgpu import _lib
from ttgpu._lib import i64
from ttgpu import autograd as AG
from oracle import tt_oracle as O

pytestmark = pytest.mark.gpu


def _vp(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("ranks", [[1, 4, 16, 20, 16, 4, 1], [1, 4, 16, 64, 80, 16, 4, 1]])
def test_need_space_reports_required_length(gpu_ctx, ranks):
    rng = np.random.default_rng(51)
    d = len(ranks) - 1
    x = O.rand_tt(rng, [4] * d, ranks)
    lib, h = gpu_ctx.lib, gpu_ctx.h
    cap = ttgpu.capacity(x.n, x.r)
    out = np.empty(cap)
    r_out = (C.c_int64 * (d + 1))()
    used = C.c_int64(-1)
    s = lib.ttg_round(h, d, i64(x.n), i64(x.r), _vp(x.core), 1e-8, None, 0, r_out, _vp(out), cap - 1,
                      C.byref(used), None)
    assert s == _lib.TTG_NEED_SPACE and used.value == cap
    dcore = torch.tensor(x.core, device="cuda:0")
    dout = torch.empty(cap, dtype=torch.float64, device="cuda:0")
    used = C.c_int64(-1)
    s = lib.ttg_round_dev(h, d, i64(x.n), i64(x.r), C.c_void_p(dcore.data_ptr()), 1e-8, None, 0, r_out,
                          C.c_void_p(dout.data_ptr()), cap - 1, C.byref(used), None)
    assert s == _lib.TTG_NEED_SPACE and used.value == cap


def test_need_space_mtimes(gpu_ctx):
    rng = np.random.default_rng(52)
    n = m = [2, 3, 2]
    ra, rx = [1, 2, 3, 1], [1, 3, 2, 1]
    a = O.TTM.from_cores([rng.standard_normal((ra[k], n[k], m[k], ra[k + 1])) for k in range(3)])
    x = O.rand_tt(rng, m, rx)
    lib, h = gpu_ctx.lib, gpu_ctx.h
    ry = [p * q for p, q in zip(ra, rx)]
    need = sum(ry[k] * n[k] * ry[k + 1] for k in range(3))
    out = np.empty(need)
    r_out = (C.c_int64 * 4)()
    used = C.c_int64(-1)
    s = lib.ttg_mtimes(h, 3, i64(n), i64(m), i64(ra), _vp(a.tt.core), i64(rx), _vp(x.core), r_out, _vp(out),
                       need - 1, C.byref(used))
    assert s == _lib.TTG_NEED_SPACE and used.value == need
    capr = ttgpu.capacity(n, ry)
    used = C.c_int64(-1)
    s = lib.ttg_mtimes_round(h, 3, i64(n), i64(m), i64(ra), _vp(a.tt.core), i64(rx), _vp(x.core), 1e-8, None, 0,
                             r_out, _vp(out), capr - 1, C.byref(used), None)
    assert s == _lib.TTG_NEED_SPACE and used.value == capr


@pytest.mark.parametrize("ranks", [[1, 3, 9, 12, 9, 3, 1], [1, 4, 16, 40, 40, 16, 4, 1]])
def test_no_convergence_status(gpu_ctx, monkeypatch, ranks):
    rng = np.random.default_rng(53)
    d = len(ranks) - 1
    x = O.rand_tt(rng, [3 if ranks[3] == 12 else 4] * d, ranks)
    z = O.tt_plus(x, x)  # every bond truncates: the SVD path runs
    monkeypatch.setenv("TTG_JACOBI_MAX_SWEEPS", "1")
    info = {}
    y = ttgpu.round(ttgpu.TTTensor(z.n, z.r, z.core), 1e-10, info=info)
    assert info["status"] == _lib.TTG_NO_CONVERGENCE
    assert y.core.size == int(O.TT(y.n, y.r, np.zeros(y.core.size)).ps[-1])
    monkeypatch.delenv("TTG_JACOBI_MAX_SWEEPS")
    info = {}
    ttgpu.round(ttgpu.TTTensor(z.n, z.r, z.core), 1e-10, info=info)
    assert info["status"] == _lib.TTG_OK


def test_mtimes_sizes_checked_before_staging(gpu_ctx):
    lib, h = gpu_ctx.lib, gpu_ctx.h
    buf = np.zeros(64)
    r_out = (C.c_int64 * 3)()
    used = C.c_int64()
    n = m = [2, 2]
    for ra, rx in (([1, -1, 1], [1, 2, 1]), ([1, 2, 1], [1, 0, 1]), ([-1, -1, -1], [-1, -1, -1])):
        s = lib.ttg_mtimes(h, 2, i64(n), i64(m), i64(ra), _vp(buf), i64(rx), _vp(buf), r_out, _vp(buf), 64,
                           C.byref(used))
        assert s == _lib.TTG_BAD_ARG, (ra, rx)
        s = lib.ttg_mtimes_round(h, 2, i64(n), i64(m), i64(ra), _vp(buf), i64(rx), _vp(buf), 1e-8, None, 0,
                                 r_out, _vp(buf), 64, C.byref(used), None)
        assert s == _lib.TTG_BAD_ARG, (ra, rx)
    s = lib.ttg_mtimes(h, 2, i64([2, -3]), i64(m), i64([1, 2, 1]), _vp(buf), i64([1, 2, 1]), _vp(buf), r_out,
                       _vp(buf), 64, C.byref(used))
    assert s == _lib.TTG_BAD_ARG


@pytest.mark.parametrize("bad", [4, -1])
def test_completion_index_out_of_range(gpu_ctx, bad):
    rng = np.random.default_rng(54)
    n, r = [3, 4, 3], [1, 2, 3, 1]
    x = O.rand_tt(rng, n, r)
    idx = np.stack([rng.integers(0, nn, size=50) for nn in n], axis=1).astype(np.int32)
    idx[17, 1] = bad  # mode 1 has size 4
    core = torch.tensor(x.core, device="cuda:0")
    didx = torch.tensor(idx, device="cuda:0")
    vals = torch.zeros(50, dtype=torch.float64, device="cuda:0")
    grad = torch.zeros_like(core)
    loss = C.c_double()
    s = gpu_ctx.lib.ttg_adjoint_completion_f64(gpu_ctx.h, 3, i64(n), i64(r), C.c_void_p(core.data_ptr()), 50,
                                               C.c_void_p(didx.data_ptr()), C.c_void_p(vals.data_ptr()),
                                               C.byref(loss), C.c_void_p(grad.data_ptr()))
    assert s == _lib.TTG_BAD_ARG
    assert "index" in gpu_ctx.lib.ttg_last_error(gpu_ctx.h).decode()
    with pytest.raises(ttgpu.TTGError):
        AG.completion_loss(core.clone().requires_grad_(True), n, r, didx, vals)


@pytest.mark.parametrize("ranks,prec", [
    ([1, 4, 16, 20, 20, 16, 4, 1], "fp64"),            # one-workgroup path
    ([1, 4, 16, 64, 96, 96, 16, 4, 1], "fp64"),        # multi-CU path
    ([1, 4, 16, 64, 96, 96, 16, 4, 1], "fp32"),        # mixed form
])
def test_round_twice_is_bitwise_identical(gpu_ctx, ranks, prec):
    rng = np.random.default_rng(55)
    d = len(ranks) - 1
    x = O.rand_tt(rng, [4] * d, ranks)
    z = O.tt_plus(x, O.tt_scal(0.5, O.rand_tt(rng, [4] * d, ranks)))
    y1 = ttgpu.round(ttgpu.TTTensor(z.n, z.r, z.core), 1e-6, precision=prec)
    y2 = ttgpu.round(ttgpu.TTTensor(z.n, z.r, z.core), 1e-6, precision=prec)
    assert y1.r.tolist() == y2.r.tolist()
    assert y1.core.tobytes() == y2.core.tobytes()


def test_fused_mtimes_round_and_batched_twice_bitwise(gpu_ctx):
    c = O.make_config(2, d=6, rank=16, seed=56)
    a, x = c["a"], c["x"]
    A = ttgpu.TTMatrix(a.n, a.m, a.tt.r, a.tt.core)
    X = ttgpu.TTTensor(x.n, x.r, x.core)
    y1 = ttgpu.mtimes_round(A, X, 1e-8)
    y2 = ttgpu.mtimes_round(A, X, 1e-8)
    assert y1.core.tobytes() == y2.core.tobytes()
    batch = O.make_batch(count=12, seed=57, dmin=3, dmax=6, nset=(2, 4), rmin=2, rmax=12)
    cores = [t.core for t in batch]
    cofs = np.concatenate([[0], np.cumsum([c_.size for c_ in cores])])
    dcore = torch.tensor(np.concatenate(cores), device="cuda:0")
    outs = []
    for _ in range(2):
        res = ttgpu.round_batched_dev(gpu_ctx, [t.n for t in batch], [t.r for t in batch], dcore, cofs[:-1], 1e-8)
        outs.append([y.cpu().numpy().tobytes() for _, y in res])
    assert outs[0] == outs[1]
