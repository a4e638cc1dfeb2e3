"""GPU parity of the adjoint kernels (completion loss and <y, A x>) with the
oracle's interface-formula gradients, through the C ABI and through the
torch.autograd.Function binding, plus FP64 central finite differences of the
GPU loss itself.

Tolerances: FP64 kernels agree with the oracle to 1e-11 relative (summation
order differs); FP32 kernels to 2e-4 relative to the largest gradient entry
(FP32 interface products over d=16 cores).
"""
import numpy as np
import pytest
import torch

from oracle import tt_oracle as O
from ttgpu import autograd as AG
from ttgpu import workloads as W

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _completion_case(seed, n, r, S):
    rng = np.random.default_rng(seed)
    x = O.rand_tt(rng, n, r)
    idx = np.stack([rng.integers(0, nn, size=S) for nn in n], axis=1).astype(np.int32)
    vals = rng.standard_normal(S)
    return x, idx, vals


def _gpu_completion(x, idx, vals, dtype):
    core = torch.tensor(x.core, dtype=dtype, device=DEV, requires_grad=True)
    loss = AG.completion_loss(core, x.n, x.r, torch.tensor(idx, device=DEV), torch.tensor(vals, device=DEV))
    (g,) = torch.autograd.grad(loss, core)
    return float(loss), g.double().cpu().numpy()


@pytest.mark.parametrize("seed,n,r,S", [
    (1, [3, 4, 3, 2], [1, 3, 4, 2, 1], 300),
    (2, [4] * 6, [1, 5, 7, 6, 7, 5, 1], 2000),
    (3, [2] * 9, [1, 2, 4, 8, 8, 8, 8, 4, 2, 1], 777),
])
def test_completion_f64_matches_oracle(gpu_ctx, seed, n, r, S):
    x, idx, vals = _completion_case(seed, n, r, S)
    loss_ref, grads = O.completion_value_grad(x, idx, vals)
    gref = np.concatenate([g.ravel(order="F") for g in grads])
    loss, g = _gpu_completion(x, idx, vals, torch.float64)
    assert abs(loss - loss_ref) <= 1e-11 * max(1.0, loss_ref)
    np.testing.assert_allclose(g, gref, rtol=0, atol=1e-11 * np.abs(gref).max())


def test_completion_f64_finite_differences(gpu_ctx):
    x, idx, vals = _completion_case(4, [3, 4, 3, 2], [1, 3, 4, 2, 1], 250)
    _, g = _gpu_completion(x, idx, vals, torch.float64)
    rng = np.random.default_rng(0)
    h = 1e-6
    for p in rng.choice(x.core.size, size=12, replace=False):
        cp, cm = x.core.copy(), x.core.copy()
        cp[p] += h
        cm[p] -= h
        lp, _ = _gpu_completion(O.TT(x.n, x.r, cp), idx, vals, torch.float64)
        lm, _ = _gpu_completion(O.TT(x.n, x.r, cm), idx, vals, torch.float64)
        fd = (lp - lm) / (2 * h)
        assert abs(g[p] - fd) <= 1e-6 * max(1.0, abs(fd))


def test_completion_config5_f32_vs_f64(gpu_ctx):
    """Config 5 (d=16, n=8, model rank 32, 65,536 samples): FP32 adjoint against the
    FP64 adjoint and the oracle on a 4,096-sample subset."""
    c = W.completion5()
    n, r, core = c["model"]
    x = O.TT(n, r, core)
    idx, vals = c["idx"], c["vals"]
    l64, g64 = _gpu_completion(x, idx, vals, torch.float64)
    l32, g32 = _gpu_completion(x, idx, vals, torch.float32)
    assert abs(l32 - l64) <= 1e-4 * l64
    assert np.abs(g32 - g64).max() <= 2e-4 * np.abs(g64).max()
    sub = slice(0, 4096)
    lref, grads = O.completion_value_grad(x, idx[sub], vals[sub])
    gref = np.concatenate([gg.ravel(order="F") for gg in grads])
    ls, gs = _gpu_completion(x, idx[sub], vals[sub], torch.float64)
    assert abs(ls - lref) <= 1e-10 * lref
    np.testing.assert_allclose(gs, gref, rtol=0, atol=1e-10 * np.abs(gref).max())


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_completion_recompute_mode_bitwise(gpu_ctx, monkeypatch, dtype):
    """With a zero workspace budget the left interfaces are recomputed core by
    core instead of stored; the loss and the gradient are bitwise those of the
    stored mode (same products, same chunks), and the gradient still matches
    the oracle."""
    x, idx, vals = _completion_case(5, [4] * 7, [1, 5, 9, 12, 12, 9, 5, 1], 3000)
    monkeypatch.delenv("TTG_ADJOINT_LBUDGET_MB", raising=False)
    l_st, g_st = _gpu_completion(x, idx, vals, dtype)
    monkeypatch.setenv("TTG_ADJOINT_LBUDGET_MB", "0")
    l_rc, g_rc = _gpu_completion(x, idx, vals, dtype)
    assert l_rc == l_st
    assert np.array_equal(g_rc, g_st)
    if dtype == torch.float64:
        lref, grads = O.completion_value_grad(x, idx, vals)
        gref = np.concatenate([gg.ravel(order="F") for gg in grads])
        assert abs(l_rc - lref) <= 1e-11 * max(1.0, lref)
        np.testing.assert_allclose(g_rc, gref, rtol=0, atol=1e-11 * np.abs(gref).max())


def _bilinear_case(seed, n, m, ry, ra, rx):
    rng = np.random.default_rng(seed)
    acores = [rng.standard_normal((ra[k], n[k], m[k], ra[k + 1])) / np.sqrt(n[k] * ra[k])
              for k in range(len(n))]
    a = O.TTM.from_cores(acores)
    x = O.rand_tt(rng, m, rx)
    y = O.rand_tt(rng, n, ry)
    return y, a, x


def _gpu_bilinear(y, a, x, dtype):
    yc = torch.tensor(y.core, dtype=dtype, device=DEV, requires_grad=True)
    xc = torch.tensor(x.core, dtype=dtype, device=DEV, requires_grad=True)
    ac = torch.tensor(a.tt.core, dtype=dtype, device=DEV)
    val = AG.bilinear(yc, xc, ac, a.n, a.m, y.r, a.r, x.r)
    gy, gx = torch.autograd.grad(val, (yc, xc))
    return float(val), gx.double().cpu().numpy(), gy.double().cpu().numpy()


@pytest.mark.parametrize("seed,d,nn,rr,rar", [(5, 4, 3, 3, 2), (6, 20, 4, 20, 5)])
def test_bilinear_matches_oracle(gpu_ctx, seed, d, nn, rr, rar):
    """d=20, n=m=4, ranks 20 (x, y) and 5 (A): the config-1 shapes."""
    ranks = lambda q: [1] + [q] * (d - 1) + [1]
    y, a, x = _bilinear_case(seed, [nn] * d, [nn] * d, ranks(rr), ranks(rar), ranks(rr))
    val_ref, gx, gy = O.bilinear_value_grad(y, a, x)
    gxr = np.concatenate([g.ravel(order="F") for g in gx])
    gyr = np.concatenate([g.ravel(order="F") for g in gy])
    val, gx64, gy64 = _gpu_bilinear(y, a, x, torch.float64)
    assert abs(val - val_ref) <= 1e-11 * max(1.0, abs(val_ref))
    np.testing.assert_allclose(gx64, gxr, rtol=0, atol=1e-11 * np.abs(gxr).max())
    np.testing.assert_allclose(gy64, gyr, rtol=0, atol=1e-11 * np.abs(gyr).max())
    val32, gx32, gy32 = _gpu_bilinear(y, a, x, torch.float32)
    assert abs(val32 - val_ref) <= 1e-4 * max(1.0, abs(val_ref))
    assert np.abs(gx32 - gxr).max() <= 2e-4 * np.abs(gxr).max()
    assert np.abs(gy32 - gyr).max() <= 2e-4 * np.abs(gyr).max()


def test_bilinear_nonsquare_finite_differences(gpu_ctx):
    y, a, x = _bilinear_case(7, [2, 3, 2], [3, 2, 2], [1, 3, 2, 1], [1, 2, 3, 1], [1, 2, 3, 1])
    _, gx, _ = _gpu_bilinear(y, a, x, torch.float64)
    h = 1e-6
    for p in range(x.core.size):
        cp, cm = x.core.copy(), x.core.copy()
        cp[p] += h
        cm[p] -= h
        vp, _, _ = _gpu_bilinear(y, a, O.TT(x.n, x.r, cp), torch.float64)
        vm, _, _ = _gpu_bilinear(y, a, O.TT(x.n, x.r, cm), torch.float64)
        assert abs(gx[p] - (vp - vm) / (2 * h)) <= 1e-6
