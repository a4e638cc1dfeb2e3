"""Sums, norms and inner products on the GPU against the FP64 oracle.

* ``a X + b Y`` (``@tt_tensor/plus.m:41-86`` with the scalars of
  ``@tt_tensor/mtimes.m:10-27`` on core 1): bitwise equal to the oracle's
  ``tt_plus(tt_scal(a, X), tt_scal(b, Y))`` (the same products, no sums), d = 1
  sums, and the boundary-rank check of ``plus.m:50-59``.
* ``round(a X + b Y)`` through all three forms: the sweeps with the sum
  generated inside sweep 1's products (TTG_GRAM=0, TTG_SMALL_PERSIST=0), the
  small-tensor form and the Gram-chain form on the assembled sum.  Bars: the
  oracle's ranks, relative difference <= 1e-12 (plain) / 1e-6 (mixed).
* ``norm`` (``@tt_tensor/norm.m:25-41``) against the oracle and the dense
  norm; ``norm(X + d W - X) = d norm(W)`` (closed form: the difference of two
  nearly equal tensors keeps its digits); ``dot`` (``@tt_tensor/dot.m:92-111``)
  against the oracle and the dense inner product, ``dot(X, X) = norm(X)^2``.
"""
import numpy as np
import pytest
import torch

import ttgpu
from oracle import tt_oracle as O
from ttgpu import workloads as W

pytestmark = pytest.mark.gpu


def _dev(x):
    t = torch.from_numpy(np.ascontiguousarray(x.core)).to("cuda:0")
    torch.cuda.synchronize()
    return t


def _t(x):
    return ttgpu.TTTensor(x.n, x.r, x.core)


@pytest.mark.parametrize("a,b", [(1.0, 1.0), (2.5, -0.75)])
def test_plus_bitwise_matches_oracle(gpu_ctx, a, b):
    rng = np.random.default_rng(101)
    n = [3, 4, 2, 5, 3]
    x = O.rand_tt(rng, n, [1, 3, 7, 6, 4, 1])
    y = O.rand_tt(rng, n, [1, 2, 5, 8, 3, 1])
    z = ttgpu.plus(_t(x), _t(y), a, b)
    ref = O.tt_plus(O.tt_scal(a, x), O.tt_scal(b, y))
    assert z.r.tolist() == ref.r.tolist()
    assert z.core.tobytes() == ref.core.tobytes()


def test_plus_d1_and_boundary_rank_check(gpu_ctx):
    rng = np.random.default_rng(102)
    u, v = O.rand_tt(rng, [7], [1, 1]), O.rand_tt(rng, [7], [1, 1])
    z = ttgpu.plus(_t(u), _t(v), 1.0, -2.0)
    np.testing.assert_array_equal(z.core, u.core - 2.0 * v.core)
    x = O.rand_tt(rng, [3, 3], [2, 4, 1])
    y = O.rand_tt(rng, [3, 3], [1, 4, 1])
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.plus(_t(x), _t(y))
    assert e.value.status == 1


def _check_round(z, ref, tol):
    assert z.r.tolist() == ref.r.tolist()
    err = O.tt_rel_diff(O.TT(z.n, z.r, z.core), ref)
    assert err <= tol, err
    return err


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_plus_round_fused_sweeps(gpu_ctx, monkeypatch, prec, tol):
    """X + 0.5 X (every bond truncates to half) and X + Y (nothing
    truncates) through the sweeps with the sum generated inside sweep 1."""
    monkeypatch.setenv("TTG_GRAM", "0")
    monkeypatch.setenv("TTG_SMALL_PERSIST", "0")
    rng = np.random.default_rng(103)
    n = [4] * 9
    x = O.rand_tt(rng, n, [1, 4, 16, 30, 30, 30, 30, 16, 4, 1])
    y = O.rand_tt(rng, n, [1, 4, 12, 20, 24, 20, 12, 8, 4, 1])
    for other, a, b, eps in ((x, 1.0, 0.5, 1e-10), (y, 1.0, 1.0, 1e-8), (y, -1.5, 0.25, 1e-3)):
        info = {}
        z = ttgpu.plus_round(_t(x), _t(other), eps, a, b, precision=prec, info=info)
        ref = O.tt_round(O.tt_plus(O.tt_scal(a, x), O.tt_scal(b, other)), eps)
        err = _check_round(z, ref, tol)
        assert info["path"] == 0
        print(f"a={a} b={b} eps={eps}: ranks {z.r.tolist()} err {err:.2e}", flush=True)


def test_plus_round_config1_full_size(gpu_ctx):
    """The X + 0.5 X workload (d=20, n=4, rank 20 -> rank 40, eps 1e-10)
    from its two summands: the oracle's ranks (those of X) and <= 1e-12."""
    c = W.round1()
    n, r, core = c["x"]
    x = O.TT(n, r, core)
    z = ttgpu.plus_round(_t(x), _t(x), c["eps"], 1.0, 0.5)
    ref = O.tt_round(O.TT(*c["z"]), c["eps"])
    _check_round(z, ref, 1e-12)
    assert z.r.tolist() == O.generic_ranks(n, r)


def test_plus_round_batch_tensors(gpu_ctx):
    """x + x for eight tensors of the 512-tensor recipe (ranks halve back)."""
    bb = W.batch4(count=8, seed=105)
    for t in bb["tensors"]:
        zz = O.TT(*t)
        x = O.TT.from_cores([np.asfortranarray(c) for c in _halves(zz)])
        z = ttgpu.plus_round(_t(x), _t(x), bb["eps"])
        ref = O.tt_round(zz, bb["eps"])
        _check_round(z, ref, 1e-12)


def _halves(z):
    d = z.d
    rx = [1] + [int(v) // 2 for v in z.r[1:-1]] + [1]
    return [z.core_k(k)[:rx[k], :, :rx[k + 1]] for k in range(d)]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_plus_round_gram_form(gpu_ctx, monkeypatch, prec, tol):
    """The mixed-vs-FP64 recipe at reduced size from its two summands: the
    assembled sum takes the Gram-chain form (no bond truncates; the plain
    form's one-workgroup path for ranks <= 64 is switched off)."""
    monkeypatch.setenv("TTG_SMALL_PERSIST", "0")
    rng = np.random.default_rng(106)
    d, nn, rank = 10, 8, 24
    xa = W.rescale_right_unfoldings(W.rand_cores(rng, [nn] * d, [1] + [rank] * (d - 1) + [1]))
    xb = W.rescale_right_unfoldings(W.rand_cores(rng, [nn] * d, [1] + [rank] * (d - 1) + [1]))
    a, b = O.TT.from_cores(xa), O.TT.from_cores(xb)
    info = {}
    z = ttgpu.plus_round(_t(a), _t(b), 1e-6, precision=prec, info=info)
    assert info["path"] == 1
    _check_round(z, O.tt_round(O.tt_plus(a, b), 1e-6), max(tol, 1e-10) if prec == "fp32" else tol)


def test_plus_round_small_tensor_form(gpu_ctx):
    rng = np.random.default_rng(107)
    n = [3, 2, 4, 3]
    x = O.rand_tt(rng, n, [1, 3, 5, 4, 1])
    y = O.rand_tt(rng, n, [1, 2, 4, 3, 1])
    z = ttgpu.plus_round(_t(x), _t(y), 1e-12, 0.5, 2.0)
    _check_round(z, O.tt_round(O.tt_plus(O.tt_scal(0.5, x), O.tt_scal(2.0, y)), 1e-12), 1e-12)


def test_norm_matches_oracle_and_dense(gpu_ctx):
    rng = np.random.default_rng(108)
    n = [3, 4, 2, 3, 2]
    x = O.rand_tt(rng, n, [1, 3, 8, 5, 4, 1])
    full = O.tt_full(x)
    v = ttgpu.norm(_t(x))
    assert abs(v - np.linalg.norm(full)) <= 1e-13 * np.linalg.norm(full)
    assert abs(v - O.tt_norm(x)) <= 1e-13 * v
    zero = O.TT(n, x.r, np.zeros_like(x.core))
    assert ttgpu.norm(_t(zero)) == 0.0
    one = O.rand_tt(rng, [9], [1, 1])
    assert abs(ttgpu.norm(_t(one)) - np.linalg.norm(one.core)) <= 1e-15 * np.linalg.norm(one.core)


def test_difference_norm_keeps_its_digits(gpu_ctx):
    """||(X + delta W) - X|| = delta ||W||: the Gram identity
    ||A||^2 + ||B||^2 - 2<A,B> would lose every digit below sqrt(u)."""
    rng = np.random.default_rng(109)
    n = [4] * 8
    x = O.rand_tt(rng, n, [1, 4, 12, 16, 16, 16, 12, 4, 1])
    w = O.rand_tt(rng, n, [1, 3, 5, 6, 6, 6, 5, 3, 1])
    for delta in (1e-4, 1e-9):
        y = O.tt_plus(x, O.tt_scal(delta, w))
        nx = O.tt_norm(x)
        got = ttgpu.plus_norm_dev(gpu_ctx, n, y.r, _dev(y), x.r, _dev(x))
        want = delta * O.tt_norm(w)
        assert abs(got - want) <= 1e-14 * nx + 1e-12 * want, (delta, got, want)
        assert abs(got - O.tt_norm(O.tt_minus(y, x))) <= 1e-14 * nx


def test_dot_matches_oracle_and_dense(gpu_ctx):
    rng = np.random.default_rng(110)
    n = [3, 2, 4, 3]
    x = O.rand_tt(rng, n, [1, 3, 6, 4, 1])
    y = O.rand_tt(rng, n, [1, 2, 5, 3, 1])
    v = ttgpu.dot(_t(x), _t(y))
    dense = O.tt_full(x) @ O.tt_full(y)
    scale = np.linalg.norm(O.tt_full(x)) * np.linalg.norm(O.tt_full(y))
    assert abs(v - dense) <= 1e-14 * scale
    assert abs(v - O.tt_dot(x, y)) <= 1e-14 * scale
    nx = ttgpu.norm(_t(x))
    assert abs(ttgpu.dot(_t(x), _t(x)) - nx * nx) <= 1e-13 * nx * nx
    with pytest.raises(ttgpu.TTGError):
        bad = O.rand_tt(rng, n, [2, 3, 6, 4, 1])
        ttgpu.dot(_t(bad), _t(bad))


def test_full_size_norm_and_difference(gpu_ctx):
    """The d=32, n=8, r=256 tensor: GPU norm vs the oracle; the GPU's
    ||Y - Y_ref|| / ||Y_ref|| of the GPU round against the oracle's value."""
    c = O.make_config(3)
    x = c["x"]
    nx = ttgpu.norm(_t(x))
    ref_n = O.tt_norm(x)
    assert abs(nx - ref_n) <= 1e-12 * ref_n
    y = ttgpu.round(_t(x), c["eps"])
    ref = O.tt_round(x, c["eps"])
    got = ttgpu.rel_diff(y, _t(ref))
    want = O.tt_rel_diff(O.TT(y.n, y.r, y.core), ref)
    print(f"norm {nx:.6e}; rel diff GPU {got:.3e} oracle {want:.3e}", flush=True)
    assert got <= 1e-12 and abs(got - want) <= 1e-14
    dv = ttgpu.dot(_t(x), _t(x))
    assert abs(dv - ref_n ** 2) <= 1e-11 * ref_n ** 2
