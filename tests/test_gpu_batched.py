"""Batched rounding of uneven TT tensors (config-4 recipe at reduced sizes)
against the oracle, tensor by tensor: identical ranks and relative Frobenius
difference <= 1e-12 (plain form)."""
import numpy as np
import pytest
import torch

import ttgpu
from oracle import tt_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count,dmax,rmax", [(6, 6, 8), (24, 10, 24)])
def test_batched_matches_oracle(gpu_ctx, count, dmax, rmax):
    batch = O.make_batch(count=count, seed=40 + count, dmin=3, dmax=dmax, nset=(2, 4, 8), rmin=2, rmax=rmax)
    eps = 1e-8
    cores = [t.core for t in batch]
    cofs = np.concatenate([[0], np.cumsum([c.size for c in cores])])
    dcore = torch.tensor(np.concatenate(cores), device="cuda:0")
    res = ttgpu.round_batched_dev(gpu_ctx, [t.n for t in batch], [t.r for t in batch], dcore,
                                  cofs[:-1], eps, precision="fp64")
    assert len(res) == count
    for t, (r_t, y_t) in zip(batch, res):
        ref = O.tt_round(t, eps)
        assert list(map(int, r_t)) == ref.r.tolist()
        y = O.TT(t.n, r_t, y_t.cpu().numpy())
        assert O.tt_rel_diff(y, ref) <= 1e-12


def _check_batch(batch, res, eps, tol=1e-12):
    assert len(res) == len(batch)
    for i, (t, (r_t, y_t)) in enumerate(zip(batch, res)):
        ref = O.tt_round(t, eps)
        assert list(map(int, r_t)) == ref.r.tolist(), f"tensor {i}"
        if len(t.n) == 1:  # rank concatenation (the metric's minus) needs d >= 2
            np.testing.assert_allclose(y_t.cpu().numpy(), ref.core, rtol=1e-14, atol=0)
            continue
        y = O.TT(t.n, r_t, y_t.cpu().numpy())
        assert O.tt_rel_diff(y, ref) <= tol, f"tensor {i}"


def _run(ctx, batch, eps, precision="fp64", info=None):
    cores = [t.core for t in batch]
    cofs = np.concatenate([[0], np.cumsum([c.size for c in cores])])
    dcore = torch.tensor(np.concatenate(cores), device="cuda:0")
    torch.cuda.synchronize()
    return ttgpu.round_batched_dev(ctx, [t.n for t in batch], [t.r for t in batch], dcore, cofs[:-1], eps,
                                   precision=precision, info=info)


def test_batched_tuner_all_modes():
    """96 tensors on a fresh context: the first call tunes on four interleaved
    quarters (three persistent launch shapes and the stream workers), so every
    execution mode rounds a quarter of the batch; the second call runs the
    chosen mode (a replayed launch when it is a persistent one)."""
    ctx = ttgpu.Context(0)
    batch = O.make_batch(count=96, seed=7, dmin=2, dmax=7, nset=(2, 3, 4, 8), rmin=1, rmax=12)
    eps = 1e-8
    info = {}
    _check_batch(batch, _run(ctx, batch, eps, info=info), eps)
    _check_batch(batch, _run(ctx, batch, eps, info=info), eps)
    _check_batch(batch, _run(ctx, batch, eps), eps)


def test_batched_edge_cases():
    """d = 1, a zero tensor, eps = 0 (no truncation), a rank above r_i n_i
    (QR cap), a tall last SVD and a bond of rank 110 (its Jacobi matrix does
    not fit in LDS), in one persistent launch."""
    rng = np.random.default_rng(11)
    t1 = O.rand_tt(rng, [7], [1, 1])
    tz = O.TT([3, 4, 2], [1, 3, 2, 1], np.zeros(3 * 3 + 3 * 4 * 2 + 2 * 2))
    tcap = O.rand_tt(rng, [2, 3, 2, 3], [1, 5, 9, 4, 1])
    ttall = O.rand_tt(rng, [4, 2, 2], [1, 4, 8, 1])
    x = O.rand_tt(rng, [3, 3, 3, 3], [1, 3, 5, 3, 1])
    tsum = O.tt_plus(x, x)
    tbig = O.rand_tt(rng, [8, 16, 16, 8], [1, 8, 110, 8, 1])
    batch = [t1, tz, tcap, ttall, tsum, tbig]
    _check_batch(batch, _run(ttgpu.default_context(0), batch, 1e-10), 1e-10)
    _check_batch(batch, _run(ttgpu.default_context(0), batch, 0.0), 0.0)


def test_batched_nonfinite_is_an_error(gpu_ctx):
    rng = np.random.default_rng(12)
    good = O.rand_tt(rng, [3, 3, 3], [1, 2, 2, 1])
    bad = O.rand_tt(rng, [3, 3, 3], [1, 2, 2, 1])
    bad.core[5] = np.nan
    with pytest.raises(ttgpu.TTGError) as e:
        _run(gpu_ctx, [good, bad], 1e-8)
    assert "NONFINITE" in str(e.value)


def test_batched_mixed_precision_uses_streams(gpu_ctx):
    batch = O.make_batch(count=8, seed=13, dmin=3, dmax=6, nset=(2, 4), rmin=2, rmax=10)
    eps = 1e-8
    res = _run(gpu_ctx, batch, eps, precision="fp32")
    for t, (r_t, y_t) in zip(batch, res):
        ref = O.tt_round(t, eps)
        y = O.TT(t.n, r_t, y_t.cpu().numpy())
        assert O.tt_rel_diff(y, ref) <= 1e-6


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_batched_rmax(gpu_ctx, prec, tol):
    """rmax per tensor (round.m:78): tensors whose cap binds go to the stream
    workers, the rest keep the batched launch; every result equals
    round(X{t}, eps, rmax{t}) of the oracle (ranks identical)."""
    batch = O.make_batch(count=40, seed=91, dmin=3, dmax=8, nset=(2, 4, 8), rmin=4, rmax=20)
    eps = 1e-10
    caps = [3 if i % 3 == 0 else (100 if i % 3 == 1 else 6) for i in range(len(batch))]
    rms = [np.full(t.d + 1, c, dtype=np.int64) for t, c in zip(batch, caps)]
    cores = [t.core for t in batch]
    cofs = np.concatenate([[0], np.cumsum([c.size for c in cores])])
    dcore = torch.tensor(np.concatenate(cores), device="cuda:0")
    torch.cuda.synchronize()
    res = ttgpu.round_batched_dev(gpu_ctx, [t.n for t in batch], [t.r for t in batch], dcore, cofs[:-1], eps,
                                  precision=prec, rmax=rms)
    binding = 0
    for i, (t, (r_t, y_t)) in enumerate(zip(batch, res)):
        ref = O.tt_round(t, eps, rms[i])
        assert list(map(int, r_t)) == ref.r.tolist(), f"tensor {i}"
        assert max(r_t) <= caps[i]
        binding += int(max(O.tt_round(t, eps).r) > caps[i])
        y = O.TT(t.n, r_t, y_t.cpu().numpy())
        assert O.tt_rel_diff(y, ref) <= tol, f"tensor {i}"
    assert binding >= 10
    # a scalar cap for the whole batch
    res2 = ttgpu.round_batched_dev(gpu_ctx, [t.n for t in batch], [t.r for t in batch], dcore, cofs[:-1], eps,
                                   precision=prec, rmax=5)
    for i, (t, (r_t, y_t)) in enumerate(zip(batch, res2)):
        assert list(map(int, r_t)) == O.tt_round(t, eps, 5).r.tolist(), f"tensor {i}"
