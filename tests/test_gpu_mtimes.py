"""GPU parity of A*x (@tt_matrix/mtimes.m:38-65) and of the fused round(A*x)."""
import numpy as np
import pytest

import ttgpu
from oracle import tt_oracle as O

pytestmark = pytest.mark.gpu


def rand_ttm(rng, n, m, r):
    d = len(n)
    cores = [rng.standard_normal((r[k], n[k], m[k], r[k + 1])) for k in range(d)]
    return O.TTM.from_cores(cores)


def test_mtimes_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(1)
    n, m = [2, 3, 4, 2], [3, 2, 2, 5]
    a = rand_ttm(rng, n, m, [1, 3, 2, 4, 1])
    x = O.rand_tt(rng, m, [1, 2, 5, 3, 1])
    ref = O.ttm_mtimes(a, x)
    y = ttgpu.mtimes(ttgpu.TTMatrix(a.n, a.m, a.tt.r, a.tt.core), ttgpu.TTTensor(x.n, x.r, x.core))
    assert y.r.tolist() == ref.r.tolist()
    np.testing.assert_allclose(y.core, ref.core, rtol=1e-13, atol=1e-14)


def test_fused_round_matches_oracle(gpu_ctx):
    c = O.make_config(2, d=6, rank=16, seed=7)
    a, x = c["a"], c["x"]
    ref = O.tt_round(O.ttm_mtimes(a, x), c["eps"])
    info = {}
    y = ttgpu.mtimes_round(ttgpu.TTMatrix(a.n, a.m, a.tt.r, a.tt.core), ttgpu.TTTensor(x.n, x.r, x.core),
                           c["eps"], info=info)
    print("fused ranks", y.r.tolist(), info, flush=True)
    assert y.r.tolist() == ref.r.tolist()
    assert O.tt_rel_diff(O.TT(y.n, y.r, y.core), ref) <= 1e-12


def test_fused_round_config2_full_size(gpu_ctx):
    """Config 2 at its real size: d=12, A rank 5 (identity Kronecker + 0.1 N(0,1)),
    x rank 64, product rank 320, eps 1e-6 — ranks identical to the oracle's
    mtimes-then-round and relative difference <= 1e-12, plain form."""
    c = O.make_config(2)
    a, x = c["a"], c["x"]
    ref = O.tt_round(O.ttm_mtimes(a, x), c["eps"])
    A = ttgpu.TTMatrix(a.n, a.m, a.tt.r, a.tt.core)
    X = ttgpu.TTTensor(x.n, x.r, x.core)
    info = {}
    y = ttgpu.mtimes_round(A, X, c["eps"], info=info)
    print("config 2 ranks", y.r.tolist(), info, flush=True)
    assert y.r.tolist() == ref.r.tolist()
    # no bond truncates below its cap: every bond (wide and right-capped tall
    # ones alike) is certified and completed without an SVD
    assert info["svd_skipped"] == len(a.n) - 1 and info["jacobi_sweeps"] == 0, info
    assert O.tt_rel_diff(O.TT(y.n, y.r, y.core), ref) <= 1e-12
    # unfused path: A*x on the GPU, then round on the GPU
    p = ttgpu.mtimes(A, X)
    assert p.r.tolist() == (a.tt.r * x.r).tolist()
    y2 = ttgpu.round(p, c["eps"])
    assert y2.r.tolist() == ref.r.tolist()
    assert O.tt_rel_diff(O.TT(y2.n, y2.r, y2.core), ref) <= 1e-12


@pytest.mark.parametrize("m", [[1, 2, 3, 4], [5, 6, 7, 9], [8, 8, 2, 1]])
def test_mtimes_mode_sizes(gpu_ctx, m):
    """Every mode-size specialisation of the product-core kernel (m = 1..4 and
    8 unrolled, others generic), element by element against the oracle."""
    rng = np.random.default_rng(sum(m))
    n = [3, 2, 4, 2]
    a = rand_ttm(rng, n, m, [1, 3, 5, 2, 1])
    x = O.rand_tt(rng, m, [1, 4, 9, 3, 1])
    ref = O.ttm_mtimes(a, x)
    y = ttgpu.mtimes(ttgpu.TTMatrix(a.n, a.m, a.tt.r, a.tt.core), ttgpu.TTTensor(x.n, x.r, x.core))
    assert y.r.tolist() == ref.r.tolist()
    np.testing.assert_allclose(y.core, ref.core, rtol=1e-13, atol=1e-14)


def test_mtimes_config2_elementwise(gpu_ctx):
    """Config 2's A*x (d=12, product rank 320, 32.8 MB of cores) element by
    element against the oracle."""
    import torch
    c = O.make_config(2)
    a, x = c["a"], c["x"]
    ref = O.ttm_mtimes(a, x)
    dev = torch.device("cuda:0")
    ctx = ttgpu.torch_context(0)
    ry, y = ttgpu.mtimes_dev(ctx, a.n, a.m, a.tt.r, torch.from_numpy(a.tt.core).to(dev), x.r,
                             torch.from_numpy(x.core).to(dev))
    assert ry.tolist() == ref.r.tolist()
    np.testing.assert_allclose(y.cpu().numpy(), ref.core, rtol=1e-13, atol=1e-14)


def test_back_to_back_device_calls(gpu_ctx):
    """Device-resident calls return before their work ends; two A*x calls of
    different shapes enqueued back to back (no host synchronisation between
    them) must each use their own core table (the staging buffer is reused
    only after the earlier copy out of it has finished)."""
    import torch
    rng = np.random.default_rng(11)
    dev = torch.device("cuda:0")
    ctx = ttgpu.torch_context(0)
    cases = []
    for n, m, ra, rx in [([2, 3, 4], [3, 2, 2], [1, 3, 2, 1], [1, 2, 5, 1]),
                         ([4, 2, 2, 3, 2], [2, 2, 3, 1, 4], [1, 2, 4, 3, 2, 1], [1, 3, 2, 6, 2, 1])]:
        a = rand_ttm(rng, n, m, ra)
        x = O.rand_tt(rng, m, rx)
        cases.append((a, x, torch.from_numpy(a.tt.core).to(dev), torch.from_numpy(x.core).to(dev)))
    torch.cuda.synchronize()
    outs = []
    for _ in range(3):
        for a, x, da, dx in cases:
            outs.append((a, x) + ttgpu.mtimes_dev(ctx, a.n, a.m, a.tt.r, da, x.r, dx))
    torch.cuda.synchronize()
    for a, x, ry, y in outs:
        ref = O.ttm_mtimes(a, x)
        assert ry.tolist() == ref.r.tolist()
        np.testing.assert_allclose(y.cpu().numpy(), ref.core, rtol=1e-13, atol=1e-14)
