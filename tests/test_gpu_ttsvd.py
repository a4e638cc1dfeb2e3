"""TT-SVD (tt_tensor(full, eps), @tt_tensor/tt_tensor.m:175-244) on the GPU.

Against the oracle's restatement (tests/test_oracle.py pins it to a dense
left-to-right projection): identical ranks and a dense difference
<= 1e-12 ||b||, plus the bound ||b - full(TT)|| <= eps ||b||.  Cases cover
wide and tall unfoldings, numerically low-rank unfoldings (Householder
fallback), a 16.7M-entry tensor, the zero tensor, d = 1, NaN input, the
512-row limit and an rmax cap.
"""
import numpy as np
import pytest
import torch

import ttgpu
from oracle import tt_oracle as O

pytestmark = pytest.mark.gpu


def _check(full, n, eps, tol=1e-12, info=None):
    t = ttgpu.tt_tensor_from_full(full, eps, n, info=info)
    ref = O.tt_from_full(full, eps, n)
    assert t.r.tolist() == ref.r.tolist()
    nb = np.linalg.norm(full)
    got = O.tt_full(O.TT(t.n, t.r, t.core))
    assert np.linalg.norm(got - O.tt_full(ref)) <= tol * nb
    assert np.linalg.norm(got - full) <= eps * nb * (1 + 1e-10) + 1e-14 * nb
    return t


@pytest.mark.parametrize("seed,n,r,eps", [
    (31, [3, 4, 2, 3, 2], [1, 3, 5, 4, 2, 1], 1e-12),
    (32, [2, 3, 4, 3], [1, 2, 6, 3, 1], 0.05),
    (33, [4, 4, 4], [1, 4, 4, 1], 0.2),
    (34, [2, 3, 4, 3], [1, 2, 6, 3, 1], 0.26),
    (35, [4, 4, 4, 2], [1, 4, 16, 2, 1], 1e-14),     # tall last unfolding (64 x 2)
])
def test_ttsvd_matches_oracle(gpu_ctx, seed, n, r, eps):
    rng = np.random.default_rng(seed)
    x = O.rand_tt(rng, n, r)
    x = O.tt_plus(x, O.tt_scal(0.1, O.rand_tt(rng, n, r)))
    t = _check(O.tt_full(x), n, eps)
    print(f"n={n} eps={eps}: ranks {t.r.tolist()}", flush=True)


def test_ttsvd_low_rank_unfoldings_take_householder(gpu_ctx):
    rng = np.random.default_rng(36)
    n = [4] * 8
    x = O.rand_tt(rng, n, [1, 3, 3, 3, 3, 3, 3, 3, 1])
    info = {}
    t = _check(O.tt_full(x), n, 1e-10, info=info)
    assert t.r.tolist() == [1, 3, 3, 3, 3, 3, 3, 3, 1]
    assert info["qr_fallbacks"] >= 1
    print("info", info, flush=True)


def test_ttsvd_16m_entries(gpu_ctx):
    """d=12, n=4 (16.7M entries): a rank-6 TT plus 1e-9 dense noise, eps 1e-6."""
    rng = np.random.default_rng(37)
    n = [4] * 12
    x = O.rand_tt(rng, n, [1, 4] + [6] * 9 + [4, 1])
    full = O.tt_full(x)
    full = full + 1e-9 * np.linalg.norm(full) / np.sqrt(full.size) * rng.standard_normal(full.size)
    t = _check(full, n, 1e-6, tol=1e-11)
    assert t.r.tolist() == [1, 4] + [6] * 9 + [4, 1]


def test_ttsvd_edge_cases(gpu_ctx):
    n = [3, 4, 5]
    # zero tensor: rank 1 everywhere (my_chop2 of a zero vector is 1); the
    # singular vectors of a zero unfolding are arbitrary unit vectors (as
    # svd's), the value is zero
    z = ttgpu.tt_tensor_from_full(np.zeros(60), 1e-8, n)
    assert z.r.tolist() == [1, 1, 1, 1] == O.tt_from_full(np.zeros(60), 1e-8, n).r.tolist()
    assert not np.any(O.tt_full(O.TT(z.n, z.r, z.core)))
    v = np.random.default_rng(38).standard_normal(17)
    one = ttgpu.tt_tensor_from_full(v, 1e-8, [17])
    assert one.r.tolist() == [1, 1] and one.core.tobytes() == v.tobytes()
    bad = np.ones(60)
    bad[7] = np.nan
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.tt_tensor_from_full(bad, 1e-8, n)
    assert e.value.status == 3
    big = np.random.default_rng(39).standard_normal(32 ** 4)
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.tt_tensor_from_full(big, 1e-8, [32] * 4)  # 1024 x 1024 unfolding
    assert e.value.status == 1


def test_ttsvd_rmax_cap(gpu_ctx):
    rng = np.random.default_rng(40)
    n = [4] * 6
    x = O.rand_tt(rng, n, [1, 4, 10, 12, 10, 4, 1])
    full = O.tt_full(x)
    t = ttgpu.tt_tensor_from_full(full, 1e-12, n, rmax=5)
    assert max(t.r.tolist()) <= 5
    # the capped result is a quasi-optimal rank-5 approximation: within the
    # sum of the discarded tails (each step's ||s(6:end)||) of full
    err = np.linalg.norm(O.tt_full(O.TT(t.n, t.r, t.core)) - full)
    assert err < np.linalg.norm(full)
