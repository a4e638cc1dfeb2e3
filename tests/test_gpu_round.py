"""GPU parity of round(X, eps, rmax) against the FP64 oracle (restatement of @tt_tensor/round.m).

Bar (plain FP64 form): identical ranks and ||Y_gpu - Y_oracle|| / ||Y_oracle|| <= 1e-12,
measured as the norm of a TT difference (@tt_tensor/minus.m:18 + norm.m:25-41).
Mixed form: relative error <= 1e-6 against the FP64 oracle.
"""
import numpy as np
import pytest

import ttgpu
from oracle import tt_oracle as O

pytestmark = pytest.mark.gpu


def to_gpu_tt(x):
    return ttgpu.TTTensor(x.n, x.r, x.core)


def to_oracle(y):
    return O.TT(y.n, y.r, y.core)


def check_plain(x, eps, rmax=None, tol=1e-12):
    ref = O.tt_round(x, eps, rmax)
    info = {}
    y = ttgpu.round(to_gpu_tt(x), eps, rmax, "fp64", info=info)
    print("ranks", y.r.tolist(), "info", info, flush=True)
    assert y.r.tolist() == ref.r.tolist()
    err = O.tt_rel_diff(to_oracle(y), ref)
    assert err <= tol, err
    return y, ref, info


def test_config1_closed_form(gpu_ctx):
    c = O.make_config(1)
    y, ref, info = check_plain(c["z"], c["eps"])
    assert y.r.tolist() == [1, 4, 16] + [20] * 15 + [16, 4, 1]
    assert O.tt_rel_diff(to_oracle(y), O.tt_scal(1.5, c["x"])) <= 1e-10


@pytest.mark.parametrize("seed,d,n,r,eps", [
    (11, 4, 3, 5, 1e-8), (12, 6, 2, 8, 1e-3), (13, 5, 4, 12, 1e-6), (14, 3, 7, 30, 0.0),
    (15, 8, 3, 17, 1e-2), (16, 2, 40, 33, 1e-10),
])
def test_random_plain(gpu_ctx, seed, d, n, r, eps):
    rng = np.random.default_rng(seed)
    ranks = [1] + [int(v) for v in rng.integers(max(1, r // 2), r + 1, size=d - 1)] + [1]
    x = O.rand_tt(rng, [n] * d, ranks)
    check_plain(x, eps)


def test_graded_truncation(gpu_ctx):
    """Truncation actually happens: graded bond spectra, eps inside the spectrum."""
    c = O.make_config(3, d=8, n=4, rank=24, seed=21)
    check_plain(c["x"], 1e-3)


def test_rmax_cap(gpu_ctx):
    rng = np.random.default_rng(5)
    x = O.rand_tt(rng, [4] * 6, [1, 4, 16, 20, 16, 4, 1])
    check_plain(x, 1e-14, rmax=7)


def test_zero_tensor(gpu_ctx):
    x = O.TT([3] * 4, [1, 3, 5, 3, 1], np.zeros(3 * 3 + 3 * 3 * 5 + 5 * 3 * 3 + 3 * 3))
    y = ttgpu.round(to_gpu_tt(x), 1e-8)
    assert y.r.tolist() == [1, 1, 1, 1, 1]
    assert np.all(y.core == 0)


def test_d1(gpu_ctx):
    rng = np.random.default_rng(3)
    x = O.rand_tt(rng, [7], [1, 1])
    y = ttgpu.round(to_gpu_tt(x), 1e-8)
    assert y.r.tolist() == [1, 1]
    np.testing.assert_allclose(y.core, x.core, rtol=1e-14)


def test_nonfinite_rejected(gpu_ctx):
    rng = np.random.default_rng(4)
    x = O.rand_tt(rng, [3] * 3, [1, 2, 2, 1])
    core = x.core.copy()
    core[3] = np.nan
    with pytest.raises(ttgpu.TTGError) as e:
        ttgpu.round(ttgpu.TTTensor(x.n, x.r, core), 1e-8)
    assert e.value.status == 3


def test_matrix_round(gpu_ctx):
    rng = np.random.default_rng(8)
    d = 5
    cores = [rng.standard_normal((1 if k == 0 else 6, 3, 2, 1 if k == d - 1 else 6)) for k in range(d)]
    a = O.TTM.from_cores(cores)
    a2 = O.TTM(a.n, a.m, O.tt_plus(a.tt, a.tt))
    ref = O.ttm_round(a2, 1e-10)
    g = ttgpu.TTMatrix(a2.n, a2.m, a2.tt.r, a2.tt.core).round(1e-10)
    assert g.r.tolist() == ref.tt.r.tolist()
    assert O.tt_rel_diff(O.TT(g.n * g.m, g.r, g.core), ref.tt) <= 1e-12


def test_mixed_small(gpu_ctx):
    c = O.make_config(3, d=10, n=8, rank=32, seed=31)
    x = c["x"]
    ref = O.tt_round(x, 1e-6)
    info = {}
    y = ttgpu.round(to_gpu_tt(x), 1e-6, precision="fp32", info=info)
    err = O.tt_rel_diff(to_oracle(y), ref)
    print("mixed ranks", y.r.tolist(), "ref", ref.r.tolist(), "err", err, info, flush=True)
    assert y.r.tolist() == ref.r.tolist()
    assert err <= 1e-6


def test_cooperative_jacobi_path(gpu_ctx, monkeypatch):
    """Rank 128 bond: the SVD runs on several cooperating workgroups (the
    sweeps; the Gram form is turned off)."""
    monkeypatch.setenv("TTG_GRAM", "0")
    rng = np.random.default_rng(9)
    x = O.rand_tt(rng, [16, 16, 16, 16], [1, 16, 128, 16, 1])
    check_plain(x, 1e-9)


@pytest.mark.parametrize("eps", [1e-6, 0.0])
def test_no_truncation_shortcut(gpu_ctx, monkeypatch, eps):
    """Bonds whose smallest singular value provably clears my_chop2's threshold
    are completed from their QR factors without an SVD; the result must equal
    the full SVD path's and the oracle's (ranks exact, 1e-12 relative)."""
    monkeypatch.setenv("TTG_GRAM", "0")  # the shortcut of the sweeps, not the Gram form
    rng = np.random.default_rng(21)
    x = O.rand_tt(rng, [4] * 8, [1, 4, 16, 64, 80, 80, 16, 4, 1])  # ranks > 64: multi-CU path
    ref = O.tt_round(x, eps)
    info = {}
    y = ttgpu.round(to_gpu_tt(x), eps, info=info)
    assert info["svd_skipped"] > 0
    monkeypatch.setenv("TTG_SVD_ALWAYS", "1")
    info2 = {}
    y2 = ttgpu.round(to_gpu_tt(x), eps, info=info2)
    assert info2["svd_skipped"] == 0
    for yy in (y, y2):
        assert yy.r.tolist() == ref.r.tolist()
        assert O.tt_rel_diff(O.TT(yy.n, yy.r, yy.core), ref) <= 1e-12


def test_truncating_bonds_keep_the_svd(gpu_ctx):
    rng = np.random.default_rng(22)
    x = O.rand_tt(rng, [3] * 6, [1, 3, 9, 12, 9, 3, 1])
    z = O.tt_plus(x, x)  # every bond truncates to half
    info = {}
    y = ttgpu.round(to_gpu_tt(z), 1e-10, info=info)
    ref = O.tt_round(z, 1e-10)
    assert y.r.tolist() == ref.r.tolist()
    assert O.tt_rel_diff(O.TT(y.n, y.r, y.core), ref) <= 1e-12
