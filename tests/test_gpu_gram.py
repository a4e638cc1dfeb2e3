"""The Gram-chain form of round (mixed precision mode) against the FP64 oracle.

The form applies when every bond provably keeps all its singular values; its
result is then X in right-orthogonal, balanced form with the capped ranks of
@tt_tensor/round.m:48 and the sweep-2 SVD sizes.  Bars: identical ranks,
relative Frobenius difference <= 1e-10 from the oracle (the form itself is
accurate to O(u sum kappa); the mode's budget is 1e-6), balancing invariant
<= 1e-6 (C_k C_k^T = nrm0^2 I, round.m:90-107).  Tensors where a bond
truncates, a binding rmax, a zero tensor or non-finite input must be handed
to the sweeps (info path 0) and still match the oracle.
"""
import numpy as np
import pytest

import ttgpu
from oracle import tt_oracle as O

pytestmark = pytest.mark.gpu


def _mixed(x, eps, rmax=None):
    info = {}
    y = ttgpu.round(ttgpu.TTTensor(x.n, x.r, x.core), eps, rmax, precision="fp32", info=info)
    return y, info


def _check(x, eps, path, tol=1e-10, rmax=None):
    """path: 1 Gram form, 0 sweeps, None either (a bond whose smallest
    singular value is near the Gram rounding floor may go either way)."""
    ref = O.tt_round(x, eps, rmax)
    y, info = _mixed(x, eps, rmax)
    print("ranks", y.r.tolist(), "info", info, flush=True)
    if path is not None:
        assert info["path"] == path
    if info["path"] == 0:
        tol = max(tol, 1e-6)
    assert y.r.tolist() == ref.r.tolist()
    yo = O.TT(y.n, y.r, y.core)
    err = O.tt_rel_diff(yo, ref)
    assert err <= tol, err
    if y.d > 1 and np.any(y.core):
        _, bres = O.balance_residual(yo)
        assert bres <= 1e-6, bres
    return y, info, err


@pytest.mark.parametrize("seed,n,r,path", [
    (61, [4] * 8, [1, 4, 16, 64, 80, 80, 16, 4, 1], None),      # right cap; bond 5 sigma_min/||X|| = 4e-6
    (62, [3, 5, 2, 7, 4, 2], [1, 3, 15, 20, 28, 8, 1], 1),       # uneven modes, right caps inside
    (63, [2] * 10, [1, 2, 4, 8, 16, 32, 16, 8, 4, 2, 1], 0),    # bond 5 has sigma_min/||X|| = 4e-12: truncates
    (64, [6, 6, 6, 6], [1, 6, 36, 6, 1], 1),                      # every bond at its cap
    (65, [4, 4, 4, 4, 4], [1, 300, 200, 70, 9, 1], 1),             # input ranks above the caps (> 256)
    (66, [8] * 6, [1, 8, 64, 256, 64, 8, 1], 1),                   # rank 256 bond
])
def test_gram_form_matches_oracle(gpu_ctx, seed, n, r, path):
    rng = np.random.default_rng(seed)
    x = O.rand_tt(rng, n, r)
    _check(x, 1e-8, path=path)


def test_gram_form_eps_zero(gpu_ctx):
    rng = np.random.default_rng(67)
    x = O.rand_tt(rng, [4] * 6, [1, 4, 16, 30, 16, 4, 1])
    _check(x, 0.0, path=1)


def test_gram_form_graded_spectra(gpu_ctx):
    """Mixed-vs-FP64 recipe at reduced size (graded core spectra): proven."""
    c = O.make_config(3, d=10, n=8, rank=32, seed=68)
    _check(c["x"], 1e-6, path=1)


def test_truncating_tensor_goes_to_the_sweeps(gpu_ctx):
    rng = np.random.default_rng(69)
    x = O.rand_tt(rng, [3] * 6, [1, 3, 9, 12, 9, 3, 1])
    z = O.tt_plus(x, x)  # every bond truncates to half: not provable
    _check(z, 1e-10, path=0, tol=1e-6)


def test_graded_spectra_threshold_crossing(gpu_ctx):
    """Graded bonds (smallest sigma/||X|| = 1.3e-2): eps = 1e-3 keeps every
    value (threshold 3.8e-4, proven: Gram form), eps = 0.1 truncates
    (threshold 3.8e-2: sweeps)."""
    c = O.make_config(3, d=8, n=4, rank=24, seed=21)
    _check(c["x"], 1e-3, path=1)
    _check(c["x"], 0.1, path=0, tol=1e-6)


def test_binding_rmax_goes_to_the_sweeps(gpu_ctx):
    rng = np.random.default_rng(70)
    x = O.rand_tt(rng, [4] * 6, [1, 4, 16, 20, 16, 4, 1])
    _check(x, 1e-14, path=0, tol=1e-6, rmax=7)


def test_zero_and_nonfinite_go_to_the_sweeps(gpu_ctx):
    x = O.TT([3] * 4, [1, 3, 5, 3, 1], np.zeros(3 * 3 + 3 * 3 * 5 + 5 * 3 * 3 + 3 * 3))
    y, info = _mixed(x, 1e-8)
    assert info["path"] == 0 and y.r.tolist() == [1, 1, 1, 1, 1] and not np.any(y.core)
    rng = np.random.default_rng(71)
    x = O.rand_tt(rng, [3] * 4, [1, 3, 5, 3, 1])
    x.core[7] = np.inf
    with pytest.raises(ttgpu.TTGError) as e:
        _mixed(x, 1e-8)
    assert e.value.status == 3


def test_gram_form_twice_bitwise(gpu_ctx):
    x = O.make_config(3, d=10, n=8, rank=40, seed=72)["x"]
    y1, i1 = _mixed(x, 1e-6)
    y2, i2 = _mixed(x, 1e-6)
    assert i1["path"] == 1 == i2["path"]
    assert y1.core.tobytes() == y2.core.tobytes()


def test_graph_replay_matches_eager(monkeypatch):
    """The Gram form's launch sequence is captured once per shape and pointer
    set and replayed: three tensors of one shape through the same device
    buffers (with other eager work between the calls), every replay bitwise
    equal to the eager run (TTG_GRAPH=0) and within 1e-10 of the oracle."""
    import torch
    ctx = ttgpu.Context(0)
    xs = [O.make_config(3, d=9, n=4, rank=20, seed=s)["x"] for s in (81, 82, 83)]
    n, r = xs[0].n, xs[0].r
    dcore = torch.empty(xs[0].core.size, dtype=torch.float64, device="cuda:0")
    out = torch.empty(ttgpu.capacity(n, r), dtype=torch.float64, device="cuda:0")
    other = O.rand_tt(np.random.default_rng(84), [3] * 5, [1, 3, 7, 7, 3, 1])
    for rep in range(2):
        for x in xs:
            dcore.copy_(torch.from_numpy(x.core))
            info = {}
            r1, y1 = ttgpu.round_dev(ctx, n, r, dcore, 1e-6, precision="fp32", out=out, info=info)
            assert info["path"] == 1
            g = y1.cpu().numpy().copy()
            ttgpu.round(ttgpu.TTTensor(other.n, other.r, other.core), 1e-8)  # eager work between replays
            monkeypatch.setenv("TTG_GRAPH", "0")
            r2, y2 = ttgpu.round_dev(ctx, n, r, dcore, 1e-6, precision="fp32", out=out)
            monkeypatch.delenv("TTG_GRAPH")
            assert r1.tolist() == r2.tolist()
            assert g.tobytes() == y2.cpu().numpy().tobytes()
            ref = O.tt_round(x, 1e-6)
            assert r1.tolist() == ref.r.tolist()
            assert O.tt_rel_diff(O.TT(n, r1, g), ref) <= 1e-10
    ctx.close()
