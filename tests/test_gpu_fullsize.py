"""Parity of the rounding workloads at their real sizes (GPU).

* Mixed-vs-FP64 workload (d=32, n=8, r=256, the sum of two rank-128 TTs with
  graded core spectra, eps 1e-6): the plain FP64 form and the mixed form
  against the FP64 oracle (restatement of @tt_tensor/round.m:20-111) on the
  same tensor the benchmark rounds.  Bars: identical ranks; relative
  Frobenius difference <= 1e-12 (plain) and <= 1e-6 (mixed); the balancing
  of round.m:90-107 (C_k C_k^T = nrm0^2 I for k >= 2) to 1e-12 (plain) and
  1e-6 (mixed).
* The 512-tensor batch (d in [8,24], n in {2,4,8,16}, ranks in [8,128], each
  tensor x + x, eps 1e-8) in one batched call: the oracle on the 32 tensors of
  most QR work (identical ranks, <= 1e-12); for all 512 tensors the closed
  form of the ranks (round(x + x) keeps the generic TT rank of x), the
  balancing invariant (<= 1e-12) and ||Y|| = nrm0^d = ||x + x|| = 2 ||x||
  (<= 1e-10 relative; ||x|| by the Gram chain of the undoubled blocks).
"""
import numpy as np
import pytest
import torch

import ttgpu
from oracle import tt_oracle as O
from ttgpu import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg3():
    c = O.make_config(3)
    x = c["x"]
    print("config 3: oracle round", flush=True)
    ref = O.tt_round(x, c["eps"])
    return x, c["eps"], ref


@pytest.mark.parametrize("prec,tol,gram", [("fp64", 1e-12, "1"), ("fp32", 1e-6, "1"), ("fp64", 1e-12, "0"),
                                           ("fp32", 1e-6, "0")])
def test_config3_full_size(gpu_ctx, cfg3, monkeypatch, prec, tol, gram):
    """Both forms, through the Gram form (every bond is proven to keep its
    singular values) and through the QR/SVD sweeps (TTG_GRAM=0)."""
    x, eps, ref = cfg3
    monkeypatch.setenv("TTG_GRAM", gram)
    print(f"config 3 d=32 n=8 r=256 eps=1e-6 form {prec} gram {gram}", flush=True)
    info = {}
    y = ttgpu.round(ttgpu.TTTensor(x.n, x.r, x.core), eps, precision=prec, info=info)
    print("ranks", y.r.tolist(), "info", info, flush=True)
    assert info["path"] == int(gram)
    assert y.r.tolist() == ref.r.tolist()
    yo = O.TT(y.n, y.r, y.core)
    err = O.tt_rel_diff(yo, ref)
    print(f"rel diff vs oracle {err:.3e}", flush=True)
    assert err <= tol
    _, bres = O.balance_residual(yo)
    assert bres <= tol, bres


def _qr_work(n, r):
    c, rk = 0.0, float(r[0])
    for j in range(len(n) - 1):
        m = rk * n[j]
        k = min(m, float(r[j + 1]))
        c += m * k * k
        rk = k
    return c


def _undoubled(n, r, core):
    """The summand x of z = x + x (rank concatenation, @tt_tensor/plus.m:41-86):
    the leading blocks of z's cores."""
    z = O.TT(n, r, core)
    d = z.d
    rx = [1] + [int(v) // 2 for v in r[1:-1]] + [1]
    cores = [z.core_k(k)[:rx[k], :, :rx[k + 1]] for k in range(d)]
    return O.TT.from_cores([np.asfortranarray(c) for c in cores]), rx


@pytest.fixture(scope="module")
def batch512(gpu_ctx):
    print("batch: generating 512 tensors", flush=True)
    bb = W.batch4()
    tens = bb["tensors"]
    cofs = np.concatenate([[0], np.cumsum([t[2].size for t in tens])])
    dall = torch.empty(int(cofs[-1]), dtype=torch.float64, device="cuda:0")
    for i, t in enumerate(tens):
        dall[int(cofs[i]):int(cofs[i + 1])].copy_(torch.from_numpy(t[2]))
    torch.cuda.synchronize()
    print("batch: one batched call", flush=True)
    res = ttgpu.round_batched_dev(gpu_ctx, [t[0] for t in tens], [t[1] for t in tens], dall, cofs[:-1],
                                  bb["eps"], precision="fp64")
    out = [(np.asarray(r_t), y_t.cpu().numpy()) for r_t, y_t in res]
    del dall, res
    torch.cuda.empty_cache()
    order = np.argsort([-_qr_work(t[0], t[1]) for t in tens], kind="stable")
    return tens, out, bb["eps"], order


def test_batch512_closed_form_ranks_balance_and_norm(batch512):
    tens, out, eps, _ = batch512
    worst_bal, worst_nrm = 0.0, 0.0
    for i, ((n, r, core), (r_t, y_t)) in enumerate(zip(tens, out)):
        x, rx = _undoubled(n, r, core)
        want = O.generic_ranks(n, rx)
        assert r_t.tolist() == want, f"tensor {i}"
        y = O.TT(n, r_t, y_t)
        nrm0, bres = O.balance_residual(y)
        worst_bal = max(worst_bal, bres)
        nz = 2.0 * O.tt_norm_gram(x)
        rel = abs(nrm0 ** len(n) - nz) / nz
        worst_nrm = max(worst_nrm, rel)
        assert bres <= 1e-12, f"tensor {i}: balance {bres}"
        assert rel <= 1e-10, f"tensor {i}: norm {rel}"
    print(f"512 tensors: ranks closed form, balance <= {worst_bal:.2e}, norm <= {worst_nrm:.2e}", flush=True)


@pytest.mark.parametrize("part", range(4))
def test_batch512_costliest_match_oracle(batch512, part):
    tens, out, eps, order = batch512
    for j in order[8 * part:8 * part + 8]:
        n, r, core = tens[j]
        ref = O.tt_round(O.TT(n, r, core), eps)
        r_t, y_t = out[j]
        assert r_t.tolist() == ref.r.tolist(), f"tensor {j}"
        err = O.tt_rel_diff(O.TT(n, r_t, y_t), ref)
        assert err <= 1e-12, f"tensor {j}: {err}"
    print(f"part {part}: 8 tensors match the oracle", flush=True)


def test_ttsolve_fixtures_through_hip_mtimes(gpu_ctx):
    """The reference-held tt-solve fixtures (tests/tt-solve/*.mat, converted):
    A x on the GPU equals the oracle's A x element by element, and satisfies
    A x ~ rhs to the accuracy the fixtures were saved at."""
    from pathlib import Path
    golden = Path(__file__).resolve().parent / "golden"
    for name, tol in (("ex2_6-10", 1e-5), ("ex6", 1e-3)):
        g = np.load(golden / f"ttsolve_{name}.npz")
        a = O.TTM(g["A_n"], g["A_m"], O.TT(g["A_n"] * g["A_m"], g["A_r"], g["A_core"]))
        x = O.TT(g["x_n"], g["x_r"], g["x_core"])
        rhs = O.TT(g["rhs_n"], g["rhs_r"], g["rhs_core"])
        ref = O.ttm_mtimes(a, x)
        y = ttgpu.mtimes(ttgpu.TTMatrix(a.n, a.m, a.tt.r, a.tt.core), ttgpu.TTTensor(x.n, x.r, x.core))
        assert y.r.tolist() == ref.r.tolist()
        np.testing.assert_allclose(y.core, ref.core, rtol=1e-12, atol=1e-13 * np.abs(ref.core).max())
        res = O.tt_rel_diff(O.TT(y.n, y.r, y.core), rhs)
        print(f"{name}: ||A x - rhs|| / ||rhs|| = {res:.2e}", flush=True)
        assert res <= tol
