"""Pin the FP64 oracle (CPU only).

Each check is independent of the oracle's own formulas: closed forms, a dense
textbook restatement of the sequential truncation, brute-force dense
products, the reference's own known-answer fixtures (tests/tt-solve/*.mat,
converted to tests/golden/ttsolve_*.npz), invariants, and central finite
differences for the adjoint formulas.  A few tests plant a mistake in a copy
of the oracle and show that a check catches it.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import tt_oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------- my_chop2
def test_my_chop2_known_answers():
    # core/my_chop2.m: tail sums from the smallest value, strict comparison
    sv = np.array([3.0, 2.0, 1.0])
    assert O.my_chop2(sv, 1.5) == 2          # 1^2 = 1 < 2.25, 1+4 = 5 >= 2.25
    assert O.my_chop2(sv, 2.3) == 1          # 5 < 5.29
    assert O.my_chop2(sv, 1.0) == 3          # 1 < 1 is false: strict
    assert O.my_chop2(sv, 0.0) == 3          # eps <= 0 -> numel
    assert O.my_chop2(np.zeros(4), 1.0) == 1  # zero vector -> 1
    assert O.my_chop2(sv, 100.0) == 0        # everything below eps (unreachable from round)


def test_planted_nonstrict_chop_is_caught():
    def chop_le(sv, eps):
        sv0 = np.cumsum(np.asarray(sv)[::-1] ** 2)
        ff = np.nonzero(sv0 <= eps ** 2)[0]
        return sv.size if ff.size == 0 else sv.size - (ff[-1] + 1)
    sv = np.array([3.0, 2.0, 1.0])
    assert chop_le(sv, 1.0) != O.my_chop2(sv, 1.0)


# ---------------------------------------------------------------- round
@pytest.mark.parametrize("seed,n,r,eps", [
    (1, [3, 4, 2, 3], [1, 3, 5, 4, 1], 1e-2),
    (2, [2, 2, 2, 2, 2], [1, 2, 4, 4, 2, 1], 0.3),
    (3, [4, 3, 5], [1, 4, 6, 1], 1e-12),
    (4, [3, 3, 3, 3], [1, 3, 9, 3, 1], 0.0),
])
def test_round_matches_dense_sequential_truncation(seed, n, r, eps):
    rng = np.random.default_rng(seed)
    x = O.rand_tt(rng, n, r)
    x = O.tt_plus(x, O.tt_scal(0.3, O.rand_tt(rng, n, r)))
    y = O.tt_round(x, eps)
    dense, ranks = O.dense_sequential_round(O.tt_full(x), n, eps)
    assert y.r.tolist() == ranks
    np.testing.assert_allclose(O.tt_full(y), dense, rtol=0, atol=1e-12 * np.linalg.norm(dense))


def test_round_error_bound_and_invariants():
    rng = np.random.default_rng(7)
    n = [3, 4, 3, 4, 3]
    x = O.rand_tt(rng, n, [1, 3, 8, 8, 3, 1])
    eps = 0.05
    info = {}
    y = O.tt_round(x, eps, info=info)
    full_x, full_y = O.tt_full(x), O.tt_full(y)
    assert np.linalg.norm(full_y - full_x) <= eps * np.linalg.norm(full_x) * (1 + 1e-12)
    # round.m:102-107: cores 2..d are nrm0-scaled right-orthonormal, ||core1|| = nrm0
    nrm0 = info["nrm0"]
    for k in range(1, y.d):
        c = y.core_k(k).reshape((y.r[k], -1), order="F")
        np.testing.assert_allclose(c @ c.T, nrm0 ** 2 * np.eye(y.r[k]), atol=1e-12 * nrm0 ** 2)
    assert abs(np.linalg.norm(y.core_k(0)) - nrm0) <= 1e-12 * nrm0
    assert abs(nrm0 ** y.d - np.linalg.norm(full_y)) <= 1e-12 * np.linalg.norm(full_y)


def test_config1_closed_form():
    c = O.make_config(1)
    y = O.tt_round(c["z"], c["eps"])
    assert y.r.tolist() == [1, 4, 16] + [20] * 15 + [16, 4, 1]
    assert O.tt_rel_diff(y, O.tt_scal(1.5, c["x"])) <= 1e-10


def test_eps0_and_rmax_and_zero_and_d1():
    rng = np.random.default_rng(2)
    x = O.rand_tt(rng, [4] * 5, [1, 4, 16, 16, 4, 1])
    y = O.tt_round(x, 0.0)
    assert y.r.tolist() == [1, 4, 16, 16, 4, 1]
    assert O.tt_rel_diff(y, x) <= 1e-13
    y = O.tt_round(x, 1e-14, rmax=6)
    assert max(y.r) <= 6
    dense, ranks = O.dense_sequential_round(O.tt_full(x), x.n, 1e-14, rmax=6)
    assert y.r.tolist() == ranks
    z = O.TT([3, 3], [1, 2, 1], np.zeros(12))
    yz = O.tt_round(z, 1e-8)
    assert yz.r.tolist() == [1, 1, 1] and not np.any(yz.core)
    x1 = O.rand_tt(rng, [6], [1, 1])
    np.testing.assert_allclose(O.tt_round(x1, 1e-8).core, x1.core, rtol=1e-14)


def test_gesdd_and_gesvd_agree_on_ranks():
    c = O.make_config(3, d=8, n=4, rank=24, seed=21)
    a = O.tt_round(c["x"], 1e-3)
    b = O.tt_round(c["x"], 1e-3, svd_driver="gesvd")
    assert a.r.tolist() == b.r.tolist()
    assert O.tt_rel_diff(a, b) <= 1e-12


def test_golden_oracle_vectors_reproduce():
    g = np.load(GOLDEN / "oracle_cfg1_d8.npz")
    y = O.tt_round(O.TT(g["n"], g["r"], g["core"]), float(g["eps"]))
    assert y.r.tolist() == g["out_r"].tolist()
    assert O.tt_rel_diff(y, O.TT(g["n"], g["out_r"], g["out_core"])) <= 1e-13
    g = np.load(GOLDEN / "oracle_cfg3_d8.npz")
    y = O.tt_round(O.TT(g["n"], g["r"], g["core"]), float(g["eps"]))
    assert y.r.tolist() == g["out_r"].tolist()


# ---------------------------------------------------------------- mtimes
def test_mtimes_matches_dense_product_nonsymmetric():
    rng = np.random.default_rng(3)
    n, m = [2, 3, 2], [3, 2, 4]
    cores = [rng.standard_normal((rr, nn, mm, rr1)) for rr, nn, mm, rr1 in
             zip([1, 2, 3], n, m, [2, 3, 1])]
    a = O.TTM.from_cores(cores)
    x = O.rand_tt(rng, m, [1, 2, 2, 1])
    y = O.ttm_mtimes(a, x)
    assert y.r.tolist() == [1, 4, 6, 1]
    np.testing.assert_allclose(O.tt_full(y), O.ttm_full(a) @ O.tt_full(x), rtol=1e-12, atol=1e-12)


def test_planted_rank_order_mistake_in_mtimes_is_caught():
    """Plant a plausible mistake: right rank index ordered (alpha', a') instead of (a', alpha')."""
    rng = np.random.default_rng(4)
    n, m = [2, 3, 2], [3, 2, 2]
    cores = [rng.standard_normal((rr, nn, mm, rr1)) for rr, nn, mm, rr1 in
             zip([1, 2, 3], n, m, [2, 3, 1])]
    a = O.TTM.from_cores(cores)
    x = O.rand_tt(rng, m, [1, 2, 2, 1])
    dense = O.ttm_full(a) @ O.tt_full(x)
    bad = []
    for i in range(3):
        mcur = a.tt.core_k(i).reshape((a.r[i] * n[i] * m[i], a.r[i + 1]), order="F").T
        mcur = mcur.reshape((a.r[i + 1] * a.r[i] * n[i], m[i]), order="F")
        vcur = np.transpose(x.core_k(i), (1, 0, 2)).reshape((m[i], -1), order="F")
        p = (mcur @ vcur).reshape((a.r[i + 1], a.r[i], n[i], x.r[i], x.r[i + 1]), order="F")
        p = np.transpose(p, (1, 3, 2, 4, 0))  # planted: wrong order on the right rank index
        bad.append(p.reshape((a.r[i] * x.r[i], n[i], a.r[i + 1] * x.r[i + 1]), order="F"))
    assert np.linalg.norm(O.tt_full(O.TT.from_cores(bad)) - dense) > 1e-3 * np.linalg.norm(dense)
    np.testing.assert_allclose(O.tt_full(O.ttm_mtimes(a, x)), dense, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name,tol", [("ex2_6-10", 1e-5), ("ex6", 1e-3)])
def test_ttsolve_fixture_residual(name, tol):
    """tests/tt-solve/README: A*x ~= rhs for the shipped solutions."""
    g = np.load(GOLDEN / f"ttsolve_{name}.npz")
    a = O.TTM(g["A_n"], g["A_m"], O.TT(g["A_n"] * g["A_m"], g["A_r"], g["A_core"]))
    x = O.TT(g["x_n"], g["x_r"], g["x_core"])
    rhs = O.TT(g["rhs_n"], g["rhs_r"], g["rhs_core"])
    res = O.tt_rel_diff(O.ttm_mtimes(a, x), rhs)
    assert res <= tol, res


# ---------------------------------------------------------------- helpers
def test_plus_scal_norm_full_consistency():
    rng = np.random.default_rng(5)
    n = [2, 3, 4]
    a = O.rand_tt(rng, n, [1, 2, 3, 1])
    b = O.rand_tt(rng, n, [1, 3, 2, 1])
    np.testing.assert_allclose(O.tt_full(O.tt_plus(a, b)), O.tt_full(a) + O.tt_full(b), atol=1e-13)
    np.testing.assert_allclose(O.tt_full(O.tt_scal(-2.5, a)), -2.5 * O.tt_full(a), atol=1e-13)
    assert abs(O.tt_norm(a) - np.linalg.norm(O.tt_full(a))) <= 1e-13 * np.linalg.norm(O.tt_full(a))
    assert abs(O.tt_dot(a, b) - O.tt_full(a) @ O.tt_full(b)) <= 1e-12
    # d = 1 (plus.m:30-33): the cores add
    u = O.rand_tt(rng, [5], [1, 1])
    v = O.rand_tt(rng, [5], [1, 1])
    np.testing.assert_allclose(O.tt_full(O.tt_plus(u, v)), O.tt_full(u) + O.tt_full(v), atol=1e-15)


# ---------------------------------------------------------------- adjoints
def _fd_grad(f, cores, k, h=1e-6, samples=12, rng=None):
    rng = rng or np.random.default_rng(0)
    flat = cores[k].ravel()
    picks = rng.choice(flat.size, size=min(samples, flat.size), replace=False)
    out = []
    for p in picks:
        cp = [c.copy() for c in cores]
        cm = [c.copy() for c in cores]
        cp[k].ravel()[p] += h
        cm[k].ravel()[p] -= h
        out.append((p, (f(cp) - f(cm)) / (2 * h)))
    return out


def test_completion_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    n, r = [3, 4, 3, 2], [1, 3, 4, 2, 1]
    x = O.rand_tt(rng, n, r)
    S = 200
    idx = np.stack([rng.integers(0, nn, size=S) for nn in n], axis=1)
    vals = rng.standard_normal(S)
    loss, grads = O.completion_value_grad(x, idx, vals)

    def f(cores):
        return O.completion_value_grad(O.TT.from_cores(cores), idx, vals)[0]
    full = O.tt_full(x).reshape(n, order="F")
    res = full[tuple(idx.T)] - vals
    assert abs(loss - 0.5 * res @ res) <= 1e-10 * max(1.0, loss)
    cores = x.cores()
    for k in range(len(n)):
        for p, fd in _fd_grad(f, cores, k, rng=rng):
            assert abs(grads[k].ravel()[p] - fd) <= 1e-6 * max(1.0, abs(fd))


def test_bilinear_gradient_matches_finite_differences():
    rng = np.random.default_rng(12)
    n, m = [2, 3, 2], [3, 2, 2]
    acores = [rng.standard_normal((ra, nn, mm, rb)) for ra, nn, mm, rb in zip([1, 2, 3], n, m, [2, 3, 1])]
    a = O.TTM.from_cores(acores)
    x = O.rand_tt(rng, m, [1, 2, 3, 1])
    y = O.rand_tt(rng, n, [1, 3, 2, 1])
    val, gx, gy = O.bilinear_value_grad(y, a, x)
    assert abs(val - O.tt_full(y) @ (O.ttm_full(a) @ O.tt_full(x))) <= 1e-12 * max(1, abs(val))
    xc, yc = x.cores(), y.cores()
    for k in range(3):
        for p, fd in _fd_grad(lambda cs: O.bilinear_value_grad(y, a, O.TT.from_cores(cs))[0], xc, k, rng=rng):
            assert abs(gx[k].ravel()[p] - fd) <= 1e-6 * max(1.0, abs(fd))
        for p, fd in _fd_grad(lambda cs: O.bilinear_value_grad(O.TT.from_cores(cs), a, x)[0], yc, k, rng=rng):
            assert abs(gy[k].ravel()[p] - fd) <= 1e-6 * max(1.0, abs(fd))


# ---------------------------------------------------------------- full-size property checkers
def _dense_unfolding_ranks(x):
    full = O.tt_full(x)
    n = [int(v) for v in x.n]
    out = [1]
    for i in range(1, len(n)):
        rows = int(np.prod(n[:i]))
        s = np.linalg.svd(full.reshape((rows, -1), order="F"), compute_uv=False)
        out.append(int(np.sum(s > 1e-10 * s[0])))
    return out + [1]


@pytest.mark.parametrize("seed,n,r", [
    (31, [2, 3, 2, 4, 2], [1, 2, 7, 9, 3, 1]),     # caps from both sides
    (32, [4, 4, 4], [1, 3, 5, 1]),                  # no cap binds
    (33, [2, 2, 2, 2, 2, 2], [1, 5, 5, 5, 5, 5, 1]),  # every bond capped
])
def test_generic_ranks_equal_dense_unfolding_ranks(seed, n, r):
    rng = np.random.default_rng(seed)
    x = O.rand_tt(rng, n, r)
    z = O.tt_plus(x, x)  # exact rank of x, concatenated to 2r
    want = _dense_unfolding_ranks(x)
    assert O.generic_ranks(n, r) == want
    assert O.tt_round(z, 1e-8).r.tolist() == want


def test_balance_residual_of_rounded_tt_and_planted_scaling():
    rng = np.random.default_rng(34)
    x = O.rand_tt(rng, [3, 4, 3, 2], [1, 3, 6, 4, 1])
    y = O.tt_round(O.tt_plus(x, x), 1e-10)
    nrm0, res = O.balance_residual(y)
    assert res <= 1e-13
    assert abs(nrm0 ** y.d - O.tt_norm(y)) <= 1e-12 * O.tt_norm(y)
    bad = y.copy()
    ps = bad.ps
    bad.core[ps[2]:ps[3]] *= 1.001  # one core off its balance
    assert O.balance_residual(bad)[1] > 1e-3


def test_tt_norm_gram_matches_qr_norm_and_dense():
    rng = np.random.default_rng(35)
    x = O.rand_tt(rng, [3, 2, 4, 3], [1, 3, 5, 2, 1])
    dense = np.linalg.norm(O.tt_full(x))
    assert abs(O.tt_norm_gram(x) - dense) <= 1e-13 * dense
    assert abs(O.tt_norm(x) - dense) <= 1e-13 * dense


# ---------------------------------------------------------------- TT-SVD
def _dense_left_projection(full, n, eps):
    """TT-SVD in exact arithmetic: bond i (left to right) projects the
    (i_1..i_i) x (rest) unfolding of the current dense tensor onto its
    leading r_i left singular vectors, r_i = my_chop2(s, eps/sqrt(d-1) ||s||)."""
    d = len(n)
    y = np.asarray(full, dtype=np.float64).copy()
    ranks = [1] * (d + 1)
    ep = eps / math.sqrt(d - 1)
    for i in range(1, d):
        rows = int(np.prod(n[:i]))
        mat = y.reshape((rows, -1), order="F")
        u, s, vh = np.linalg.svd(mat, full_matrices=False)
        r1 = O.my_chop2(s, np.linalg.norm(s) * ep)
        ranks[i] = r1
        y = (u[:, :r1] @ (u[:, :r1].T @ mat)).reshape(-1, order="F")
    return y, ranks


@pytest.mark.parametrize("seed,n,r,eps", [
    (31, [3, 4, 2, 3, 2], [1, 3, 5, 4, 2, 1], 1e-12),
    (32, [2, 3, 4, 3], [1, 2, 6, 3, 1], 0.05),
    (33, [4, 4, 4], [1, 4, 4, 1], 0.2),
])
def test_tt_from_full_matches_dense_left_projection(seed, n, r, eps):
    rng = np.random.default_rng(seed)
    x = O.rand_tt(rng, n, r)
    x = O.tt_plus(x, O.tt_scal(0.1, O.rand_tt(rng, n, r)))
    full = O.tt_full(x)
    t = O.tt_from_full(full, eps, n)
    dense, ranks = _dense_left_projection(full, n, eps)
    assert t.r.tolist() == ranks
    np.testing.assert_allclose(O.tt_full(t), dense, rtol=0, atol=1e-12 * np.linalg.norm(full))
    assert np.linalg.norm(O.tt_full(t) - full) <= eps * np.linalg.norm(full) * (1 + 1e-12)
    if eps == 1e-12:  # nothing truncates: the generic ranks of a sum of two rank-r TTs
        assert t.r.tolist() == O.generic_ranks(n, [1] + [2 * v for v in r[1:-1]] + [1])


def test_planted_right_to_left_ttsvd_is_caught():
    """A right-to-left TT-SVD (a plausible mistake: round.m's direction) gives
    other ranks on a tensor whose two directions disagree."""
    rng = np.random.default_rng(34)
    n = [2, 3, 4, 3]
    x = O.rand_tt(rng, n, [1, 2, 6, 3, 1])
    x = O.tt_plus(x, O.tt_scal(0.1, O.rand_tt(rng, n, [1, 2, 6, 3, 1])))
    full = O.tt_full(x)
    found = False
    for eps in (0.26, 0.3):
        _, left = _dense_left_projection(full, n, eps)
        _, right = O.dense_sequential_round(full, n, eps)
        if left != right:
            found = True
            assert O.tt_from_full(full, eps, n).r.tolist() == left
    assert found
