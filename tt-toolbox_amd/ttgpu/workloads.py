"""Seeded synthetic inputs of the TT rounding / TT-matvec workloads (host, numpy).

Returned as plain arrays in the tt_tensor / tt_matrix layout (flat
column-major cores).  Recipes (seeds are the defaults):

1. ``round1``: X random, d=20, n=4, internal rank 20, cores N(0,1)/sqrt(n*r);
   Z = X + 0.5*X (rank 40 by concatenation); eps 1e-10.  (seed 1)
2. ``matvec2``: A tt_matrix d=12, n=m=4, rank 5; core k = E_k + 0.1*N(0,1)
   where E_k[a,:,:,b] = I_4 for a == b ("identity Kronecker" channel);
   x random d=12, n=4, rank 64, N(0,1)/sqrt(4*64); eps 1e-6.  (seed 2)
3. ``mixed3``: X = Xa + Xb, each a random rank-128 TT (d=32, n=8,
   N(0,1)/sqrt(n*r)) whose cores' right unfoldings r_k x (n r_{k+1}) are
   re-factored as U diag(2^(-j/16)) V^T, j = 0..min-1; eps 1e-6.  (seed 3)
4. ``batch4``: 512 TTs, d in [8,24], n_k in {2,4,8,16}, bond ranks in
   [8,128], N(0,1)/sqrt(n*r), each doubled by self-addition; eps 1e-8. (seed 4)
5. ``completion5``: target rank 8, d=16, n=8, N(0,1)/sqrt(8*8); 65,536
   uniform multi-indices; values + N(0, 1e-3^2) noise; model rank 32 with
   N(0,1)/sqrt(8*32).  (seed 5)

Sum of TTs ("X + Y") is rank concatenation with block-diagonal cores and
block row / block column end cores, as @tt_tensor/plus.m:41-86 builds it.
"""
from __future__ import annotations

import math

import numpy as np


def ps_of(n, r):
    n = np.asarray(n, dtype=np.int64)
    r = np.asarray(r, dtype=np.int64)
    return np.concatenate([[0], np.cumsum(n * r[:-1] * r[1:])]).astype(np.int64)


def flatten(cores):
    n = [c.shape[1] for c in cores]
    r = [cores[0].shape[0]] + [c.shape[2] for c in cores]
    return np.asarray(n, np.int64), np.asarray(r, np.int64), np.concatenate(
        [np.asarray(c, np.float64).ravel(order="F") for c in cores])


def unflatten(n, r, core):
    ps = ps_of(n, r)
    return [core[ps[k]:ps[k + 1]].reshape((r[k], n[k], r[k + 1]), order="F") for k in range(len(n))]


def rand_cores(rng, n, r):
    out = []
    for k in range(len(n)):
        rint = r[1] if k == 0 else r[k]
        out.append(rng.standard_normal((r[k], n[k], r[k + 1])) / math.sqrt(n[k] * max(rint, 1)))
    return out


def concat(a, b, alpha=1.0, beta=1.0):
    """Cores of alpha*A + beta*B by rank concatenation (scalars applied to core 1)."""
    d = len(a)
    out = []
    for k in range(d):
        ra0, n, ra1 = a[k].shape
        rb0, _, rb1 = b[k].shape
        r0 = ra0 if k == 0 else ra0 + rb0
        r1 = ra1 if k == d - 1 else ra1 + rb1
        pp = np.zeros((r0, n, r1))
        sa = alpha if k == 0 else 1.0
        sb = beta if k == 0 else 1.0
        pp[:ra0, :, :ra1] = sa * a[k]
        pp[r0 - rb0:, :, r1 - rb1:] = sb * b[k]
        out.append(pp)
    return out


def rescale_right_unfoldings(cores, base=2.0, step=16.0):
    out = []
    for c in cores:
        rk, nk, rk1 = c.shape
        u, s, vh = np.linalg.svd(c.reshape((rk, nk * rk1), order="F"), full_matrices=False)
        s_new = base ** (-np.arange(s.size) / step)
        out.append(((u * s_new[None, :]) @ vh).reshape((rk, nk, rk1), order="F"))
    return out


def _ranks(d, rr):
    return [1] + [rr] * (d - 1) + [1]


def round1(d=20, n=4, rank=20, seed=1):
    rng = np.random.default_rng(seed)
    x = rand_cores(rng, [n] * d, _ranks(d, rank))
    z = concat(x, x, 1.0, 0.5)
    return {"x": flatten(x), "z": flatten(z), "eps": 1e-10}


def matvec2(d=12, n=4, rank=64, seed=2):
    rng = np.random.default_rng(seed)
    ra = _ranks(d, 5)
    acores = []
    for k in range(d):
        e = np.zeros((ra[k], n, n, ra[k + 1]))
        for a in range(min(ra[k], ra[k + 1])):
            e[a, :, :, a] = np.eye(n)
        acores.append(e + 0.1 * rng.standard_normal(e.shape))
    x = rand_cores(rng, [n] * d, _ranks(d, rank))
    am = np.asarray([n] * d, np.int64)
    aflat = np.concatenate([c.ravel(order="F") for c in acores])
    return {"a": (am, am.copy(), np.asarray(ra, np.int64), aflat), "x": flatten(x), "eps": 1e-6}


def mixed3(d=32, n=8, rank=128, seed=3):
    rng = np.random.default_rng(seed)
    xa = rescale_right_unfoldings(rand_cores(rng, [n] * d, _ranks(d, rank)))
    xb = rescale_right_unfoldings(rand_cores(rng, [n] * d, _ranks(d, rank)))
    return {"x": flatten(concat(xa, xb)), "xa": flatten(xa), "xb": flatten(xb), "eps": 1e-6}


def batch4(count=512, seed=4, dmin=8, dmax=24, nset=(2, 4, 8, 16), rmin=8, rmax=128):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        d = int(rng.integers(dmin, dmax + 1))
        n = [int(v) for v in rng.choice(nset, size=d)]
        r = [1] + [int(v) for v in rng.integers(rmin, rmax + 1, size=d - 1)] + [1]
        x = rand_cores(rng, n, r)
        out.append(flatten(concat(x, x)))
    return {"tensors": out, "eps": 1e-8}


def bilinear1(d=20, n=4, rank=20, arank=5, seed=6):
    """<y, A x> at the config-1 shapes: x, y rank 20, A rank 5, n = m = 4."""
    rng = np.random.default_rng(seed)
    ra = _ranks(d, arank)
    acores = [rng.standard_normal((ra[k], n, n, ra[k + 1])) / math.sqrt(n * ra[k]) for k in range(d)]
    x = rand_cores(rng, [n] * d, _ranks(d, rank))
    y = rand_cores(rng, [n] * d, _ranks(d, rank))
    am = np.asarray([n] * d, np.int64)
    aflat = np.concatenate([c.ravel(order="F") for c in acores])
    return {"a": (am, am.copy(), np.asarray(ra, np.int64), aflat), "x": flatten(x), "y": flatten(y)}


def completion5(d=16, n=8, target_rank=8, model_rank=32, samples=65536, noise=1e-3, seed=5):
    rng = np.random.default_rng(seed)
    tgt = rand_cores(rng, [n] * d, _ranks(d, target_rank))
    idx = rng.integers(0, n, size=(samples, d)).astype(np.int32)
    vals = np.ones(samples)
    left = np.ones((samples, 1))
    for k in range(d):
        left = np.einsum("sa,sab->sb", left, np.transpose(tgt[k][:, idx[:, k], :], (1, 0, 2)))
    vals = left[:, 0] + noise * rng.standard_normal(samples)
    model = rand_cores(rng, [n] * d, _ranks(d, model_rank))
    return {"target": flatten(tgt), "model": flatten(model), "idx": idx, "vals": vals}
