"""CPU oracle for TT rounding and TT-matrix x TT-tensor contraction (FP64).

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``libttgpu.so`` and the ``ttgpu`` host package)
never calls it and has no CPU fallback.

Every function restates the reference MATLAB code line by line; the file:line
it follows is cited in its docstring (paths relative to the TT-Toolbox 2.2.2
tree).  MATLAB built-ins map onto the same LAPACK call classes:

* ``qr(A,0)``       -> ``numpy.linalg.qr(A, 'reduced')``  (dgeqrf + dorgqr)
* ``svd(A,'econ')`` -> ``numpy.linalg.svd(A, full_matrices=False)`` (dgesdd;
  ``svd_driver='gesvd'`` switches to dgesvd through scipy as a cross-check)
* ``reshape``/``permute`` -> Fortran-order ``reshape``/``transpose``.

Layout (``@tt_tensor/tt_tensor.m:33-41``): a TT tensor is ``d``, mode sizes
``n[d]``, ranks ``r[d+1]`` and one flat column-major ``core`` vector; core k
is ``r[k] x n[k] x r[k+1]`` starting at ``ps[k]`` (0-based here, MATLAB's
``ps`` minus one).  A TT matrix (``@tt_matrix/tt_matrix.m:33-39``) is a TT
tensor with mode sizes ``n*m`` whose core k is ``r[k] x n[k] x m[k] x r[k+1]``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "TT", "TTM", "my_chop2", "tt_round", "ttm_round", "ttm_mtimes",
    "tt_plus", "tt_scal", "tt_minus", "tt_norm", "tt_full", "ttm_full",
    "tt_rel_diff", "tt_dot", "rand_tt", "dense_sequential_round", "tt_from_full",
    "completion_value_grad", "bilinear_value_grad", "make_config",
]


# ---------------------------------------------------------------------------
# containers
# ---------------------------------------------------------------------------
def _ps(n, r):
    """``ps=cumsum([1;n.*r(1:d).*r(2:d+1)])`` (``@tt_tensor/tt_tensor.m:143``), 0-based."""
    n = np.asarray(n, dtype=np.int64)
    r = np.asarray(r, dtype=np.int64)
    return np.concatenate([[0], np.cumsum(n * r[:-1] * r[1:])]).astype(np.int64)


@dataclass
class TT:
    """``tt_tensor`` struct (``@tt_tensor/tt_tensor.m:33-41``)."""
    n: np.ndarray
    r: np.ndarray
    core: np.ndarray

    def __post_init__(self):
        self.n = np.asarray(self.n, dtype=np.int64).copy()
        self.r = np.asarray(self.r, dtype=np.int64).copy()
        self.core = np.ascontiguousarray(self.core, dtype=np.float64).ravel()
        if self.r.size != self.n.size + 1:
            raise ValueError("r must have d+1 entries")
        if self.core.size != int(self.ps[-1]):
            raise ValueError("core length does not match n and r")

    @property
    def d(self):
        return int(self.n.size)

    @property
    def ps(self):
        return _ps(self.n, self.r)

    def core_k(self, k):
        """Core k as an ``r_k x n_k x r_{k+1}`` column-major array (a view)."""
        ps = self.ps
        return self.core[ps[k]:ps[k + 1]].reshape(
            (self.r[k], self.n[k], self.r[k + 1]), order="F")

    def cores(self):
        return [self.core_k(k).copy() for k in range(self.d)]

    @staticmethod
    def from_cores(cores):
        n = [c.shape[1] for c in cores]
        r = [cores[0].shape[0]] + [c.shape[2] for c in cores]
        flat = np.concatenate([np.asarray(c, dtype=np.float64).ravel(order="F") for c in cores])
        return TT(n, r, flat)

    def copy(self):
        return TT(self.n, self.r, self.core.copy())


@dataclass
class TTM:
    """``tt_matrix`` struct (``@tt_matrix/tt_matrix.m:33-39``): ``tt`` with modes ``n.*m``."""
    n: np.ndarray
    m: np.ndarray
    tt: TT = field(repr=False)

    def __post_init__(self):
        self.n = np.asarray(self.n, dtype=np.int64).copy()
        self.m = np.asarray(self.m, dtype=np.int64).copy()
        if not np.array_equal(self.tt.n, self.n * self.m):
            raise ValueError("tt mode sizes must be n.*m")

    @property
    def d(self):
        return int(self.n.size)

    @property
    def r(self):
        return self.tt.r

    def core_k(self, k):
        """``@tt_matrix/core.m:23-28``: core k as ``r_k x n_k x m_k x r_{k+1}``."""
        c = self.tt.core_k(k)
        return c.reshape((self.tt.r[k], self.n[k], self.m[k], self.tt.r[k + 1]), order="F")

    @staticmethod
    def from_cores(cores):
        n = [c.shape[1] for c in cores]
        m = [c.shape[2] for c in cores]
        tcores = [c.reshape((c.shape[0], c.shape[1] * c.shape[2], c.shape[3]), order="F") for c in cores]
        return TTM(n, m, TT.from_cores(tcores))


# ---------------------------------------------------------------------------
# rank rule and rounding
# ---------------------------------------------------------------------------
def my_chop2(sv, eps):
    """``core/my_chop2.m:18-34``: minimal r with ``sqrt(sum(sv(r+1:end).^2)) < eps``.

    Zero vector -> 1; ``eps <= 0`` -> ``numel(sv)``; otherwise the reversed
    cumulative sum of squares is compared strictly with ``eps^2`` and
    ``r = numel(sv) - ff(end)``.  The cumulative sum is taken sequentially from
    the smallest value, in the order MATLAB's ``cumsum`` uses.
    """
    sv = np.asarray(sv, dtype=np.float64).ravel()
    if np.linalg.norm(sv) == 0:
        return 1
    if eps <= 0:
        return int(sv.size)
    sv0 = np.cumsum(sv[::-1] ** 2)
    ff = np.nonzero(sv0 < eps ** 2)[0]
    if ff.size == 0:
        return int(sv.size)
    return int(sv.size - (ff[-1] + 1))


def _qr0(a):
    """MATLAB ``qr(a,0)``: economy QR through LAPACK dgeqrf+dorgqr."""
    return np.linalg.qr(a, mode="reduced")


def _svd_econ(a, driver="gesdd"):
    """MATLAB ``svd(a,'econ')`` returning ``u, s, v`` (not ``v'``)."""
    if driver == "gesdd":
        u, s, vh = np.linalg.svd(a, full_matrices=False)
    else:
        import scipy.linalg
        u, s, vh = scipy.linalg.svd(a, full_matrices=False, lapack_driver="gesvd")
    return u, s, vh.T


def tt_round(tt, eps, rmax=None, info=None, svd_driver="gesdd"):
    """``@tt_tensor/round.m`` restated line by line.

    * ``:20-32`` arguments: ``rmax`` defaults to ``prod(size(tt))``, which no
      TT rank can exceed, so ``None`` means "no cap"; a scalar is expanded to
      ``d+1`` entries; ``rmax[i]`` caps bond i in the second sweep only (``:78``).
    * ``:39-53`` left-to-right QR sweep with per-bond norm extraction and the
      rank cap ``r(i+1)=size(core0,2)``.
    * ``:56`` ``ep=eps/sqrt(d-1)``.
    * ``:59-89`` right-to-left SVD sweep with ``my_chop2(s,norm(s)*ep)``.
    * ``:94-99`` first-core normalisation, ``:102-107`` balance by
      ``exp(mean(log|nrm|))``.

    ``info`` (a dict) receives per-bond singular values, thresholds and the
    relative margin of each rank decision (how far the tail sum is from the
    threshold), plus the QR-capped ranks.
    """
    d = tt.d
    n = tt.n.copy()
    r = tt.r.copy()
    if rmax is None:
        rmax_v = np.full(d + 1, np.iinfo(np.int64).max, dtype=np.int64)
    else:
        rmax_v = np.asarray(rmax, dtype=np.int64).ravel()
        if rmax_v.size == 1:
            rmax_v = np.full(d + 1, int(rmax_v[0]), dtype=np.int64)
    cores = tt.cores()
    nrm = np.zeros(d)
    # left-to-right orthogonalisation (round.m:39-53)
    core0 = cores[0]
    for i in range(d - 1):
        core0 = core0.reshape((r[i] * n[i], r[i + 1]), order="F")
        q, ru = _qr0(core0)
        nrm[i + 1] = np.linalg.norm(ru, "fro")
        if nrm[i + 1] != 0:
            ru = ru / nrm[i + 1]
        core1 = cores[i + 1].reshape((r[i + 1], n[i + 1] * r[i + 2]), order="F")
        core1 = ru @ core1
        r[i + 1] = q.shape[1]
        cores[i] = q.reshape((r[i], n[i], r[i + 1]), order="F")
        cores[i + 1] = core1.reshape((r[i + 1], n[i + 1], r[i + 2]), order="F")
        core0 = cores[i + 1]
    if info is not None:
        info["r_qr"] = r.copy()
        info["sv"] = [None] * (d + 1)
        info["thr"] = [None] * (d + 1)
        info["margin"] = [None] * (d + 1)
    ep = eps / math.sqrt(d - 1) if d > 1 else math.inf
    # right-to-left truncation (round.m:59-89)
    core0 = cores[d - 1]
    for i in range(d - 1, 0, -1):
        core0 = core0.reshape((r[i], n[i] * r[i + 1]), order="F")
        core1 = cores[i - 1].reshape((r[i - 1] * n[i - 1], r[i]), order="F")
        u, s, v = _svd_econ(core0, svd_driver)
        thr = np.linalg.norm(s) * ep
        r1 = my_chop2(s, thr)
        r1 = int(min(r1, rmax_v[i]))
        if info is not None:
            info["sv"][i] = s.copy()
            info["thr"][i] = thr
            info["margin"][i] = _chop_margin(s, thr)
        u = u[:, :r1] * s[:r1][None, :]
        v = v[:, :r1]
        r[i] = r1
        core1 = core1 @ u
        core0 = v.T
        cores[i] = core0.reshape((r[i], n[i], r[i + 1]), order="F")
        cores[i - 1] = core1.reshape((r[i - 1], n[i - 1], r[i]), order="F")
        core0 = cores[i - 1]
    # normalise the first core and balance (round.m:94-107)
    nrm[0] = np.linalg.norm(cores[0].ravel())
    if nrm[0] != 0:
        cores[0] = cores[0] / nrm[0]
    with np.errstate(divide="ignore"):
        nrm0 = np.sum(np.log(np.abs(nrm)))
    nrm0 = math.exp(nrm0 / d) if np.isfinite(nrm0) else 0.0
    cores = [c * nrm0 for c in cores]
    out = TT.from_cores(cores)
    if info is not None:
        info["nrm"] = nrm.copy()
        info["nrm0"] = nrm0
    return out


def _chop_margin(s, thr):
    """Relative distance of the deciding tail sum from ``thr^2`` (rank-decision margin)."""
    if thr <= 0 or np.linalg.norm(s) == 0:
        return math.inf
    c = np.cumsum(s[::-1] ** 2)
    return float(np.min(np.abs(c - thr ** 2)) / thr ** 2)


def ttm_round(ttm, eps, rmax=None):
    """``@tt_matrix/round.m:23-27``: round the flattened ``tt`` with modes ``n.*m``."""
    return TTM(ttm.n, ttm.m, tt_round(ttm.tt, eps, rmax))


# ---------------------------------------------------------------------------
# contraction
# ---------------------------------------------------------------------------
def ttm_mtimes(a, x):
    """``@tt_matrix/mtimes.m:38-65`` (tt_matrix * tt_tensor).

    Per core: ``mcur`` is ``A_i`` reshaped to ``[rA n m, rA']``, transposed and
    reshaped to ``[rA' rA n, m]``; ``vcur`` is ``x_i`` permuted ``[2,1,3]`` to
    ``[m, rx rx']``; ``pcur=mcur*vcur`` is reshaped to
    ``[rA',rA,n,rx,rx']`` and permuted ``[2,4,3,1,5]``.  The product rank
    index is ``a + rA*alpha`` (A-rank fastest); no rounding.
    """
    if a.d != x.d or not np.array_equal(a.m, x.n):
        raise ValueError("inconsistent sizes")
    d = a.d
    n, m = a.n, a.m
    rm, rv = a.r, x.r
    rp = rm * rv
    out = []
    for i in range(d):
        mcur = a.tt.core_k(i).reshape((rm[i] * n[i] * m[i], rm[i + 1]), order="F")
        mcur = mcur.T
        mcur = mcur.reshape((rm[i + 1] * rm[i] * n[i], m[i]), order="F")
        vcur = x.core_k(i)
        vcur = np.transpose(vcur, (1, 0, 2)).reshape((m[i], rv[i] * rv[i + 1]), order="F")
        pcur = mcur @ vcur
        pcur = pcur.reshape((rm[i + 1], rm[i], n[i], rv[i], rv[i + 1]), order="F")
        pcur = np.transpose(pcur, (1, 3, 2, 0, 4))
        out.append(pcur.reshape((rp[i], n[i], rp[i + 1]), order="F"))
    return TT.from_cores(out)


# ---------------------------------------------------------------------------
# input assembly and parity metrics
# ---------------------------------------------------------------------------
def tt_plus(b, c):
    """``@tt_tensor/plus.m:41-86``: rank concatenation, block-diagonal cores.

    Boundary ranks must agree (``:50-59``); ``p1`` goes to the leading block
    ``pp(1:r1(i),:,1:r1(i+1))`` and ``p2`` to the trailing block
    ``pp(r(i)-r2(i)+1:r(i),:,r(i+1)-r2(i+1)+1:r(i+1))`` (``:82-83``).
    """
    if b.d != c.d or not np.array_equal(b.n, c.n):
        raise ValueError("inconsistent sizes")
    d = b.d
    if d == 1:  # :30-33: one core each, the cores add
        if not np.array_equal(b.r, c.r):
            raise ValueError("inconsistent sizes")
        return TT(b.n.copy(), b.r.copy(), b.core + c.core)
    r1, r2 = b.r, c.r
    r = r1 + r2
    if r1[0] != r2[0]:
        raise ValueError("Inconsistent sizes in mode 1")
    r[0] = r1[0]
    if r1[d] != r2[d]:
        raise ValueError("Inconsistent sizes in the last mode")
    r[d] = r1[d]
    out = []
    for i in range(d):
        pp = np.zeros((r[i], b.n[i], r[i + 1]))
        p1 = b.core_k(i)
        p2 = c.core_k(i)
        pp[:r1[i], :, :r1[i + 1]] = p1
        pp[r[i] - r2[i]:, :, r[i + 1] - r2[i + 1]:] = p2
        out.append(pp)
    return TT.from_cores(out)


def tt_scal(alpha, b):
    """``@tt_tensor/mtimes.m:10-27`` (scalar * tt_tensor): scales core 1 only."""
    cores = b.cores()
    cores[0] = alpha * cores[0]
    return TT.from_cores(cores)


def tt_minus(b, c):
    """``@tt_tensor/minus.m:18``: ``a=b+(-1.0)*c``."""
    return tt_plus(b, tt_scal(-1.0, c))


def tt_norm(tt):
    """``@tt_tensor/norm.m:25-41``: QR-sweep norm with ``max(nrm,1e-308)`` guard."""
    d = tt.d
    n, r = tt.n, tt.r.copy()
    cores = tt.cores()
    nrm = np.zeros(d)
    core0 = cores[0]
    for i in range(d - 1):
        core0 = core0.reshape((r[i] * n[i], r[i + 1]), order="F")
        q, ru = _qr0(core0)
        nrm[i] = max(np.linalg.norm(ru, "fro"), 1e-308)
        ru = ru / nrm[i]
        core1 = cores[i + 1].reshape((r[i + 1], n[i + 1] * r[i + 2]), order="F")
        core0 = ru @ core1
        r[i + 1] = q.shape[1]
    nrm[d - 1] = np.linalg.norm(core0.ravel())
    return float(np.prod(nrm))


def tt_full(tt):
    """``@tt_tensor/full.m:19-35``: dense column-major vector (boundary ranks 1)."""
    d, n, r = tt.d, tt.n, tt.r
    a = tt.core_k(0).reshape(-1, order="F")
    for i in range(1, d):
        cr = tt.core_k(i).reshape((r[i], n[i] * r[i + 1]), order="F")
        a = a.reshape((a.size // r[i], r[i]), order="F") @ cr
    return a.reshape(-1, order="F")


def ttm_full(ttm):
    """``@tt_matrix/full.m:17-25``: dense ``prod(n) x prod(m)`` matrix (rows follow n)."""
    d = ttm.d
    a = tt_full(ttm.tt)
    v = np.stack([ttm.n, ttm.m], axis=1).ravel()  # [n1 m1 n2 m2 ...]
    a = a.reshape(tuple(v), order="F")
    prm = np.arange(2 * d).reshape((2, d)).T.ravel()  # MATLAB prm = [1 d+1 2 d+2 ...] - 1
    inv = np.argsort(prm)
    a = np.transpose(a, inv)  # ipermute(a, prm)
    return a.reshape((int(np.prod(ttm.n)), int(np.prod(ttm.m))), order="F")


def tt_rel_diff(y, yref):
    """``norm(y-yref)/norm(yref)`` through TT arithmetic (``minus.m:18`` + ``norm.m:25-41``)."""
    den = tt_norm(yref)
    return tt_norm(tt_minus(y, yref)) / den if den > 0 else tt_norm(y)


def tt_dot(a, b):
    """``@tt_tensor/dot.m:92-111`` for boundary ranks 1: left interface chain."""
    p = np.ones((1, 1))
    for i in range(a.d):
        ca, cb = a.core_k(i), b.core_k(i)
        p = np.einsum("ab,aic,bie->ce", p, ca, cb)
    return float(p[0, 0])


def tt_from_full(b, eps=1e-14, n=None, svd_driver="gesdd"):
    """``@tt_tensor/tt_tensor.m:175-244`` (TT-SVD, boundary ranks 1): left to
    right, ``m = n(i) r(i)``, ``[u,s,v] = svd(c,'econ')``,
    ``r1 = my_chop2(s, ep*norm(s))`` with ``ep = eps/sqrt(d-1)``, core ``u``,
    ``c = (v*diag(s))'``; the last core is what remains.  ``b`` is the flat
    column-major tensor (``b(:)``), ``n`` its mode sizes."""
    b = np.asarray(b, dtype=np.float64).reshape(-1, order="F")
    n = np.asarray(n, dtype=np.int64)
    d = n.size
    if d == 1:
        return TT(n.copy(), np.array([1, 1]), b.copy())
    r = np.ones(d + 1, dtype=np.int64)
    ep = eps / math.sqrt(d - 1)
    c = b
    cores = []
    for i in range(d - 1):
        m = int(n[i] * r[i])
        c = c.reshape((m, c.size // m), order="F")
        u, s, v = _svd_econ(c, svd_driver)
        r1 = my_chop2(s, ep * np.linalg.norm(s))
        u, s, v = u[:, :r1], s[:r1], v[:, :r1]
        r[i + 1] = r1
        cores.append(u.reshape((r[i], n[i], r1), order="F"))
        c = (v * s[None, :]).T
    cores.append(c.reshape((r[d - 1], n[d - 1], 1), order="F"))
    return TT.from_cores(cores)


# ---------------------------------------------------------------------------
# size-independent properties of a rounded TT (checks at full size)
# ---------------------------------------------------------------------------
def generic_ranks(n, r):
    """Ranks ``round`` returns for a tensor of exact generic TT rank when no
    singular value falls under the threshold: the sweep-1 caps of
    ``@tt_tensor/round.m:48`` (``r_{i+1} <- min(r_i n_i, r_{i+1})``, left to
    right), then the sweep-2 caps ``r_i <- min(r_i, n_i r_{i+1})`` (right to
    left, the SVD of an ``r_i x n_i r_{i+1}`` unfolding has at most that many
    values).  For a random TT and for ``x + x`` (rank concatenation of a
    random ``x``) these are the ranks of the dense unfoldings."""
    n = [int(v) for v in n]
    rr = [int(v) for v in r]
    d = len(n)
    for i in range(d - 1):
        rr[i + 1] = min(rr[i] * n[i], rr[i + 1])
    for i in range(d - 1, 0, -1):
        rr[i] = min(rr[i], n[i] * rr[i + 1])
    return rr


def balance_residual(tt):
    """Departure of a rounded TT from the balancing of ``@tt_tensor/round.m:90-107``:
    cores 2..d have orthogonal right unfoldings scaled by ``nrm0``
    (``C_k C_k^T = nrm0^2 I``) and ``||C_1||_F = nrm0``, with
    ``nrm0 = ||C_1||_F``.  Returns ``(nrm0, max_k max|C_k C_k^T / nrm0^2 - I|)``."""
    nrm0 = float(np.linalg.norm(tt.core_k(0)))
    if nrm0 == 0.0:
        return 0.0, 0.0
    worst = 0.0
    for k in range(1, tt.d):
        c = tt.core_k(k).reshape((tt.r[k], -1), order="F")
        g = (c @ c.T) / (nrm0 * nrm0)
        worst = max(worst, float(np.max(np.abs(g - np.eye(g.shape[0])))))
    return nrm0, worst


def tt_norm_gram(tt):
    """``||X||_F`` by the left Gram chain ``G_k = sum_j C_k[:,j,:]^T G_{k-1} C_k[:,j,:]``
    (the interface products of ``@tt_tensor/dot.m:92-111`` with both
    arguments equal).  Cheap at low rank; accurate to ``sqrt(u)`` relative at worst."""
    g = np.ones((1, 1))
    for k in range(tt.d):
        rk, nk, rk1 = int(tt.r[k]), int(tt.n[k]), int(tt.r[k + 1])
        c = tt.core_k(k)
        t = g @ c.reshape((rk, nk * rk1), order="F")                # G_{k-1} x_1 C_k
        g = c.reshape((rk * nk, rk1), order="F").T @ t.reshape((rk * nk, rk1), order="F")
    return float(np.sqrt(max(g[0, 0], 0.0)))


# ---------------------------------------------------------------------------
# textbook cross-check: sequential truncated SVD on the dense tensor
# ---------------------------------------------------------------------------
def dense_sequential_round(full, n, eps, rmax=None):
    """Dense restatement of the truncation that ``round.m`` performs.

    In exact arithmetic, the SVD taken at bond i of the second sweep of
    ``round.m`` has the singular values of the unfolding
    ``(i_1..i_{i-1}) x (i_i..i_d)`` of the tensor whose bonds i+1..d have
    already been truncated; truncation projects that unfolding onto its
    leading ``r1`` right singular vectors.  This routine does exactly that on
    the dense tensor (bonds d-1..1), with ``my_chop2`` and ``ep=eps/sqrt(d-1)``.
    Returns the dense result and the ranks ``[1, r_1 .. r_{d-1}, 1]``.
    """
    n = [int(v) for v in n]
    d = len(n)
    y = np.asarray(full, dtype=np.float64).copy()
    ranks = [1] * (d + 1)
    ep = eps / math.sqrt(d - 1) if d > 1 else math.inf
    for i in range(d - 1, 0, -1):
        rows = int(np.prod(n[:i]))
        mat = y.reshape((rows, -1), order="F")
        u, s, vh = np.linalg.svd(mat, full_matrices=False)
        r1 = my_chop2(s, np.linalg.norm(s) * ep)
        if rmax is not None:
            r1 = min(r1, int(np.broadcast_to(np.asarray(rmax), (d + 1,))[i]))
        ranks[i] = r1
        mat = (mat @ vh[:r1].T) @ vh[:r1]
        y = mat.reshape(-1, order="F")
    return y, ranks


# ---------------------------------------------------------------------------
# generators (seeded, host side)
# ---------------------------------------------------------------------------
def rand_tt(rng, n, r, scale="nr"):
    """Random TT with i.i.d. N(0,1) cores scaled by ``1/sqrt(n_k r_k)`` (interior rank)."""
    n = [int(v) for v in n]
    r = [int(v) for v in r]
    cores = []
    for k in range(len(n)):
        c = rng.standard_normal((r[k], n[k], r[k + 1]))
        if scale == "nr":
            c /= math.sqrt(n[k] * max(r[k], r[k + 1] if k == 0 else r[k]))
        cores.append(c)
    return TT.from_cores([np.asfortranarray(c) for c in cores])


def _ranks(d, rr):
    return [1] + [rr] * (d - 1) + [1]


def _rescale_right_unfoldings(tt, base=2.0, step=16.0):
    """Replace each core's right-unfolding singular values by ``base^(-j/step)``."""
    cores = []
    for c in tt.cores():
        rk, nk, rk1 = c.shape
        mat = c.reshape((rk, nk * rk1), order="F")
        u, s, vh = np.linalg.svd(mat, full_matrices=False)
        s_new = base ** (-np.arange(s.size) / step)
        cores.append(((u * s_new[None, :]) @ vh).reshape((rk, nk, rk1), order="F"))
    return TT.from_cores(cores)


def make_config(cfg, d=None, rank=None, n=None, seed=None):
    """Seeded inputs for the five workloads (sizes overridable for small parity runs).

    1: ``Z = X + 0.5 X``, X: n=4, r=20, d=20, eps 1e-10 (seed 1)
    2: ``A`` (n=m=4, rA=5, identity Kronecker + 0.1 N(0,1)), ``x`` (n=4, r=64),
       d=12, eps 1e-6 (seed 2)
    3: ``X = Xa + Xb``, each rank 128, d=32, n=8; each summand core's right
       unfolding gets singular values ``2^(-j/16)``; eps 1e-6 (seed 3)
    """
    if cfg == 1:
        d = d or 20
        n = n or 4
        rank = rank or 20
        rng = np.random.default_rng(1 if seed is None else seed)
        x = rand_tt(rng, [n] * d, _ranks(d, rank))
        z = tt_plus(x, tt_scal(0.5, x))
        return {"x": x, "z": z, "eps": 1e-10}
    if cfg == 2:
        d = d or 12
        n = n or 4
        rank = rank or 64
        rng = np.random.default_rng(2 if seed is None else seed)
        ra = _ranks(d, 5)
        acores = []
        for k in range(d):
            e = np.zeros((ra[k], n, n, ra[k + 1]))
            for a in range(min(ra[k], ra[k + 1])):
                e[a, :, :, a] = np.eye(n)
            acores.append(e + 0.1 * rng.standard_normal(e.shape))
        a = TTM.from_cores(acores)
        x = rand_tt(rng, [n] * d, _ranks(d, rank))
        return {"a": a, "x": x, "eps": 1e-6}
    if cfg == 3:
        d = d or 32
        n = n or 8
        rank = rank or 128
        rng = np.random.default_rng(3 if seed is None else seed)
        xa = _rescale_right_unfoldings(rand_tt(rng, [n] * d, _ranks(d, rank)))
        xb = _rescale_right_unfoldings(rand_tt(rng, [n] * d, _ranks(d, rank)))
        return {"x": tt_plus(xa, xb), "eps": 1e-6}
    raise ValueError(f"unknown config {cfg}")


def make_batch(count=512, seed=4, dmin=8, dmax=24, nset=(2, 4, 8, 16), rmin=8, rmax=128):
    """Batch of uneven TT tensors, each doubled to rank 2r by self-addition (eps 1e-8)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        d = int(rng.integers(dmin, dmax + 1))
        n = [int(v) for v in rng.choice(nset, size=d)]
        r = [1] + [int(v) for v in rng.integers(rmin, rmax + 1, size=d - 1)] + [1]
        x = rand_tt(rng, n, r)
        out.append(tt_plus(x, x))
    return out


# ---------------------------------------------------------------------------
# adjoint references (the reference has no gradients; formulas from its
# interface products)
# ---------------------------------------------------------------------------
def completion_value_grad(tt, idx, vals):
    """Loss ``0.5*||P_Omega(X) - y||^2`` and its core gradients.

    Entry evaluation is the product of ``r_k x r_{k+1}`` slices
    (``@tt_tensor/subsref.m:116-122``); the per-sample left and right interface
    vectors follow ``tt_regression/reg_als.m:22-39,66-92``.  ``idx`` is an
    ``S x d`` integer array of 0-based multi-indices.  Returns ``(loss, grads)``
    with ``grads[k]`` shaped like core k.
    """
    d = tt.d
    idx = np.asarray(idx, dtype=np.int64)
    S = idx.shape[0]
    cores = tt.cores()
    left = [np.ones((S, 1))]
    for k in range(d - 1):
        left.append(np.einsum("sa,sab->sb", left[k], np.transpose(cores[k][:, idx[:, k], :], (1, 0, 2))))
    right = [None] * d
    right[d - 1] = np.ones((S, 1))
    for k in range(d - 1, 0, -1):
        right[k - 1] = np.einsum("sab,sb->sa", np.transpose(cores[k][:, idx[:, k], :], (1, 0, 2)), right[k])
    val = np.einsum("sa,sab,sb->s", left[d - 1], np.transpose(cores[d - 1][:, idx[:, d - 1], :], (1, 0, 2)), right[d - 1])
    res = val - vals
    loss = 0.5 * float(res @ res)
    grads = []
    for k in range(d):
        g = np.zeros_like(cores[k])
        contrib = np.einsum("s,sa,sb->sab", res, left[k], right[k])
        for i in range(tt.n[k]):
            sel = idx[:, k] == i
            g[:, i, :] = contrib[sel].sum(axis=0)
        grads.append(g)
    return loss, grads


def bilinear_value_grad(y, a, x):
    """``<y, A x>`` and its gradients w.r.t. the cores of x and of y.

    The left interface chain ``phi(ry, rA, rx)`` is ``core/tt_mvdot.m:16-57``
    (contract y, A and x core by core); the right chain is its mirror.  The
    gradient for core k contracts both interfaces with ``A_k`` and the core
    of the other vector.
    """
    d = x.d
    lphi = [np.ones((1, 1, 1))]
    for k in range(d):
        yk, ak, xk = y.core_k(k), a.core_k(k), x.core_k(k)
        lphi.append(np.einsum("cax,cie,aijb,xjf->ebf", lphi[k], yk, ak, xk))
    val = float(lphi[d][0, 0, 0])
    rphi = [None] * (d + 1)
    rphi[d] = np.ones((1, 1, 1))
    for k in range(d - 1, -1, -1):
        yk, ak, xk = y.core_k(k), a.core_k(k), x.core_k(k)
        rphi[k] = np.einsum("cie,aijb,xjf,ebf->cax", yk, ak, xk, rphi[k + 1])
    gx, gy = [], []
    for k in range(d):
        yk, ak, xk = y.core_k(k), a.core_k(k), x.core_k(k)
        gx.append(np.einsum("cax,cie,aijb,ebf->xjf", lphi[k], yk, ak, rphi[k + 1]))
        gy.append(np.einsum("cax,aijb,xjf,ebf->cie", lphi[k], ak, xk, rphi[k + 1]))
    return val, gx, gy
