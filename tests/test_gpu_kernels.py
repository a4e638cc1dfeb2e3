"""Building blocks of the rounding sweep against numpy/LAPACK (kernel-level hooks)."""
import ctypes as C

import numpy as np
import pytest

import ttgpu

pytestmark = pytest.mark.gpu


def f64p(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("prec,ta,tb,M,N,K", [
    (0, 0, 0, 37, 51, 29), (0, 1, 0, 64, 64, 2048), (0, 0, 1, 70, 33, 100), (0, 1, 1, 17, 90, 45),
    (1, 0, 0, 65, 130, 40), (1, 1, 0, 256, 256, 2048), (2, 1, 0, 256, 256, 2048), (2, 0, 0, 100, 70, 300),
    (0, 0, 0, 2048, 256, 256), (0, 0, 1, 2048, 256, 256), (0, 1, 0, 256, 2048, 256),
])
def test_gemm(gpu_ctx, prec, ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    Af, Bf = np.asfortranarray(A), np.asfortranarray(B)
    Cm = np.zeros((M, N), order="F")
    s = gpu_ctx.lib.ttg_test_gemm(gpu_ctx.h, prec, ta, tb, M, N, K, f64p(Af), Af.shape[0], f64p(Bf),
                                  Bf.shape[0], f64p(Cm), M)
    gpu_ctx.check(s)
    ref = (A.T if ta else A) @ (B.T if tb else B)
    tol = 1e-12 if prec == 0 else 1e-5
    assert np.max(np.abs(Cm - ref)) <= tol * np.max(np.abs(ref)) * np.sqrt(K), np.max(np.abs(Cm - ref))


@pytest.mark.parametrize("ta,tb,M,N,K", [(0, 0, 45, 70, 130), (1, 0, 33, 64, 257), (0, 1, 100, 64, 31),
                                          (1, 1, 70, 65, 96), (0, 0, 2048, 256, 256), (1, 0, 256, 256, 2048)])
def test_gemm_padded_leading_dims(gpu_ctx, ta, tb, M, N, K):
    """Operands and result inside larger column-major arrays (leading
    dimensions past the row counts): the tile loads must read only the
    window, and the store only touch C's window."""
    rng = np.random.default_rng(M + 3 * N + 5 * K)
    ar, ac = (K, M) if ta else (M, K)
    br, bc = (N, K) if tb else (K, N)
    Abig = np.asfortranarray(rng.standard_normal((ar + 5, ac)))
    Bbig = np.asfortranarray(rng.standard_normal((br + 7, bc)))
    Cbig = np.asfortranarray(np.full((M + 3, N), 7.0))  # the hook clears C on the device first
    s = gpu_ctx.lib.ttg_test_gemm(gpu_ctx.h, 0, ta, tb, M, N, K, f64p(Abig), ar + 5, f64p(Bbig), br + 7,
                                  f64p(Cbig), M + 3)
    gpu_ctx.check(s)
    A, B = Abig[:ar], Bbig[:br]
    ref = (A.T if ta else A) @ (B.T if tb else B)
    assert np.max(np.abs(Cbig[:M] - ref)) <= 1e-12 * np.max(np.abs(ref)) * np.sqrt(K)
    assert np.all(Cbig[M:] == 0.0)


@pytest.mark.parametrize("m,n,rank", [(40, 7, 7), (160, 40, 20), (8, 30, 8), (300, 64, 64), (64, 64, 10),
                                         (2048, 256, 128), (1000, 100, 100), (70, 90, 50)])
def test_householder(gpu_ctx, m, n, rank):
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, rank)) @ rng.standard_normal((rank, n))
    A = np.asfortranarray(A)
    kq = min(m, n)
    Q = np.zeros((m, kq), order="F")
    R = np.zeros((kq, n), order="F")
    gpu_ctx.check(gpu_ctx.lib.ttg_test_householder(gpu_ctx.h, m, n, f64p(A), f64p(Q), f64p(R)))
    assert np.max(np.abs(Q.T @ Q - np.eye(kq))) <= 1e-13
    assert np.max(np.abs(Q @ R - A)) <= 1e-13 * np.max(np.abs(A)) * m
    assert np.max(np.abs(np.tril(R, -1))) == 0.0
    # LAPACK agrees up to column signs
    q2, r2 = np.linalg.qr(A)
    np.testing.assert_allclose(np.abs(np.diag(R)[:rank]), np.abs(np.diag(r2)[:rank]), rtol=1e-10)


def _chol(gpu_ctx, G, shift=0.0):
    k = G.shape[0]
    G = np.asfortranarray(G)
    R = np.zeros((k, k), order="F")
    flag = np.zeros(1, dtype=np.int32)
    stats = np.zeros(2)
    gpu_ctx.check(gpu_ctx.lib.ttg_test_chol(gpu_ctx.h, k, f64p(G), shift, f64p(R), f64p(flag), f64p(stats)))
    return R, int(flag[0]), stats


@pytest.mark.parametrize("k", [5, 32, 33, 40, 65, 96, 100, 224, 255, 256, 320, 480, 512])
def test_cholesky(gpu_ctx, k):
    """Against LAPACK, with the diagonal-spread and max|R - I| statistics
    (k <= 256: the kernel with the overlapped block-row update, incl. one
    block, a last block of one column and full blocks; k <= 480: fused
    pivot/panel kernel; 512: the panel-after-pivot kernel)."""
    rng = np.random.default_rng(k)
    A = rng.standard_normal((3 * k, k))
    G = A.T @ A
    R, flag, stats = _chol(gpu_ctx, G)
    assert flag == 0
    ref = np.linalg.cholesky(G).T
    np.testing.assert_allclose(R, ref, rtol=1e-10, atol=1e-10 * np.max(np.abs(ref)))
    assert np.max(np.abs(np.tril(R, -1))) == 0.0
    np.testing.assert_allclose(stats[0], np.max(np.abs(np.diag(ref))) / np.min(np.abs(np.diag(ref))), rtol=1e-8)
    np.testing.assert_allclose(stats[1], np.max(np.abs(ref - np.eye(k))), rtol=1e-8)


def test_cholesky_near_identity(gpu_ctx):
    """A near-identity Gram (the second CholeskyQR pass): R - I to full
    relative accuracy of the perturbation."""
    rng = np.random.default_rng(3)
    k = 200
    E = rng.standard_normal((k, k)) * 1e-9
    G = np.eye(k) + E + E.T
    R, flag, stats = _chol(gpu_ctx, G)
    assert flag == 0
    ref = np.linalg.cholesky(G).T
    assert np.max(np.abs(R - ref)) <= 4e-15
    np.testing.assert_allclose(stats[1], np.max(np.abs(ref - np.eye(k))), rtol=1e-6)


def test_cholesky_flags_indefinite(gpu_ctx):
    G = np.diag([4.0, 1.0, -1.0, 2.0, 1.0, 1.0])
    _, flag, _ = _chol(gpu_ctx, G)
    assert flag == 3
    # a failing pivot in the fourth 32-column block
    d = np.ones(150)
    d[100] = -2.0
    _, flag, stats = _chol(gpu_ctx, np.diag(d))
    assert flag == 101
    assert np.isinf(stats[0]) and np.isinf(stats[1])


@pytest.mark.parametrize("m,k,logcond", [(2048, 256, 3), (2048, 256, 7), (300, 40, 10), (256, 256, 5),
                                          (1000, 100, 12), (77, 33, 4)])
def test_chol_then_trsm_with_block_inverses(gpu_ctx, m, k, logcond):
    """One CholeskyQR pass: R = chol(A^T A), X = A R^{-1} through the
    inverses of R's diagonal blocks plus one refinement step; the residual
    ||X R - A|| stays at rounding level however ill-conditioned A is, and
    X matches LAPACK's solve to kappa u."""
    rng = np.random.default_rng(m + 3 * k)
    U, _ = np.linalg.qr(rng.standard_normal((m, k)))
    V, _ = np.linalg.qr(rng.standard_normal((k, k)))
    A = np.asfortranarray((U * np.logspace(0, -logcond / 2, k)) @ V.T)
    G = np.asfortranarray(A.T @ A)
    R = np.zeros((k, k), order="F")
    X = np.zeros((m, k), order="F")
    used = np.zeros(1, dtype=np.int32)
    gpu_ctx.check(gpu_ctx.lib.ttg_test_chol_trsm(gpu_ctx.h, m, k, f64p(G), f64p(A), f64p(R), f64p(X), f64p(used)))
    assert int(used[0]) == 1
    u = np.finfo(float).eps
    res = np.linalg.norm(X @ R - A) / (np.linalg.norm(X) * np.linalg.norm(R))
    assert res <= 50 * u, res
    ref = np.linalg.solve(R.T, A.T).T
    cond = np.linalg.cond(R)
    assert np.linalg.norm(X - ref) <= 50 * cond * u * np.linalg.norm(ref)


@pytest.mark.parametrize("prec,m,k,logcond", [(0, 2048, 256, 8), (0, 300, 40, 12), (1, 2048, 256, 3), (0, 1000, 320, 6), (0, 50, 7, 10)])
def test_trsm_backward_stable(gpu_ctx, prec, m, k, logcond):
    """X = A R^{-1}: the residual ||X R - A|| stays O(u ||X|| ||R||) even for ill-conditioned R."""
    rng = np.random.default_rng(m + k)
    U, _ = np.linalg.qr(rng.standard_normal((k, k)))
    Vt, _ = np.linalg.qr(rng.standard_normal((k, k)))
    B = (U * np.logspace(0, -logcond, k)) @ Vt
    R = np.linalg.qr(B)[1]
    R = np.asfortranarray(R * np.sign(np.diag(R))[:, None])
    A = np.asfortranarray(rng.standard_normal((m, k)) @ R)
    X = np.zeros((m, k), order="F")
    gpu_ctx.check(gpu_ctx.lib.ttg_test_trsm(gpu_ctx.h, prec, m, k, f64p(A), f64p(R), f64p(X)))
    u = 1.1e-16 if prec == 0 else 6e-8
    res = np.linalg.norm(X @ R - A) / (np.linalg.norm(X) * np.linalg.norm(R))
    print("trsm residual", res, "cond", np.linalg.cond(R), flush=True)
    assert res <= 10 * u * np.sqrt(k)


@pytest.mark.parametrize("k,trans,deficient", [(6, 0, 0), (16, 1, 3), (40, 1, 20), (64, 0, 0), (128, 1, 0), (256, 0, 100), (300, 1, 0)])
def test_jacobi_svd(gpu_ctx, k, trans, deficient):
    rng = np.random.default_rng(k + trans)
    S = rng.standard_normal((k, k))
    if deficient:
        S = S[:, :k - deficient] @ rng.standard_normal((k - deficient, k))
    S = np.asfortranarray(S)
    kpad = C.c_int64()
    cap = 2 * k * (k + 64)
    Y = np.zeros(cap)
    sweeps = np.zeros(1, dtype=np.int32)
    err = np.zeros(1, dtype=np.int32)
    gpu_ctx.check(gpu_ctx.lib.ttg_test_jacobi(gpu_ctx.h, k, f64p(S), trans, f64p(Y), cap, C.byref(kpad),
                                              f64p(sweeps), f64p(err)))
    kp = kpad.value
    Ym = Y[:2 * k * kp].reshape((2 * k, kp), order="F")
    top, J = Ym[:k, :k], Ym[k:, :k]
    M = S.T if trans else S
    print("sweeps", sweeps[0], "err", err[0], "kpad", kp, flush=True)
    assert err[0] == 0
    np.testing.assert_allclose(M @ J, top, atol=1e-12 * np.linalg.norm(M))
    np.testing.assert_allclose(J.T @ J, np.eye(k), atol=1e-12)
    sig = np.sort(np.linalg.norm(top, axis=0))[::-1]
    ref = np.linalg.svd(M, compute_uv=False)
    np.testing.assert_allclose(sig, ref, atol=1e-13 * ref[0] * k)
    G = top.T @ top
    off = G - np.diag(np.diag(G))
    assert np.max(np.abs(off)) <= 1e-12 * ref[0] ** 2
