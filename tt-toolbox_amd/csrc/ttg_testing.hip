// Kernel-level test hooks (declared in include/ttgpu_testing.h).  Host
// arrays in and out; each hook runs exactly one building block so that the
// GPU tests can compare it with numpy/LAPACK.  Not part of the TT API.
#include "ttg_internal.h"
#include <algorithm>

namespace ttg {
void gemm_f32_f64out(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K,
                     const float *A, int64_t lda, const float *B, int64_t ldb, double *C,
                     int64_t ldc, float *ws, int64_t wse, int cus);
}  // namespace ttg
using namespace ttg;

namespace {
template <class T> T *up(const T *h, size_t n, hipStream_t s) {
  T *d = nullptr;
  TTG_HIP(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
  if (h && n) TTG_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return d;
}
template <class T> void down(T *h, const T *d, size_t n, hipStream_t s) {
  if (n) TTG_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s));
}
struct Frees {
  std::vector<void *> p;
  ~Frees() {
    for (auto q : p) (void)hipFree(q);
  }
  template <class T> T *add(T *q) {
    p.push_back((void *)q);
    return q;
  }
};
}  // namespace

#define GUARD(ctx, ...)                                    \
  if (!ctx) return TTG_BAD_ARG;                            \
  try {                                                    \
    TTG_HIP(hipSetDevice(ctx->device));                    \
    __VA_ARGS__;                                           \
    TTG_HIP(hipStreamSynchronize(ctx->stream));            \
    return TTG_OK;                                         \
  } catch (const ttg::Error &e) {                          \
    ctx->last_error = e.what();                            \
    return e.status;                                       \
  }

extern "C" {

// C = op(A) op(B), prec 0 = FP64, 1 = FP32 operands (A, B, C given as double)
ttg_status ttg_test_gemm(ttg_ctx *ctx, int prec, int ta, int tb, int64_t M, int64_t N, int64_t K,
                         const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
                         int64_t ldc) {
  GUARD(ctx, {
    hipStream_t s = ctx->stream;
    Frees f;
    const int64_t na = ta ? lda * M : lda * K, nb = tb ? ldb * K : ldb * N, nc = ldc * N;
    const int64_t wse = 32 * M * N + 1024;
    if (prec == 0) {
      double *dA = f.add(up(A, na, s)), *dB = f.add(up(B, nb, s)), *dC = f.add(up<double>(nullptr, nc, s));
      double *ws = f.add(up<double>(nullptr, wse, s));
      TTG_HIP(hipMemsetAsync(dC, 0, nc * sizeof(double), s));
      gemm<double>(s, ta, tb, M, N, K, 1.0, dA, lda, dB, ldb, 0.0, dC, ldc, ws, wse, ctx->num_cus);
      down(C, dC, nc, s);
    } else if (prec == 2) {  // FP64 operands rounded to FP32 while staged, FP32 MFMA
      double *dA = f.add(up(A, na, s)), *dB = f.add(up(B, nb, s)), *dC = f.add(up<double>(nullptr, nc, s));
      float *ws = f.add(up<float>(nullptr, wse, s));
      TTG_HIP(hipMemsetAsync(dC, 0, nc * sizeof(double), s));
      gemm_f64_via_f32(s, ta, tb, M, N, K, dA, lda, dB, ldb, dC, ldc, ws, wse, ctx->num_cus);
      down(C, dC, nc, s);
    } else {
      double *dA = f.add(up(A, na, s)), *dB = f.add(up(B, nb, s)), *dC = f.add(up<double>(nullptr, nc, s));
      float *fA = f.add(up<float>(nullptr, na, s)), *fB = f.add(up<float>(nullptr, nb, s));
      float *ws = f.add(up<float>(nullptr, wse, s));
      convert<float, double>(s, fA, dA, na);
      convert<float, double>(s, fB, dB, nb);
      TTG_HIP(hipMemsetAsync(dC, 0, nc * sizeof(double), s));
      gemm_f32_f64out(s, ta, tb, M, N, K, fA, lda, fB, ldb, dC, ldc, ws, wse, ctx->num_cus);
      down(C, dC, nc, s);
    }
  })
}

ttg_status ttg_test_householder(ttg_ctx *ctx, int64_t m, int64_t n, const double *A, double *Q,
                                double *R) {
  GUARD(ctx, {
    hipStream_t s = ctx->stream;
    Frees f;
    const int64_t kq = std::min(m, n);
    double *dA = f.add(up(A, m * n, s)), *dQ = f.add(up<double>(nullptr, m * kq, s));
    double *dR = f.add(up<double>(nullptr, kq * n, s)), *tau = f.add(up<double>(nullptr, kq + 1, s));
    double *ws = f.add(up<double>(nullptr, householder_ws_elems(m, n, ctx->num_cus), s));
    householder_qr(s, m, n, dA, dQ, dR, tau, ws, ctx->num_cus);
    down(Q, dQ, m * kq, s);
    down(R, dR, kq * n, s);
  })
}

// R (k x k) upper Cholesky factor of G + shift
ttg_status ttg_test_chol(ttg_ctx *ctx, int64_t k, const double *G, double shift_rel, double *R,
                         int *flag, double *stats) {
  GUARD(ctx, {
    hipStream_t s = ctx->stream;
    Frees f;
    double *dG = f.add(up(G, k * k, s)), *dR = f.add(up<double>(nullptr, k * k, s));
    double *dst = f.add(up<double>(nullptr, 8, s));
    int *dfl = f.add(up<int>(nullptr, 4, s));
    chol_factor(s, k, dG, k, shift_rel, dR, dfl, dst);
    down(R, dR, k * k, s);
    down(flag, dfl, 1, s);
    down(stats, dst, 2, s);
  })
}

// X = A R^{-1}, R upper triangular; prec 0 = FP64, 1 = FP32 data
ttg_status ttg_test_trsm(ttg_ctx *ctx, int prec, int64_t m, int64_t k, const double *A,
                         const double *R, double *X) {
  GUARD(ctx, {
    hipStream_t s = ctx->stream;
    Frees f;
    double *dA = f.add(up(A, m * k, s)), *dR = f.add(up(R, k * k, s));
    double *dX = f.add(up<double>(nullptr, m * k, s));
    if (prec == 0) {
      trsm_right_upper<double>(s, m, k, dA, m, dR, dX, m);
    } else {
      float *fA = f.add(up<float>(nullptr, m * k, s)), *fX = f.add(up<float>(nullptr, m * k, s));
      convert<float, double>(s, fA, dA, m * k);
      trsm_right_upper<float>(s, m, k, fA, m, dR, fX, m);
      convert<double, float>(s, dX, fX, m * k);
    }
    down(X, dX, m * k, s);
  })
}

// Jacobi SVD of S (k x k); returns Y (2k x kpad, column-major), sweeps, kpad.
ttg_status ttg_test_jacobi(ttg_ctx *ctx, int64_t k, const double *S, int transS, double *Y,
                           int64_t ycap, int64_t *kpad_out, int *sweeps, int *err) {
  GUARD(ctx, {
    hipStream_t s = ctx->stream;
    Frees f;
    JacobiPlan pl = jacobi_plan(k);
    TTG_CHECK(pl.rows * pl.kpad <= ycap, TTG_NEED_SPACE, "Y capacity");
    double *dS = f.add(up(S, k * k, s)), *dY = f.add(up<double>(nullptr, pl.rows * pl.kpad, s));
    int *dsw = f.add(up<int>(nullptr, 4, s)), *derr = f.add(up<int>(nullptr, 4, s));
    unsigned *bar = f.add(up<unsigned>(nullptr, jacobi_bar_words(60), s));
    double *ws = f.add(up<double>(nullptr, jacobi_polar_ws_elems(pl), s));
    jacobi_svd(s, pl, dS, transS != 0, dY, dsw, bar, derr, 60, ws);
    down(Y, dY, pl.rows * pl.kpad, s);
    down(sweeps, dsw, 1, s);
    down(err, derr, 1, s);
    *kpad_out = pl.kpad;
  })
}

// Jacobi with timestamps of the first sweep: times[4*step + e], e = after load,
// after inner rounds, after the J update, after the barrier (100 MHz ticks).
ttg_status ttg_test_jacobi_timing(ttg_ctx *ctx, int64_t k, const double *S, int64_t *times,
                                  int64_t ntimes) {
  GUARD(ctx, {
    hipStream_t s = ctx->stream;
    Frees f;
    JacobiPlan pl = jacobi_plan(k);
    double *dS = f.add(up(S, k * k, s)), *dY = f.add(up<double>(nullptr, pl.rows * pl.kpad, s));
    int *dsw = f.add(up<int>(nullptr, 4, s)), *derr = f.add(up<int>(nullptr, 4, s));
    unsigned *bar = f.add(up<unsigned>(nullptr, jacobi_bar_words(60), s));
    double *ws = f.add(up<double>(nullptr, jacobi_polar_ws_elems(pl), s));
    unsigned long long *dbg = f.add(up<unsigned long long>(nullptr, ntimes, s));
    TTG_HIP(hipMemsetAsync(dbg, 0, ntimes * 8, s));
    jacobi_svd(s, pl, dS, true, dY, dsw, bar, derr, 60, ws, dbg);
    down(reinterpret_cast<unsigned long long *>(times), dbg, ntimes, s);
  })
}

// Cholesky with timestamps: times[3b] block b start, [3b+1] after the diagonal
// factor, [3b+2] after the panel solve, [3 nblocks] end (100 MHz ticks).
ttg_status ttg_test_chol_timing(ttg_ctx *ctx, int64_t k, const double *G, int64_t *times,
                                int64_t ntimes) {
  GUARD(ctx, {
    hipStream_t s = ctx->stream;
    Frees f;
    TTG_CHECK(ntimes >= 3 * ((k + 31) / 32) + 1, TTG_BAD_ARG, "times too short");
    double *dG = f.add(up(G, k * k, s)), *dR = f.add(up<double>(nullptr, k * k, s));
    double *dst = f.add(up<double>(nullptr, 8, s));
    int *dfl = f.add(up<int>(nullptr, 4, s));
    unsigned long long *dbg = f.add(up<unsigned long long>(nullptr, ntimes, s));
    TTG_HIP(hipMemsetAsync(dbg, 0, ntimes * 8, s));
    chol_factor(s, k, dG, k, 0.0, dR, dfl, dst, dbg);
    down(reinterpret_cast<unsigned long long *>(times), dbg, ntimes, s);
  })
}

// R = chol(G) (upper) and X = A R^{-1} with the inverses of R's diagonal
// blocks that the factorisation leaves (k <= 256), as in one CholeskyQR pass
ttg_status ttg_test_chol_trsm(ttg_ctx *ctx, int64_t m, int64_t k, const double *G, const double *A,
                              double *R, double *X, int *used_w) {
  GUARD(ctx, {
    hipStream_t s = ctx->stream;
    Frees f;
    double *dG = f.add(up(G, k * k, s)), *dR = f.add(up<double>(nullptr, k * k, s));
    double *dA = f.add(up(A, m * k, s)), *dX = f.add(up<double>(nullptr, m * k, s));
    double *dst = f.add(up<double>(nullptr, 8, s));
    int *dfl = f.add(up<int>(nullptr, 4, s));
    double *dW = f.add(up<double>(nullptr, (k / 32 + 1) * 33 * 32, s));
    const bool w = chol_factor(s, k, dG, k, 0.0, dR, dfl, dst, nullptr, dW);
    trsm_right_upper<double>(s, m, k, dA, m, dR, dX, m, nullptr, w ? dW : nullptr);
    *used_w = w ? 1 : 0;
    down(R, dR, k * k, s);
    down(X, dX, m * k, s);
  })
}

// per-phase clock stamps of workgroup 0 of the FP64 triangular solve
ttg_status ttg_test_trsm_timing(ttg_ctx *ctx, int64_t m, int64_t k, const double *A, const double *R,
                                int64_t *times, int64_t ntimes) {
  GUARD(ctx, {
    hipStream_t s = ctx->stream;
    Frees f;
    const int64_t nb = (k + 31) / 32;
    TTG_CHECK(ntimes >= 8 * nb + 3, TTG_BAD_ARG, "times too short");
    double *dA = f.add(up(A, m * k, s)), *dR = f.add(up(R, k * k, s));
    double *dX = f.add(up<double>(nullptr, m * k, s));
    unsigned long long *dbg = f.add(up<unsigned long long>(nullptr, ntimes, s));
    TTG_HIP(hipMemsetAsync(dbg, 0, ntimes * 8, s));
    trsm_right_upper<double>(s, m, k, dA, m, dR, dX, m, dbg);
    down(reinterpret_cast<unsigned long long *>(times), dbg, ntimes, s);
  })
}

}  // extern "C"
