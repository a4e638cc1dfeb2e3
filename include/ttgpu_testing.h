/*
 * Kernel-level test hooks of libttgpu.so.  No reference counterpart: they
 * expose one building block each (GEMM, Householder QR, Cholesky + triangular
 * inverse, Jacobi SVD) so the GPU tests can compare it with LAPACK.  Host
 * arrays, column-major, FP64 on the boundary.
 */
#ifndef TTGPU_TESTING_H
#define TTGPU_TESTING_H
#include "ttgpu.h"
#ifdef __cplusplus
extern "C" {
#endif
// prec 0: FP64; 1: FP32 data and MFMA, FP64 result; 2: FP64 data rounded to FP32 for the MFMA
ttg_status ttg_test_gemm(ttg_ctx *ctx, int prec, int ta, int tb, int64_t M, int64_t N, int64_t K,
                         const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
                         int64_t ldc);
ttg_status ttg_test_householder(ttg_ctx *ctx, int64_t m, int64_t n, const double *A, double *Q,
                                double *R);
ttg_status ttg_test_chol(ttg_ctx *ctx, int64_t k, const double *G, double shift_rel, double *R,
                         int *flag, double *stats);
// R = chol(G), X = A R^{-1} through the block inverses the factorisation leaves
// (*used_w = 1 when it left them: k <= 256)
ttg_status ttg_test_chol_trsm(ttg_ctx *ctx, int64_t m, int64_t k, const double *G, const double *A,
                              double *R, double *X, int *used_w);
ttg_status ttg_test_trsm(ttg_ctx *ctx, int prec, int64_t m, int64_t k, const double *A,
                         const double *R, double *X);
ttg_status ttg_test_jacobi(ttg_ctx *ctx, int64_t k, const double *S, int transS, double *Y,
                           int64_t ycap, int64_t *kpad_out, int *sweeps, int *err);
ttg_status ttg_test_jacobi_timing(ttg_ctx *ctx, int64_t k, const double *S, int64_t *times,
                                  int64_t ntimes);
ttg_status ttg_test_chol_timing(ttg_ctx *ctx, int64_t k, const double *G, int64_t *times,
                                int64_t ntimes);
ttg_status ttg_test_trsm_timing(ttg_ctx *ctx, int64_t m, int64_t k, const double *A, const double *R,
                                int64_t *times, int64_t ntimes);
#ifdef __cplusplus
}
#endif
#endif
