// ttgpu device kernels for gfx950 (CDNA4, wave64).
//
// Dense building blocks of the TT rounding sweep:
//   * MFMA GEMM (FP64 16x16x4, FP32 16x16x4), column-major, split-K with a
//     fixed-order reduction (results are bitwise reproducible run to run);
//   * one-workgroup Householder QR (LAPACK dgeqr2 + dorg2r semantics);
//   * one-workgroup Cholesky of a k x k Gram matrix, wave-per-column
//     triangular inverse;
//   * cooperative block one-sided Jacobi SVD (Hestenes rotations, de Rijk /
//     Rutishauser formulas) with column blocks resident in LDS;
//   * descending sort of singular values + the my_chop2 rank rule;
//   * gathers, transposes, norms and the final balance of round.m:94-107.
#include "ttg_internal.h"
#include "ttg_device.h"
#include <cmath>
#include <cfloat>
#include <algorithm>

namespace ttg {

thread_local Profiler *g_prof = nullptr;

void Profiler::begin(hipStream_t s) {
  if (pool.size() < 2 * (used + 1)) {
    hipEvent_t a, b;
    TTG_HIP(hipEventCreate(&a));
    TTG_HIP(hipEventCreate(&b));
    pool.push_back(a);
    pool.push_back(b);
  }
  TTG_HIP(hipEventRecord(pool[2 * used], s));
}
void Profiler::end(hipStream_t s, double fl, double by) {
  TTG_HIP(hipEventRecord(pool[2 * used + 1], s));
  flops.push_back(fl);
  bytes.push_back(by);
  ++used;
}

// ---------------------------------------------------------------------------
// MFMA GEMM
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <class T> struct MF;
// Accumulator layouts (measured on gfx950 with an in-bounds one-wave probe):
//   f64 16x16x4: lane l, element v holds D[l/16 + 4v][l%16]
//   f32 16x16x4: lane l, element v holds D[4(l/16) + v][l%16]
template <> struct MF<double> {
  typedef v4d V;
  static __device__ __forceinline__ V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct MF<float> {
  typedef v4f V;
  static __device__ __forceinline__ V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

constexpr int GBM = 64, GBN = 64, GBK = 16;

// C(MxN) = alpha*op(A)*op(B) + beta*C.  Block 64x64, 4 waves (2x2), each wave
// a 32x32 sub-tile made of 2x2 MFMA 16x16x4 tiles.  Operand layout of
// v_mfma_*_16x16x4: lane l supplies A[l%16][l/16] and B[l/16][l%16]; the
// accumulator element v of lane l is D[MF<T>::row(l, v)][l%16].
// Split-K: blockIdx.z selects K range [z*kchunk, ...); with gridDim.z > 1 the
// raw partial product goes to P + z*M*N (ld M) and k_splitk_reduce finishes.
template <class T, class TO, bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, int kchunk, T alpha,
                                              const T *__restrict__ A, long lda,
                                              const T *__restrict__ B, long ldb, T beta,
                                              TO *__restrict__ C, long ldc, T *__restrict__ P) {
  __shared__ T As[GBK][GBM + 4];
  __shared__ T Bs[GBK][GBN + 4];
  typedef typename MF<T>::V V;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  V acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = V{0, 0, 0, 0};

  for (int k0 = kbeg; k0 < kend; k0 += GBK) {
    if (!TA) {
      const int mm = tid & 63, kk = tid >> 6;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = kk + 4 * i, gm = m0 + mm, gk = k0 + k;
        As[k][mm] = (gm < M && gk < kend) ? A[gm + (long)gk * lda] : T(0);
      }
    } else {
      const int kk = tid & 15, mm = tid >> 4;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = mm + 16 * i, gm = m0 + m, gk = k0 + kk;
        As[kk][m] = (gm < M && gk < kend) ? A[gk + (long)gm * lda] : T(0);
      }
    }
    if (!TB) {
      const int kk = tid & 15, nn = tid >> 4;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int n = nn + 16 * i, gn = n0 + n, gk = k0 + kk;
        Bs[kk][n] = (gn < N && gk < kend) ? B[gk + (long)gn * ldb] : T(0);
      }
    } else {
      const int nn = tid & 63, kk = tid >> 6;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = kk + 4 * i, gn = n0 + nn, gk = k0 + k;
        Bs[k][nn] = (gn < N && gk < kend) ? B[gn + (long)gk * ldb] : T(0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      T a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kk + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kk + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = MF<T>::op(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  const bool split = gridDim.z > 1;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + wm * 32 + i * 16 + MF<T>::row(lane, v);
        const int gn = n0 + wn * 32 + j * 16 + (lane & 15);
        if (gm < M && gn < N) {
          if (split) {
            P[(long)blockIdx.z * M * N + gm + (long)gn * M] = acc[i][j][v];
          } else {
            T val = alpha * acc[i][j][v];
            if (beta != T(0)) val += beta * T(C[gm + (long)gn * ldc]);
            C[gm + (long)gn * ldc] = TO(val);
          }
        }
      }
}

// Fixed-order sum of the split-K partials (accumulated in double).
template <class T, class TO>
__global__ void k_splitk_reduce(int M, int N, int S, const T *__restrict__ P, double alpha,
                                double beta, TO *__restrict__ C, long ldc) {
  const long total = (long)M * N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < S; ++z) s += (double)P[(long)z * total + idx];
    const int gm = (int)(idx % M);
    const long gn = idx / M;
    double v = alpha * s;
    if (beta != 0.0) v += beta * (double)C[gm + gn * ldc];
    C[gm + gn * ldc] = (TO)v;
  }
}

static int splitk_count(int64_t M, int64_t N, int64_t K, int num_cus) {
  const int64_t tiles = ((M + GBM - 1) / GBM) * ((N + GBN - 1) / GBN);
  if (tiles >= num_cus || K < 512) return 1;
  int64_t s = (num_cus + tiles - 1) / tiles;
  s = std::min<int64_t>(s, K / 256);
  s = std::min<int64_t>(s, 32);
  return (int)std::max<int64_t>(s, 1);
}

int64_t gemm_ws_elems(int64_t M, int64_t N, int64_t K, int num_cus) {
  const int s = splitk_count(M, N, K, num_cus);
  return s > 1 ? (int64_t)s * M * N : 0;
}

template <class T, class TO>
static void gemm_impl(hipStream_t st, bool ta, bool tb, int64_t M, int64_t N, int64_t K, T alpha,
                      const T *A, int64_t lda, const T *B, int64_t ldb, T beta, TO *C, int64_t ldc,
                      T *ws, int64_t ws_elems, int num_cus) {
  if (M <= 0 || N <= 0) return;
  int S = splitk_count(M, N, K, num_cus);
  if (S > 1 && (ws == nullptr || (int64_t)S * M * N > ws_elems)) S = 1;
  int kchunk = (int)K;
  if (S > 1) {
    kchunk = (int)(((K + S - 1) / S + GBK - 1) / GBK * GBK);
    S = (int)((K + kchunk - 1) / kchunk);
  }
  if (K <= 0) kchunk = 0;
  dim3 grid((unsigned)((M + GBM - 1) / GBM), (unsigned)((N + GBN - 1) / GBN), (unsigned)S);
  ProfScope prof(KC_GEMM, st, 2.0 * (double)M * (double)N * (double)K,
                 (double)sizeof(T) * ((double)M * K + (double)K * N) + (double)sizeof(TO) * M * N);
#define TTG_GEMM_LAUNCH(TA_, TB_)                                                            \
  hipLaunchKernelGGL((k_gemm<T, TO, TA_, TB_>), grid, dim3(256), 0, st, (int)M, (int)N, (int)K, \
                     kchunk, alpha, A, (long)lda, B, (long)ldb, beta, C, (long)ldc, ws)
  if (!ta && !tb) TTG_GEMM_LAUNCH(false, false);
  else if (ta && !tb) TTG_GEMM_LAUNCH(true, false);
  else if (!ta && tb) TTG_GEMM_LAUNCH(false, true);
  else TTG_GEMM_LAUNCH(true, true);
#undef TTG_GEMM_LAUNCH
  TTG_HIP(hipGetLastError());
  if (S > 1) {
    const long total = (long)M * N;
    const int blocks = (int)std::min<long>((total + 255) / 256, 4096);
    hipLaunchKernelGGL((k_splitk_reduce<T, TO>), dim3(blocks), dim3(256), 0, st, (int)M, (int)N,
                       S, (const T *)ws, (double)alpha, (double)beta, C, (long)ldc);
    TTG_HIP(hipGetLastError());
  }
}

template <>
void gemm<double>(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K, double alpha,
                  const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
                  double *C, int64_t ldc, double *ws, int64_t wse, int cus) {
  gemm_impl<double, double>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, ws, wse, cus);
}
template <>
void gemm<float>(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K, float alpha,
                 const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
                 int64_t ldc, float *ws, int64_t wse, int cus) {
  gemm_impl<float, float>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, ws, wse, cus);
}
// FP32 operands, FP64 result (Gram matrices of the mixed form).
void gemm_f32_f64out(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K,
                     const float *A, int64_t lda, const float *B, int64_t ldb, double *C,
                     int64_t ldc, float *ws, int64_t wse, int cus) {
  gemm_impl<float, double>(s, ta, tb, M, N, K, 1.0f, A, lda, B, ldb, 0.0f, C, ldc, ws, wse, cus);
}

// ---------------------------------------------------------------------------
// elementwise helpers
// ---------------------------------------------------------------------------
template <class TO, class TI>
__global__ void k_convert(TO *y, const TI *x, long count) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
    y[i] = (TO)x[i];
}
template <class TO, class TI> void convert(hipStream_t s, TO *y, const TI *x, int64_t count) {
  if (count <= 0) return;
  const int blocks = (int)std::min<int64_t>((count + 255) / 256, 8192);
  hipLaunchKernelGGL((k_convert<TO, TI>), dim3(blocks), dim3(256), 0, s, y, x, (long)count);
  TTG_HIP(hipGetLastError());
}
template void convert<float, double>(hipStream_t, float *, const double *, int64_t);
template void convert<double, float>(hipStream_t, double *, const float *, int64_t);
template void convert<double, double>(hipStream_t, double *, const double *, int64_t);
template void convert<float, float>(hipStream_t, float *, const float *, int64_t);

// B (cols x rows) = A^T, A rows x cols; 32x32 LDS tiles.
template <class T>
__global__ void k_transpose(const T *__restrict__ A, int rows, int cols, T *__restrict__ B) {
  __shared__ T tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: ty 0..7
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + tx, c = c0 + j;
    if (r < rows && c < cols) tile[j][tx] = A[r + (long)c * rows];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + tx, r = r0 + j;
    if (r < rows && c < cols) B[c + (long)r * cols] = tile[tx][j];
  }
}
template <class T> void transpose(hipStream_t s, const T *A, int64_t rows, int64_t cols, T *B) {
  if (rows <= 0 || cols <= 0) return;
  dim3 grid((unsigned)((rows + 31) / 32), (unsigned)((cols + 31) / 32));
  hipLaunchKernelGGL((k_transpose<T>), grid, dim3(256), 0, s, A, (int)rows, (int)cols, B);
  TTG_HIP(hipGetLastError());
}
template void transpose<double>(hipStream_t, const double *, int64_t, int64_t, double *);
template void transpose<float>(hipStream_t, const float *, int64_t, int64_t, float *);

// Frobenius norm in one block (count is at most a few million here).
template <class T>
__global__ __launch_bounds__(1024) void k_fro_norm(const T *__restrict__ x, long count, double *out) {
  __shared__ double red[16];
  double s = 0.0;
  for (long i = threadIdx.x; i < count; i += blockDim.x) {
    const double v = (double)x[i];
    s += v * v;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = sqrt(s);
}
template <class T> void fro_norm(hipStream_t s, const T *x, int64_t count, double *out) {
  hipLaunchKernelGGL((k_fro_norm<T>), dim3(1), dim3(1024), 0, s, x, (long)count, out);
  TTG_HIP(hipGetLastError());
}
template void fro_norm<double>(hipStream_t, const double *, int64_t, double *);
template void fro_norm<float>(hipStream_t, const float *, int64_t, double *);

template <class T>
__global__ void k_scale_by_inv(T *x, long count, const double *nrm) {
  const double v = nrm[0];
  if (v == 0.0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
    x[i] = (T)((double)x[i] / v);
}
template <class T> void scale_by_inv(hipStream_t s, T *x, int64_t count, const double *nrm) {
  if (count <= 0) return;
  const int blocks = (int)std::min<int64_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL((k_scale_by_inv<T>), dim3(blocks), dim3(256), 0, s, x, (long)count, nrm);
  TTG_HIP(hipGetLastError());
}
template void scale_by_inv<double>(hipStream_t, double *, int64_t, const double *);
template void scale_by_inv<float>(hipStream_t, float *, int64_t, const double *);

__global__ void k_check_finite(const double *x, long count, int *flag) {
  int bad = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
    if (!isfinite(x[i])) bad = 1;
  if (bad) atomicOr(flag, 1);
}
void check_finite(hipStream_t s, const double *x, int64_t count, int *flag) {
  if (count <= 0) return;
  const int blocks = (int)std::min<int64_t>((count + 255) / 256, 2048);
  hipLaunchKernelGGL(k_check_finite, dim3(blocks), dim3(256), 0, s, x, (long)count, flag);
  TTG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Householder QR, one workgroup (robust path: rank-deficient / wide / small)
// ---------------------------------------------------------------------------
// LAPACK dlarfg: beta = -sign(alpha)*||x||, tau = (beta-alpha)/beta,
// v = x/(alpha-beta) with v0 = 1; tau = 0 when the sub-column is zero.
__global__ __launch_bounds__(1024) void k_householder_qr(int m, int n, double *__restrict__ A,
                                                         double *__restrict__ Q,
                                                         double *__restrict__ R,
                                                         double *__restrict__ tau) {
  __shared__ double red[16];
  __shared__ double sh_tau, sh_scale;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  const int kq = min(m, n);
  for (int j = 0; j < kq; ++j) {
    double *aj = A + (long)j * m;
    double ss = 0.0;
    for (int i = j + 1 + tid; i < m; i += blockDim.x) ss += aj[i] * aj[i];
    ss = block_sum(ss, red);
    if (tid == 0) {
      const double alpha = aj[j];
      if (ss == 0.0) {
        sh_tau = 0.0;
        sh_scale = 0.0;
      } else {
        const double nrm = sqrt(alpha * alpha + ss);
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        sh_tau = (beta - alpha) / beta;
        sh_scale = 1.0 / (alpha - beta);
        aj[j] = beta;
      }
      tau[j] = sh_tau;
    }
    __syncthreads();
    const double tj = sh_tau, sc = sh_scale;
    if (tj != 0.0)
      for (int i = j + 1 + tid; i < m; i += blockDim.x) aj[i] *= sc;
    __syncthreads();
    if (tj != 0.0) {
      for (int c = j + 1 + wv; c < n; c += nw) {
        double *ac = A + (long)c * m;
        double w = (lane == 0) ? ac[j] : 0.0;
        for (int i = j + 1 + lane; i < m; i += 64) w += aj[i] * ac[i];
        w = wave_sum(w) * tj;
        if (lane == 0) ac[j] -= w;
        for (int i = j + 1 + lane; i < m; i += 64) ac[i] -= w * aj[i];
      }
    }
    __syncthreads();
  }
  // R = upper trapezoid (kq x n)
  for (long idx = tid; idx < (long)kq * n; idx += blockDim.x) {
    const int i = (int)(idx % kq), c = (int)(idx / kq);
    R[idx] = (i <= c) ? A[i + (long)c * m] : 0.0;
  }
  __syncthreads();
  // Q = H_0 ... H_{kq-1} I(:,0:kq)  (dorg2r backward accumulation)
  for (int j = kq - 1; j >= 0; --j) {
    const double tj = tau[j];
    const double *v = A + (long)j * m;  // v[j] = 1 implicit, v[j+1..] stored
    if (j < kq - 1 && tj != 0.0) {
      for (int c = j + 1 + wv; c < kq; c += nw) {
        double *qc = Q + (long)c * m;
        double w = (lane == 0) ? qc[j] : 0.0;
        for (int i = j + 1 + lane; i < m; i += 64) w += v[i] * qc[i];
        w = wave_sum(w) * tj;
        if (lane == 0) qc[j] -= w;
        for (int i = j + 1 + lane; i < m; i += 64) qc[i] -= w * v[i];
      }
    }
    __syncthreads();
    double *qj = Q + (long)j * m;
    for (int i = tid; i < m; i += blockDim.x) {
      double val;
      if (i < j) val = 0.0;
      else if (i == j) val = 1.0 - tj;
      else val = -tj * v[i];
      qj[i] = val;
    }
    __syncthreads();
  }
}
void householder_qr(hipStream_t s, int64_t m, int64_t n, double *A, double *Q, double *R,
                    double *tau) {
  const double kq = (double)std::min(m, n);
  ProfScope prof(KC_HQR, s, 4.0 * m * n * kq - 2.0 * (m + n) * kq * kq + 4.0 / 3.0 * kq * kq * kq,
                 8.0 * ((double)m * n + m * kq + kq * n));
  hipLaunchKernelGGL(k_householder_qr, dim3(1), dim3(1024), 0, s, (int)m, (int)n, A, Q, R, tau);
  TTG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Cholesky G + shift*trace(G)/k*I = R^T R (upper R), one workgroup, nb = 32.
// stats[0] = max|r_jj| / min|r_jj|; stats[1] = max_ij |R_ij - delta_ij|.
// ---------------------------------------------------------------------------
constexpr int CNB = 32;
constexpr int CHOL_MAXK = 512;
__global__ __launch_bounds__(1024) void k_chol(int k, const double *__restrict__ G, long ldg,
                                               double shift_rel, double *__restrict__ R,
                                               int *flag, double *stats) {
  __shared__ double D[CNB][CNB + 1];
  __shared__ double red[16];
  __shared__ int sh_fail;
  extern __shared__ double Ps[];  // CNB x (k) panel
  const int tid = threadIdx.x;
  double tr = 0.0;
  for (int j = tid; j < k; j += blockDim.x) tr += G[j + (long)j * ldg];
  tr = block_sum(tr, red);
  const double shift = shift_rel * tr / (k > 0 ? k : 1);
  for (long idx = tid; idx < (long)k * k; idx += blockDim.x) {
    const int i = (int)(idx % k), j = (int)(idx / k);
    double v = 0.0;
    if (i <= j) v = G[i + (long)j * ldg] + (i == j ? shift : 0.0);
    R[idx] = v;
  }
  if (tid == 0) sh_fail = 0;
  __syncthreads();
  for (int J0 = 0; J0 < k; J0 += CNB) {
    const int jb = min(CNB, k - J0);
    for (int idx = tid; idx < CNB * CNB; idx += blockDim.x) {
      const int i = idx % CNB, j = idx / CNB;
      D[i][j] = (i < jb && j < jb && i <= j) ? R[(J0 + i) + (long)(J0 + j) * k] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < jb; ++c) {
      if (tid == 0) {
        const double p = D[c][c];
        if (!(p > 0.0) || !isfinite(p)) {
          sh_fail = J0 + c + 1;
        } else {
          D[c][c] = sqrt(p);
        }
      }
      __syncthreads();
      if (sh_fail) break;
      const double dc = D[c][c];
      if (tid > c && tid < jb) D[c][tid] /= dc;
      __syncthreads();
      for (int idx = tid; idx < CNB * CNB; idx += blockDim.x) {
        const int i = idx % CNB, j = idx / CNB;
        if (i > c && i <= j && j < jb) D[i][j] -= D[c][i] * D[c][j];
      }
      __syncthreads();
    }
    if (sh_fail) break;
    for (int idx = tid; idx < jb * jb; idx += blockDim.x) {
      const int i = idx % jb, j = idx / jb;
      if (i <= j) R[(J0 + i) + (long)(J0 + j) * k] = D[i][j];
    }
    const int w = k - J0 - jb;
    if (w > 0) {
      // panel: solve D^T P = R[J0:J0+jb, J0+jb:k]  (forward substitution per column)
      for (int c = tid; c < w; c += blockDim.x) {
        const long col = (long)(J0 + jb + c) * k;
        for (int t = 0; t < jb; ++t) {
          double v = R[J0 + t + col];
          for (int l = 0; l < t; ++l) v -= D[l][t] * Ps[l * w + c];
          Ps[t * w + c] = v / D[t][t];
        }
      }
      __syncthreads();
      for (int idx = tid; idx < jb * w; idx += blockDim.x) {
        const int t = idx % jb, c = idx / jb;
        R[J0 + t + (long)(J0 + jb + c) * k] = Ps[t * w + c];
      }
      // trailing update of the upper triangle
      const long tri = (long)w * (w + 1) / 2;
      for (long idx = tid; idx < tri; idx += blockDim.x) {
        // map idx -> (i, j) with i <= j, column-major over the upper triangle
        int j = (int)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
        while ((long)j * (j + 1) / 2 > idx) --j;
        while ((long)(j + 1) * (j + 2) / 2 <= idx) ++j;
        const int i = (int)(idx - (long)j * (j + 1) / 2);
        double s = 0.0;
        for (int t = 0; t < jb; ++t) s += Ps[t * w + i] * Ps[t * w + j];
        R[(J0 + jb + i) + (long)(J0 + jb + j) * k] -= s;
      }
    }
    __syncthreads();
  }
  __syncthreads();
  double dmax = 0.0, dmin = INFINITY, dev = 0.0;
  if (!sh_fail) {
    for (int j = tid; j < k; j += blockDim.x) {
      const double v = fabs(R[j + (long)j * k]);
      dmax = fmax(dmax, v);
      dmin = fmin(dmin, v);
    }
    for (long idx = tid; idx < (long)k * k; idx += blockDim.x) {
      const int i = (int)(idx % k), j = (int)(idx / k);
      dev = fmax(dev, fabs(R[idx] - (i == j ? 1.0 : 0.0)));
    }
  }
  dmax = block_max(dmax, red);
  dmin = -block_max(-dmin, red);
  dev = block_max(dev, red);
  if (tid == 0) {
    flag[0] = sh_fail;
    stats[0] = sh_fail ? INFINITY : (dmin > 0.0 ? dmax / dmin : INFINITY);
    stats[1] = sh_fail ? INFINITY : dev;
  }
}

// Upper-triangular inverse: one wave per column c solves R x = e_c by
// column-oriented back substitution (x held in registers, 8 rows per lane).
constexpr int TRI_ROWS_PER_LANE = CHOL_MAXK / 64;
__global__ __launch_bounds__(256) void k_trtri(int k, const double *__restrict__ R,
                                               double *__restrict__ X) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (c >= k) return;
  double x[TRI_ROWS_PER_LANE];
#pragma unroll
  for (int q = 0; q < TRI_ROWS_PER_LANE; ++q) {
    const int i = lane + 64 * q;
    x[q] = (i == c) ? 1.0 : 0.0;
  }
  for (int j = c; j >= 0; --j) {
    const double *rj = R + (long)j * k;
    const double rjj = rj[j];
    // broadcast x[j] from its owner lane
    const int q = j >> 6, owner = j & 63;
    double bj = 0.0;
#pragma unroll
    for (int qq = 0; qq < TRI_ROWS_PER_LANE; ++qq)
      if (qq == q) bj = x[qq];
    bj = __shfl(bj, owner, 64);
    const double xj = bj / rjj;
#pragma unroll
    for (int qq = 0; qq < TRI_ROWS_PER_LANE; ++qq) {
      const int i = lane + 64 * qq;
      if (i < j) x[qq] -= xj * rj[i];
      else if (i == j) x[qq] = xj;
    }
  }
#pragma unroll
  for (int q = 0; q < TRI_ROWS_PER_LANE; ++q) {
    const int i = lane + 64 * q;
    if (i < k) X[i + (long)c * k] = (i <= c) ? x[q] : 0.0;
  }
}

void chol_factor(hipStream_t s, int64_t k, const double *G, int64_t ldg, double shift_rel,
                 double *R, int *flag, double *stats) {
  TTG_CHECK(k <= CHOL_MAXK, TTG_BAD_ARG, "rank above 512 not supported by the Cholesky kernel");
  const size_t lds = (size_t)CNB * (size_t)std::max<int64_t>(k, 1) * sizeof(double);
  ProfScope prof(KC_CHOL, s, (double)k * k * k / 3.0, 8.0 * 2.0 * (double)k * k);
  TTG_HIP(hipFuncSetAttribute((const void *)k_chol, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_chol, dim3(1), dim3(1024), lds, s, (int)k, G, (long)ldg, shift_rel, R,
                     flag, stats);
  TTG_HIP(hipGetLastError());
}
void tri_inverse(hipStream_t s, int64_t k, const double *R, double *X) {
  TTG_CHECK(k <= CHOL_MAXK, TTG_BAD_ARG, "rank above 512 not supported by the triangular inverse");
  const int waves = 4;
  ProfScope prof(KC_TRTRI, s, (double)k * k * k / 3.0, 8.0 * 2.0 * (double)k * k);
  hipLaunchKernelGGL(k_trtri, dim3((unsigned)((k + waves - 1) / waves)), dim3(64 * waves), 0, s,
                     (int)k, R, X);
  TTG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// One-sided block Jacobi SVD (cooperative; p workgroups, 2p column blocks)
// ---------------------------------------------------------------------------
// Y is rows x kpad (rows = 2k): rows [0,k) hold the matrix S being
// orthogonalised, rows [k,2k) accumulate the rotations (start = identity).
// Each outer step gives workgroup w the block pair (round-robin "circle"
// schedule over 2p blocks); the 2b columns sit in LDS while one inner sweep
// of scalar rotations runs (b pairs per round, one wave per pair).  A sweep
// is 2p-1 outer steps; iteration stops after a sweep without rotations.
__device__ __forceinline__ void circle_pair(int players, int round, int idx, int &a, int &b) {
  const int P1 = players - 1;
  if (idx == 0) {
    a = P1;
    b = round;
  } else {
    a = (round + idx) % P1;
    b = (round - idx + P1) % P1;
  }
}

// Software grid barrier for a cooperative launch.  bar[0] = arrivals,
// bar[1] = generation.  Bounded spin: after ~4 s a timeout flag is raised
// and every workgroup leaves the kernel at its next check.
__device__ bool grid_sync(unsigned *bar, unsigned nblocks, unsigned &gen, int *err) {
  __syncthreads();
  __shared__ int sh_ok;
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned target = gen + 1;
    const unsigned prev = atomicAdd(&bar[0], 1u);
    int ok = 1;
    if (prev == nblocks - 1) {
      atomicExch(&bar[0], 0u);
      __threadfence();
      atomicExch(&bar[1], target);
    } else {
      const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != target) {
        __builtin_amdgcn_s_sleep(2);
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          ok = 0;
          break;
        }
        if (__builtin_amdgcn_s_memrealtime() - t0 > 400000000ull) {  // 100 MHz clock: 4 s
          atomicExch(err, 2);
          ok = 0;
          break;
        }
      }
    }
    __threadfence();
    gen = target;
    sh_ok = ok;
  }
  __syncthreads();
  return sh_ok != 0;
}

__global__ __launch_bounds__(1024) void k_jacobi_setup(int k, int kpad, const double *__restrict__ S,
                                                       int transS, double *__restrict__ Y) {
  const int rows = 2 * k;
  const long total = (long)rows * kpad;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % rows), j = (int)(idx / rows);
    double v = 0.0;
    if (j < k) {
      if (i < k) v = transS ? S[j + (long)i * k] : S[i + (long)j * k];
      else v = (i - k == j) ? 1.0 : 0.0;
    }
    Y[idx] = v;
  }
}

__global__ __launch_bounds__(1024) void k_jacobi(int k, int b, int p, double *__restrict__ Y,
                                                 double tol, int max_sweeps, int *sweeps_out,
                                                 unsigned *bar, int *err) {
  extern __shared__ double L[];  // 2b columns x rows
  const int rows = 2 * k;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  const int w = blockIdx.x;
  const int nblk = 2 * p;
  const int lc = 2 * b;
  __shared__ int sh_rot;
  __shared__ int sh_total;
  unsigned gen = 0;
  int sweep = 0;
  bool converged = false;
  for (sweep = 0; sweep < max_sweeps && !converged; ++sweep) {
    if (tid == 0) sh_rot = 0;
    __syncthreads();
    const int outer = (nblk > 1) ? nblk - 1 : 1;
    for (int step = 0; step < outer; ++step) {
      int bA = 0, bB = 1;
      if (nblk > 1) circle_pair(nblk, step, w, bA, bB);
      // load the two blocks
      for (long idx = tid; idx < (long)lc * rows; idx += blockDim.x) {
        const int c = (int)(idx / rows), i = (int)(idx % rows);
        const int gcol = (c < b) ? bA * b + c : bB * b + (c - b);
        L[idx] = Y[i + (long)gcol * rows];
      }
      __syncthreads();
      // one inner sweep over the 2b local columns
      for (int t = 0; t < lc - 1; ++t) {
        for (int pr = wv; pr < b; pr += nw) {
          int pa, pb;
          circle_pair(lc, t, pr, pa, pb);
          double *cp = L + (long)pa * rows;
          double *cq = L + (long)pb * rows;
          double al = 0.0, be = 0.0, ga = 0.0;
          for (int i = lane; i < k; i += 64) {
            const double x = cp[i], y = cq[i];
            al += x * x;
            be += y * y;
            ga += x * y;
          }
          al = wave_sum(al);
          be = wave_sum(be);
          ga = wave_sum(ga);
          if (fabs(ga) > tol * sqrt(al) * sqrt(be)) {
            const double zeta = (be - al) / (2.0 * ga);
            double tt;
            if (fabs(zeta) > 1e150) tt = 0.5 / zeta;
            else tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
            for (int i = lane; i < rows; i += 64) {
              const double x = cp[i], y = cq[i];
              cp[i] = cs * x - sn * y;
              cq[i] = sn * x + cs * y;
            }
            if (lane == 0) atomicAdd(&sh_rot, 1);
          }
        }
        __syncthreads();
      }
      // store back
      for (long idx = tid; idx < (long)lc * rows; idx += blockDim.x) {
        const int c = (int)(idx / rows), i = (int)(idx % rows);
        const int gcol = (c < b) ? bA * b + c : bB * b + (c - b);
        Y[i + (long)gcol * rows] = L[idx];
      }
      if (p > 1) {
        if (!grid_sync(bar, p, gen, err)) return;
      } else {
        __syncthreads();
      }
    }
    // sweep-level convergence: sum rotation counts over workgroups
    if (p > 1) {
      if (tid == 0) atomicAdd(&bar[2 + sweep], (unsigned)sh_rot);
      if (!grid_sync(bar, p, gen, err)) return;
      if (tid == 0)
        sh_total = (int)__hip_atomic_load(&bar[2 + sweep], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
    } else {
      if (tid == 0) sh_total = sh_rot;
      __syncthreads();
    }
    converged = (sh_total == 0);
  }
  if (w == 0 && tid == 0) {
    sweeps_out[0] = sweep;
    if (!converged) atomicExch(err, 1);
  }
}

JacobiPlan jacobi_plan(int64_t k) {
  JacobiPlan pl;
  pl.k = k;
  pl.rows = 2 * k;
  const size_t budget = 144 * 1024;
  // largest block size that fits LDS, capped at 16 columns per wave-pair slot
  int64_t bmax = (int64_t)(budget / (sizeof(double) * 2 * (size_t)pl.rows));
  if (bmax < 1) bmax = 1;
  int64_t b;
  if (2 * bmax >= k) {
    b = (k + 1) / 2;  // single workgroup
    if (b < 1) b = 1;
  } else {
    b = std::min<int64_t>(bmax, 16);
  }
  pl.b = (int)b;
  pl.p = (int)((k + 2 * b - 1) / (2 * b));
  if (pl.p < 1) pl.p = 1;
  pl.kpad = 2 * (int64_t)pl.b * pl.p;
  return pl;
}
size_t jacobi_ws_bytes(const JacobiPlan &pl) { return (size_t)pl.rows * pl.kpad * sizeof(double); }

void jacobi_svd(hipStream_t s, const JacobiPlan &pl, const double *S, bool transS, double *Y,
                int *sweeps_dev, unsigned *barrier_dev, int *err_dev, int max_sweeps) {
  const int k = (int)pl.k, kpad = (int)pl.kpad;
  {
    const long total = (long)pl.rows * kpad;
    const int blocks = (int)std::min<long>((total + 1023) / 1024, 2048);
    hipLaunchKernelGGL(k_jacobi_setup, dim3(blocks), dim3(1024), 0, s, k, kpad, S, transS ? 1 : 0, Y);
    TTG_HIP(hipGetLastError());
  }
  TTG_HIP(hipMemsetAsync(barrier_dev, 0, sizeof(unsigned) * (2 + max_sweeps), s));
  TTG_HIP(hipMemsetAsync(err_dev, 0, sizeof(int), s));
  // algorithmic flops of the small SVD (R-SVD of the k x k factor, GVL Fig. 8.6.1: 20k^3)
  ProfScope prof(KC_JACOBI, s, 20.0 * (double)k * k * k, 8.0 * 4.0 * (double)k * k);
  const double tol = sqrt((double)std::max(k, 1)) * DBL_EPSILON;
  const size_t lds = (size_t)2 * pl.b * pl.rows * sizeof(double);
  int b = pl.b, p = pl.p;
  int kk = k;
  double *Yp = Y;
  int ms = max_sweeps;
  void *args[] = {&kk, &b, &p, &Yp, (void *)&tol, &ms, &sweeps_dev, &barrier_dev, &err_dev};
  TTG_HIP(hipFuncSetAttribute((const void *)k_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (p > 1) {
    TTG_HIP(hipLaunchCooperativeKernel((const void *)k_jacobi, dim3(p), dim3(1024), args, lds, s));
  } else {
    hipLaunchKernelGGL(k_jacobi, dim3(1), dim3(1024), lds, s, kk, b, p, Yp, tol, ms, sweeps_dev,
                       barrier_dev, err_dev);
    TTG_HIP(hipGetLastError());
  }
}

// ---------------------------------------------------------------------------
// singular values: column norms, descending rank sort, my_chop2
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_sort_chop(const double *__restrict__ Y, long ldy, int k,
                                                    int kpad, double ep, long rmax,
                                                    double *__restrict__ sv, int *__restrict__ perm,
                                                    int *out_i, double *out_d) {
  extern __shared__ double sig[];  // kpad
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  for (int j = wv; j < kpad; j += nw) {
    const double *col = Y + (long)j * ldy;
    double s = 0.0;
    for (int i = lane; i < k; i += 64) s += col[i] * col[i];
    s = wave_sum(s);
    if (lane == 0) sig[j] = sqrt(s);
  }
  __syncthreads();
  for (int j = tid; j < kpad; j += blockDim.x) {
    const double v = sig[j];
    int pos = 0;
    for (int i = 0; i < kpad; ++i) {
      const double u = sig[i];
      pos += (u > v) || (u == v && i < j);
    }
    if (pos < k) {
      sv[pos] = v;
      perm[pos] = j;
    }
  }
  __syncthreads();
  if (tid == 0) {
    // norm(s) and my_chop2(s, norm(s)*ep)   (core/my_chop2.m:18-34)
    double ss = 0.0;
    for (int j = 0; j < k; ++j) ss += sv[j] * sv[j];
    const double nrm = sqrt(ss);
    const double e = nrm * ep;
    long r;
    if (nrm == 0.0) {
      r = 1;
    } else if (!(e > 0.0)) {
      r = k;
    } else {
      const double e2 = e * e;
      double c = 0.0;
      long K = 0;
      for (int j = k - 1; j >= 0; --j) {
        c += sv[j] * sv[j];
        if (c < e2) K = k - j;
      }
      r = k - K;
    }
    if (r > rmax) r = rmax;
    out_i[0] = (int)r;
    out_d[0] = nrm;
  }
}
void sort_chop(hipStream_t s, const double *Y, int64_t ldy, int64_t k, int64_t kpad, double ep,
               int64_t rmax, double *sv, int *perm, int *out_i, double *out_d) {
  hipLaunchKernelGGL(k_sort_chop, dim3(1), dim3(1024), (size_t)kpad * sizeof(double), s, Y,
                     (long)ldy, (int)k, (int)kpad, ep, (long)rmax, sv, perm, out_i, out_d);
  TTG_HIP(hipGetLastError());
}

template <class T>
__global__ void k_gather_cols(const double *__restrict__ Y, long ldy, long row0, long rows,
                              const int *__restrict__ perm, long r1, T *__restrict__ dst, long ldd,
                              int trans) {
  const long total = rows * r1;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long i = idx % rows, c = idx / rows;
    const double v = Y[row0 + i + (long)perm[c] * ldy];
    if (trans) dst[c + i * r1] = (T)v;
    else dst[i + c * ldd] = (T)v;
  }
}
template <class T>
void gather_cols(hipStream_t s, const double *Y, int64_t ldy, int64_t row0, int64_t rows,
                 const int *perm, int64_t r1, T *dst, int64_t ldd) {
  if (rows * r1 <= 0) return;
  const int blocks = (int)std::min<int64_t>((rows * r1 + 255) / 256, 4096);
  hipLaunchKernelGGL((k_gather_cols<T>), dim3(blocks), dim3(256), 0, s, Y, (long)ldy, (long)row0,
                     (long)rows, perm, (long)r1, dst, (long)ldd, 0);
  TTG_HIP(hipGetLastError());
}
template <class T>
void gather_cols_t(hipStream_t s, const double *Y, int64_t ldy, int64_t row0, int64_t rows,
                   const int *perm, int64_t r1, T *dst) {
  if (rows * r1 <= 0) return;
  const int blocks = (int)std::min<int64_t>((rows * r1 + 255) / 256, 4096);
  hipLaunchKernelGGL((k_gather_cols<T>), dim3(blocks), dim3(256), 0, s, Y, (long)ldy, (long)row0,
                     (long)rows, perm, (long)r1, dst, (long)r1, 1);
  TTG_HIP(hipGetLastError());
}
template void gather_cols<double>(hipStream_t, const double *, int64_t, int64_t, int64_t, const int *, int64_t, double *, int64_t);
template void gather_cols<float>(hipStream_t, const double *, int64_t, int64_t, int64_t, const int *, int64_t, float *, int64_t);
template void gather_cols_t<double>(hipStream_t, const double *, int64_t, int64_t, int64_t, const int *, int64_t, double *);
template void gather_cols_t<float>(hipStream_t, const double *, int64_t, int64_t, int64_t, const int *, int64_t, float *);

// ---------------------------------------------------------------------------
// final normalisation + balance (round.m:94-107) and packing
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_balance_factor(int d, double *nrm, double *scales) {
  __shared__ double red[16];
  (void)red;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < d; ++i) s += log(fabs(nrm[i]));
    const double nrm0 = exp(s / d);
    scales[0] = (nrm[0] != 0.0) ? nrm0 / nrm[0] : nrm0;
    scales[1] = nrm0;
  }
}
template <class TI>
__global__ void k_pack(int d, TI *const *src, const long *sizes, const long *ofs,
                       const double *scales, double *out) {
  const int kcore = blockIdx.y;
  const long sz = sizes[kcore];
  const TI *p = src[kcore];
  const double sc = (kcore == 0) ? scales[0] : scales[1];
  double *o = out + ofs[kcore];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < sz; i += (long)gridDim.x * blockDim.x)
    o[i] = sc * (double)p[i];
}
template <class TI>
void finalize_balance(hipStream_t s, int64_t d, TI *const *src_dev, const int64_t *sizes_dev,
                      const int64_t *ofs_dev, int64_t size0, double *nrm_dev, double *out,
                      int64_t total) {
  (void)total;
  // nrm[0] = ||core_0||_F is computed by the caller into nrm_dev[0]
  double *scales = nrm_dev + d;
  hipLaunchKernelGGL(k_balance_factor, dim3(1), dim3(64), 0, s, (int)d, nrm_dev, scales);
  TTG_HIP(hipGetLastError());
  (void)size0;
  hipLaunchKernelGGL((k_pack<TI>), dim3(256, (unsigned)d), dim3(256), 0, s, (int)d, src_dev,
                     (const long *)sizes_dev, (const long *)ofs_dev, scales, out);
  TTG_HIP(hipGetLastError());
}
template void finalize_balance<double>(hipStream_t, int64_t, double *const *, const int64_t *, const int64_t *, int64_t, double *, double *, int64_t);
template void finalize_balance<float>(hipStream_t, int64_t, float *const *, const int64_t *, const int64_t *, int64_t, double *, double *, int64_t);

// ---------------------------------------------------------------------------
// mtimes core (@tt_matrix/mtimes.m:38-65 layout)
// Y[(a + ra0*al) + ra0*rx0*(j + n*(b + ra1*be))] = sum_kk A[a,j,kk,b] X[al,kk,be]
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_mtimes_core(long ra0, long n, long m, long ra1, const double *__restrict__ A,
                              long rx0, long rx1, const double *__restrict__ X, T *__restrict__ Y) {
  const long r0 = ra0 * rx0, r1 = ra1 * rx1;
  const long total = r0 * n * r1;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long row = idx % r0;
    long rest = idx / r0;
    const long j = rest % n;
    const long col = rest / n;
    const long a = row % ra0, al = row / ra0;
    const long bb = col % ra1, be = col / ra1;
    double s = 0.0;
    for (long kk = 0; kk < m; ++kk)
      s += A[a + ra0 * (j + n * (kk + m * bb))] * X[al + rx0 * (kk + m * be)];
    Y[idx] = (T)s;
  }
}
template <class T>
void mtimes_core(hipStream_t s, int64_t ra0, int64_t n, int64_t m, int64_t ra1, const double *A,
                 int64_t rx0, int64_t rx1, const double *X, T *Y) {
  const int64_t total = ra0 * rx0 * n * ra1 * rx1;
  if (total <= 0) return;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 16384);
  hipLaunchKernelGGL((k_mtimes_core<T>), dim3(blocks), dim3(256), 0, s, (long)ra0, (long)n,
                     (long)m, (long)ra1, A, (long)rx0, (long)rx1, X, Y);
  TTG_HIP(hipGetLastError());
}
template void mtimes_core<double>(hipStream_t, int64_t, int64_t, int64_t, int64_t, const double *, int64_t, int64_t, const double *, double *);
template void mtimes_core<float>(hipStream_t, int64_t, int64_t, int64_t, int64_t, const double *, int64_t, int64_t, const double *, float *);

}  // namespace ttg
