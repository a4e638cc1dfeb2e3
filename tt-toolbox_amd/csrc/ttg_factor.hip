// Cholesky factorisation and triangular solve for Cholesky-QR (gfx950).
//
//   k_chol2 : one workgroup, G + shift = R^T R (upper R), blocked by 32.
//             The 32x32 diagonal block is factorised by wave 0 alone (LDS,
//             wave-level ordering, no workgroup barriers), the panel is a
//             forward substitution and the trailing update an MFMA SYRK.
//   k_trsm  : X = A R^{-1} (R upper, right side), 32 rows per workgroup held
//             in LDS; per 32-column block J: B = A_J - X_{<J} R_{<J,J}
//             (MFMA), then row-wise substitution with R_JJ (8 lanes per
//             row).  No inverse of R is formed, so the residual
//             ||X R - A|| stays O(u ||X|| ||R||) for any conditioning.
#include "ttg_internal.h"
#include "ttg_device.h"
#include <cmath>
#include <algorithm>

namespace ttg {

typedef double v4d_f __attribute__((ext_vector_type(4)));
typedef float v4f_f __attribute__((ext_vector_type(4)));

template <class T> struct MFf;
template <> struct MFf<double> {
  typedef v4d_f V;
  static __device__ __forceinline__ V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // accumulator element v of lane l is D[row(l,v)][l%16] (measured on gfx950)
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct MFf<float> {
  typedef v4f_f V;
  static __device__ __forceinline__ V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

constexpr int CB = 32;  // block size

// ---------------------------------------------------------------------------
// Cholesky
// ---------------------------------------------------------------------------
// R (k x k, ld k) receives the upper factor (lower triangle zero).
// stats[0] = max|r_jj| / min|r_jj|, stats[1] = max_ij |R_ij - delta_ij|.
__global__ __launch_bounds__(256) void k_chol2(int k, const double *__restrict__ G, long ldg,
                                               double shift_rel, double *__restrict__ R, int *flag,
                                               double *stats) {
  __shared__ double D[CB][CB + 1];
  __shared__ double red[8];
  __shared__ int sh_fail;
  extern __shared__ double Ps[];  // CB x w panel, Ps[t*w + c]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double tr = 0.0;
  for (int j = tid; j < k; j += blockDim.x) tr += G[j + (long)j * ldg];
  tr = block_sum(tr, red);
  const double shift = shift_rel * tr / (k > 0 ? k : 1);
  for (long idx = tid; idx < (long)k * k; idx += blockDim.x) {
    const int i = (int)(idx % k), j = (int)(idx / k);
    R[idx] = (i <= j) ? G[i + (long)j * ldg] + (i == j ? shift : 0.0) : 0.0;
  }
  if (tid == 0) sh_fail = 0;
  __syncthreads();
  for (int J0 = 0; J0 < k; J0 += CB) {
    const int jb = min(CB, k - J0);
    for (int idx = tid; idx < CB * CB; idx += blockDim.x) {
      const int i = idx % CB, j = idx / CB;
      D[i][j] = (i < jb && j < jb && i <= j) ? R[(J0 + i) + (long)(J0 + j) * k] : (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    if (wv == 0) {
      // wave 0: unblocked upper Cholesky of the diagonal block
      int fail = 0;
      for (int c = 0; c < jb; ++c) {
        const double p = D[c][c];
        if (!(p > 0.0) || !isfinite(p)) {
          fail = c + 1;
          break;
        }
        const double dc = sqrt(p);
        wave_sync_lds();
        if (lane == 0) D[c][c] = dc;
        if (lane > c && lane < jb) D[c][lane] /= dc;
        wave_sync_lds();
        const int rem = jb - c - 1;
        for (int e = lane; e < rem * rem; e += 64) {
          const int i = c + 1 + e % rem, j = c + 1 + e / rem;
          if (i <= j) D[i][j] -= D[c][i] * D[c][j];
        }
        wave_sync_lds();
      }
      if (fail && lane == 0) sh_fail = J0 + fail;
    }
    __syncthreads();
    if (sh_fail) break;
    for (int idx = tid; idx < jb * jb; idx += blockDim.x) {
      const int i = idx % jb, j = idx / jb;
      if (i <= j) R[(J0 + i) + (long)(J0 + j) * k] = D[i][j];
    }
    const int w = k - J0 - jb;
    if (w > 0) {
      // panel: D^T P = R[J0:J0+jb, J0+jb:k], one column per thread, P kept in LDS
      for (int c = tid; c < w; c += blockDim.x) {
        const long col = (long)(J0 + jb + c) * k;
        for (int t = 0; t < jb; ++t) {
          double v = R[J0 + t + col];
          for (int l = 0; l < t; ++l) v -= D[l][t] * Ps[l * w + c];
          Ps[t * w + c] = v / D[t][t];
        }
        for (int t = jb; t < CB; ++t) Ps[t * w + c] = 0.0;
      }
      __syncthreads();
      for (int idx = tid; idx < jb * w; idx += blockDim.x) {
        const int t = idx % jb, c = idx / jb;
        R[J0 + t + (long)(J0 + jb + c) * k] = Ps[t * w + c];
      }
      // trailing SYRK on 16x16 tiles of the upper triangle (MFMA FP64)
      const int nt = (w + 15) / 16;
      const int ntiles = nt * (nt + 1) / 2;
      for (int tile = wv; tile < ntiles; tile += 4) {
        int tj = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
        while (tj * (tj + 1) / 2 > tile) --tj;
        while ((tj + 1) * (tj + 2) / 2 <= tile) ++tj;
        const int ti = tile - tj * (tj + 1) / 2;
        v4d_f acc = {0, 0, 0, 0};
        const int ia = ti * 16 + (lane & 15), jbcol = tj * 16 + (lane & 15);
#pragma unroll
        for (int kk = 0; kk < CB; kk += 4) {
          const int t = kk + (lane >> 4);
          const double a = (ia < w) ? Ps[t * w + ia] : 0.0;
          const double b = (jbcol < w) ? Ps[t * w + jbcol] : 0.0;
          acc = MFf<double>::op(a, b, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = ti * 16 + MFf<double>::row(lane, v);
          const int j = tj * 16 + (lane & 15);
          if (i < w && j < w && i <= j) R[(J0 + jb + i) + (long)(J0 + jb + j) * k] -= acc[v];
        }
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

constexpr int CHOL2_MAXK = 512;

void chol_factor(hipStream_t s, int64_t k, const double *G, int64_t ldg, double shift_rel,
                 double *R, int *flag, double *stats) {
  TTG_CHECK(k <= CHOL2_MAXK, TTG_BAD_ARG, "rank above 512 not supported by the Cholesky kernel");
  const size_t lds = (size_t)CB * (size_t)std::max<int64_t>(k, 1) * sizeof(double);
  ProfScope prof(KC_CHOL, s, (double)k * k * k / 3.0, 8.0 * 2.0 * (double)k * k);
  TTG_HIP(hipFuncSetAttribute((const void *)k_chol2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_chol2, dim3(1), dim3(256), lds, s, (int)k, G, (long)ldg, shift_rel, R, flag, stats);
  TTG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// TRSM: X (m x k) = A R^{-1}
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_trsm(int m, int k, const T *__restrict__ A, long lda,
                                              const double *__restrict__ R, T *__restrict__ X,
                                              long ldx) {
  extern __shared__ char smem[];
  typedef typename MFf<T>::V V;
  const int kp = (k + CB - 1) / CB * CB;
  T *Xs = reinterpret_cast<T *>(smem);   // 32 x kp, column-major (Xs[col*32 + row])
  T *Rs = Xs + (size_t)CB * kp;           // 32 x 32 chunk of R, row-major (Rs[t*32 + n])
  double *Ds = reinterpret_cast<double *>(Rs + CB * CB);  // R_JJ, row-major (Ds[l*32 + j])
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tr = wv & 1, tc = wv >> 1;
  const int r0 = blockIdx.x * CB;
  const int srow = tid >> 3, sq = tid & 7;  // substitution: 8 lanes per row
  for (int idx = tid; idx < CB * kp; idx += blockDim.x) {
    const int row = idx % CB, col = idx / CB;
    const int gr = r0 + row;
    Xs[idx] = (gr < m && col < k) ? A[gr + (long)col * lda] : T(0);
  }
  for (int J0 = 0; J0 < k; J0 += CB) {
    const int jb = min(CB, k - J0);
    V acc = {0, 0, 0, 0};
    for (int kc = 0; kc < J0; kc += CB) {
      __syncthreads();
      for (int idx = tid; idx < CB * CB; idx += blockDim.x) {
        const int t = idx % CB, n = idx / CB;
        Rs[t * CB + n] = (n < jb) ? (T)R[(kc + t) + (long)(J0 + n) * k] : T(0);
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < CB; kk += 4) {
        const int t = kk + (lane >> 4);
        const T a = Xs[(kc + t) * CB + tr * 16 + (lane & 15)];
        const T b = Rs[t * CB + tc * 16 + (lane & 15)];
        acc = MFf<T>::op(a, b, acc);
      }
    }
    __syncthreads();
    // B = A_J - X_{<J} R_{<J,J} (in place); stage R_JJ
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = tr * 16 + MFf<T>::row(lane, v), col = tc * 16 + (lane & 15);
      Xs[(J0 + col) * CB + row] -= acc[v];
    }
    for (int idx = tid; idx < CB * CB; idx += blockDim.x) {
      const int l = idx % CB, j = idx / CB;  // column-major source R(J0+l, J0+j)
      Ds[l * CB + j] = (l < jb && j < jb) ? R[(J0 + l) + (long)(J0 + j) * k] : (l == j ? 1.0 : 0.0);
    }
    __syncthreads();
    // X_J R_JJ = B: row srow, lane sq owns columns sq + 8s
    double bb[4], xx[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      bb[s4] = (double)Xs[(J0 + sq + 8 * s4) * CB + srow];
      xx[s4] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      double part = 0.0;
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4)
        if (sq + 8 * s4 < j) part += xx[s4] * Ds[(sq + 8 * s4) * CB + j];
      part += dpp_d<0xB1>(part);
      part += dpp_d<0x4E>(part);
      part += dpp_d<0x141>(part);
      if (sq == (j & 7)) xx[j >> 3] = (bb[j >> 3] - part) / Ds[j * CB + j];
    }
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
      if (sq + 8 * s4 < jb) Xs[(J0 + sq + 8 * s4) * CB + srow] = (T)xx[s4];
    __syncthreads();
  }
  for (int idx = tid; idx < CB * k; idx += blockDim.x) {
    const int row = idx % CB, col = idx / CB;
    const int gr = r0 + row;
    if (gr < m) X[gr + (long)col * ldx] = Xs[col * CB + row];
  }
}

template <class T>
void trsm_right_upper(hipStream_t s, int64_t m, int64_t k, const T *A, int64_t lda, const double *R,
                      T *X, int64_t ldx) {
  if (m <= 0 || k <= 0) return;
  const int64_t kp = (k + CB - 1) / CB * CB;
  const size_t lds = ((size_t)CB * kp + CB * CB) * sizeof(T) + CB * CB * sizeof(double);
  TTG_CHECK(lds <= 160 * 1024, TTG_BAD_ARG, "rank too large for the triangular solve");
  ProfScope prof(KC_TRTRI, s, (double)m * k * k, (double)sizeof(T) * 2.0 * m * k + 8.0 * k * k);
  TTG_HIP(hipFuncSetAttribute((const void *)k_trsm<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_trsm<T>), dim3((unsigned)((m + CB - 1) / CB)), dim3(256), lds, s, (int)m, (int)k,
                     A, (long)lda, R, X, (long)ldx);
  TTG_HIP(hipGetLastError());
}
template void trsm_right_upper<double>(hipStream_t, int64_t, int64_t, const double *, int64_t, const double *, double *, int64_t);
template void trsm_right_upper<float>(hipStream_t, int64_t, int64_t, const float *, int64_t, const double *, float *, int64_t);

}  // namespace ttg
