// Cholesky factorisation and triangular solve for Cholesky-QR (gfx950).
//
//   k_chol6 : k <= 256, one workgroup, G + shift = R^T R (upper R), blocked by
//             32, left-looking: the diagonal block is factorised in registers
//             of wave 0 (which also forms R_JJ^{-T}) while waves 1..7 update
//             the next block row from the rows above by FP64 MFMA (R's rows
//             kept transposed in its lower triangle meanwhile); the panel is
//             one MFMA product with R_JJ^{-T}.
//   k_chol4 : k <= 480, right-looking: panel columns solved in registers of
//             waves 1..7 beside the pivot loop, MFMA SYRK trailing update.
//   k_chol3 : any k <= 512, right-looking, panel solved after the pivots.
//   k_trsm2 : X = A R^{-1} (R upper, right side), 32 rows per workgroup;
//             per 32-column block J: B = A_J - X_{<J} R_{<J,J} (MFMA), then
//             a right-looking substitution with R_JJ, one row per lane of
//             wave 0, while waves 1..3 stage block J+1.  k_trsm covers
//             k > 256.  No inverse of R is formed, so the residual
//             ||X R - A|| stays O(u ||X|| ||R||) for any conditioning.
//   k_trsm3 : X = A R^{-1} for FP64 right after k_chol6, which left the
//             inverses W_J = R_JJ^{-1} of the 32x32 diagonal blocks: per
//             block B = A_J - X_{<J} R_{<J,J}, X0 = B W_J, then one
//             refinement X_J = X0 + (B - X0 R_JJ) W_J, all by MFMA, so the
//             serial substitution of k_trsm2 disappears and the residual
//             stays at rounding level.
#include "ttg_internal.h"
#include "ttg_device.h"
#include <cmath>
#include <algorithm>
#include <type_traits>

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
constexpr int TRS = CB + 1;  // odd LDS row stride of staged 32-row blocks

// ---------------------------------------------------------------------------
// Cholesky
// ---------------------------------------------------------------------------
// One workgroup of 1024 threads; R (k x k, ld k) lives in global memory (it
// stays in L2) and receives the upper factor with a zero lower triangle.
// Per 32-column block J0:
//   diag : wave 0 factorises R_JJ with lane j holding column j in registers;
//          row c of the factor is published through LDS (Dl, row-major) and
//          every lane applies the rank-1 update to its column.
//   panel: R_J,rest <- R_JJ^{-T} R_J,rest by forward substitution, one column
//          per thread, staged in LDS (Ps) with coalesced global traffic.
//   syrk : R_rest,rest -= Ps^T Ps on 16x16 upper tiles (FP64 MFMA).
// stats[0] = max|r_jj| / min|r_jj|, stats[1] = max_ij |R_ij - delta_ij|,
// flag = 1-based index of the first non-positive pivot (0 = success).
constexpr int CHOL_THREADS = 512;

template <bool STAMP>
__global__ __launch_bounds__(CHOL_THREADS) void k_chol3(int k, const double *__restrict__ G, long ldg,
                                                        double shift_rel, double *__restrict__ R,
                                                        int *flag, double *stats,
                                                        unsigned long long *dbg) {
  extern __shared__ __align__(16) char smem[];  // 16-byte aligned: FP64 LDS accesses must not straddle
  double *Dl = reinterpret_cast<double *>(smem);  // CB x CB, Dl[t * CB + l] = R(J0 + t, J0 + l)
  double *Ps = Dl + CB * CB;  // CB x wps, Ps[t * wps + c] = R(J0 + t, J0 + CB + c)
  __shared__ double red[CHOL_THREADS / 64];
  __shared__ double dinv[CB];
  __shared__ int sh_fail;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr int NW = CHOL_THREADS / 64;
  double tr = 0.0;
  for (int j = tid; j < k; j += CHOL_THREADS) tr += G[j + (long)j * ldg];
  tr = block_sum(tr, red);
  const double shift = shift_rel * tr / (k > 0 ? k : 1);
  for (int j = wv; j < k; j += NW)
    for (int i = lane; i < k; i += 64)
      R[i + (long)j * k] = (i <= j) ? G[i + (long)j * ldg] + (i == j ? shift : 0.0) : 0.0;
  if (tid == 0) sh_fail = 0;
  double dmax = 0.0, dmin = INFINITY, dev = 0.0;
  for (int J0 = 0; J0 < k; J0 += CB) {
    const int jb = min(CB, k - J0);
    __syncthreads();
    if (STAMP && dbg && tid == 0) dbg[3 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
    // stage the diagonal block column-major in Dl (identity-padded past jb)
    for (int idx = tid; idx < CB * CB; idx += CHOL_THREADS) {
      const int i = idx & (CB - 1), j = idx >> 5;
      Dl[idx] = (i < jb && j < jb) ? (i <= j ? R[(J0 + i) + (long)(J0 + j) * k] : 0.0) : (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    if (wv == 0) {
      const int j = lane & (CB - 1);
      double col[CB];
#pragma unroll
      for (int i = 0; i < CB; ++i) col[i] = Dl[j * CB + i];
      wave_sync_lds();
      unsigned failmask = 0;
      double myinv = 1.0;
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        double p = readlane_d(col[c], c);
        // straight-line code only (branches here push col[] into scratch): a
        // failed pivot is recorded in a bit mask and replaced by a harmless 1
        const bool bad = !(p > 0.0) || !isfinite(p);
        failmask |= bad ? (1u << c) : 0u;
        p = bad ? 1.0 : p;
        // 1/sqrt(p) from the hardware estimate and two Newton steps (full
        // precision for the positive normal pivots that reach this point)
        double inv = __builtin_amdgcn_rsq(p);
        {
          const double h = 0.5 * p;
          double e = __builtin_fma(-h * inv, inv, 0.5);
          inv = __builtin_fma(inv, e, inv);
          e = __builtin_fma(-h * inv, inv, 0.5);
          inv = __builtin_fma(inv, e, inv);
        }
        const double dc = p * inv;
        const double rc = (j == c) ? dc : (j > c ? col[c] * inv : 0.0);
        col[c] = rc;
        myinv = (j == c) ? inv : myinv;
        // lanes 32..63 mirror lanes 0..31 and store identical values (no exec branch
        // inside the unrolled loop: branches there push col[] into scratch)
        Dl[c * CB + j] = rc;
        wave_sync_lds();
#pragma unroll
        for (int i = c + 1; i < CB; ++i) col[i] -= Dl[c * CB + i] * rc;
        __builtin_amdgcn_sched_barrier(0);
      }
      if (lane < CB) dinv[j] = myinv;
      if (lane == 0) sh_fail = failmask ? J0 + __builtin_ctz(failmask) + 1 : 0;
    }
    __syncthreads();
    if (STAMP && dbg && tid == 0) dbg[3 * (J0 / CB) + 1] = __builtin_amdgcn_s_memrealtime();
    if (sh_fail) break;
    // diagonal block to R (+ statistics)
    for (int idx = tid; idx < CB * CB; idx += CHOL_THREADS) {
      const int t = idx & (CB - 1), l = idx >> 5;
      if (t < jb && l < jb && t <= l) {
        const double v = Dl[t * CB + l];
        R[(J0 + t) + (long)(J0 + l) * k] = v;
        if (t == l) {
          dmax = fmax(dmax, fabs(v));
          dmin = fmin(dmin, fabs(v));
          dev = fmax(dev, fabs(v - 1.0));
        } else {
          dev = fmax(dev, fabs(v));
        }
      }
    }
    const int w = k - J0 - jb;
    if (w <= 0) continue;
    const int wpad = (w + 15) & ~15, wps = wpad + 1;
    // stage the panel rows (coalesced: 32 consecutive threads read one column)
    for (int idx = tid; idx < CB * wpad; idx += CHOL_THREADS) {
      const int t = idx & (CB - 1), c = idx >> 5;
      Ps[t * wps + c] = (t < jb && c < w) ? R[(J0 + t) + (long)(J0 + jb + c) * k] : 0.0;
    }
    __syncthreads();
    for (int c = tid; c < wpad; c += CHOL_THREADS) {
      double b[CB];
#pragma unroll
      for (int t = 0; t < CB; ++t) b[t] = Ps[t * wps + c];
#pragma unroll
      for (int t = 0; t < CB; ++t) {
        const double x = b[t] * dinv[t];
        b[t] = x;
#pragma unroll
        for (int l = t + 1; l < CB; ++l) b[l] -= Dl[t * CB + l] * x;
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int t = 0; t < CB; ++t) {
        const double v = (t < jb && c < w) ? b[t] : 0.0;
        Ps[t * wps + c] = v;
        dev = fmax(dev, fabs(v));
      }
    }
    __syncthreads();
    if (STAMP && dbg && tid == 0) dbg[3 * (J0 / CB) + 2] = __builtin_amdgcn_s_memrealtime();
    for (int idx = tid; idx < CB * w; idx += CHOL_THREADS) {
      const int t = idx & (CB - 1), c = idx >> 5;
      if (t < jb) R[(J0 + t) + (long)(J0 + jb + c) * k] = Ps[t * wps + c];
    }
    // trailing SYRK on the upper 16x16 tiles, four tiles per wave at a time:
    // their R entries are loaded first and their MFMA chains interleaved, so
    // the L2 latency of the read-modify-write is paid once per four tiles.
    // The product is formed transposed (tile^T = Ps_j^T Ps_i) so that 16
    // consecutive lanes touch 16 consecutive rows of R.
    const int nt = wpad >> 4;
    const int ntiles = nt * (nt + 1) / 2;
    for (int base = wv; base < ntiles; base += 4 * NW) {
      int ti[4], tj[4];
      bool ok[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int tile = base + q * NW;
        ok[q] = tile < ntiles;
        int c = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
        while (c * (c + 1) / 2 > tile) --c;
        while ((c + 1) * (c + 2) / 2 <= tile) ++c;
        tj[q] = c;
        ti[q] = tile - c * (c + 1) / 2;
      }
      const int i_l = lane & 15, j_l = lane >> 4;
      double old[4][4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = ti[q] * 16 + i_l, jj = tj[q] * 16 + j_l + 4 * v;
          old[q][v] = (ok[q] && i < w && jj < w && i <= jj) ? R[(J0 + jb + i) + (long)(J0 + jb + jj) * k] : 0.0;
        }
      v4d_f acc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = v4d_f{0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < CB; kk += 4) {
        const int t = kk + (lane >> 4);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          acc[q] = MFf<double>::op(Ps[t * wps + tj[q] * 16 + i_l], Ps[t * wps + ti[q] * 16 + i_l], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          // acc[q][v] = tile^T[j_l + 4v][i_l] = tile[i_l][j_l + 4v]
          const int i = ti[q] * 16 + i_l, jj = tj[q] * 16 + j_l + 4 * v;
          if (ok[q] && i < w && jj < w && i <= jj) R[(J0 + jb + i) + (long)(J0 + jb + jj) * k] = old[q][v] - acc[q][v];
        }
    }
  }
  __syncthreads();
  if (STAMP && dbg && tid == 0) dbg[3 * ((k + CB - 1) / CB)] = __builtin_amdgcn_s_memrealtime();
  const int failed = sh_fail;
  dmax = block_max(dmax, red);
  dmin = -block_max(-dmin, red);
  dev = block_max(dev, red);
  if (tid == 0) {
    flag[0] = failed;
    stats[0] = failed ? INFINITY : (dmin > 0.0 ? dmax / dmin : INFINITY);
    stats[1] = failed ? INFINITY : dev;
  }
}

// ---------------------------------------------------------------------------
// k_chol4: k_chol3 with the panel solve fused into the pivot loop.  Per
// 32-column block the diagonal block and the panel rows are staged in LDS
// together; wave 0 factorises the diagonal block one pivot at a time and,
// after each workgroup barrier, waves 1..7 apply that pivot (row c of R and
// 1 / r_cc) to their panel columns held in registers while wave 0 already
// computes the next pivot.  The panel's 32-step substitution chain thus runs
// beside the diagonal factorisation instead of after it.  Same operations in
// the same order per column as k_chol3: identical factors.
// One panel column per thread of waves 1..7: panels up to 448 columns (k <= 480).
// jobs != nullptr: workgroup b factorises the matrix of jobs[b] (batched, no
// shift; the jobs' Wout and fro are not written).
// ---------------------------------------------------------------------------
template <bool STAMP>
__global__ __launch_bounds__(CHOL_THREADS) void k_chol4(int k_, const double *__restrict__ G_, long ldg_,
                                                        double shift_rel, double *__restrict__ R_,
                                                        int *flag_, double *stats_,
                                                        unsigned long long *dbg,
                                                        const CholJob *__restrict__ jobs = nullptr) {
  const CholJob *jb = jobs ? jobs + blockIdx.x : nullptr;
  const int k = jb ? jb->k : k_;
  const double *__restrict__ G = jb ? jb->G : G_;
  const long ldg = jb ? (long)jb->ldg : ldg_;
  double *__restrict__ R = jb ? jb->R : R_;
  int *flag = jb ? jb->flag : flag_;
  double *stats = jb ? jb->stats : stats_;
  extern __shared__ __align__(16) char smem[];  // 16-byte aligned: FP64 LDS accesses must not straddle
  double *Dl = reinterpret_cast<double *>(smem);  // CB x CB, Dl[t * CB + l] = R(J0 + t, J0 + l)
  double *Ps = Dl + CB * CB;  // CB x wps, Ps[t * wps + c] = R(J0 + t, J0 + CB + c)
  __shared__ double red[CHOL_THREADS / 64];
  __shared__ double dinv[CB];
  __shared__ int sh_fail;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr int NW = CHOL_THREADS / 64;
  double tr = 0.0;
  for (int j = tid; j < k; j += CHOL_THREADS) tr += G[j + (long)j * ldg];
  tr = block_sum(tr, red);
  const double shift = shift_rel * tr / (k > 0 ? k : 1);
  for (int j = wv; j < k; j += NW)
    for (int i = lane; i < k; i += 64)
      R[i + (long)j * k] = (i <= j) ? G[i + (long)j * ldg] + (i == j ? shift : 0.0) : 0.0;
  if (tid == 0) sh_fail = 0;
  double dmax = 0.0, dmin = INFINITY, dev = 0.0;
  for (int J0 = 0; J0 < k; J0 += CB) {
    const int jb = min(CB, k - J0);
    const int w = k - J0 - jb;  // > 0 only when jb == CB
    const int wpad = (w + 15) & ~15, wps = wpad + 1;
    __syncthreads();
    if (STAMP && dbg && tid == 0) dbg[3 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
    // diagonal block column-major in Dl (identity-padded past jb); panel rows in Ps
    for (int idx = tid; idx < CB * CB; idx += CHOL_THREADS) {
      const int i = idx & (CB - 1), j = idx >> 5;
      Dl[idx] = (i < jb && j < jb) ? (i <= j ? R[(J0 + i) + (long)(J0 + j) * k] : 0.0) : (i == j ? 1.0 : 0.0);
    }
    for (int idx = tid; idx < CB * wpad; idx += CHOL_THREADS) {
      const int t = idx & (CB - 1), c = idx >> 5;
      Ps[t * wps + c] = (c < w) ? R[(J0 + t) + (long)(J0 + jb + c) * k] : 0.0;
    }
    __syncthreads();
    const int j = lane & (CB - 1);
    const int pt = tid - 64;
    double v[CB];  // wave 0: column j of the diagonal block; waves 1..7: panel column pt
    if (wv == 0) {
#pragma unroll
      for (int i = 0; i < CB; ++i) v[i] = Dl[j * CB + i];
    } else if (w > 0) {
      const int c1 = pt < wpad ? pt : 0;  // threads past the panel repeat column 0 and store nothing
#pragma unroll
      for (int t = 0; t < CB; ++t) v[t] = Ps[t * wps + c1];
    }
    unsigned failmask = 0;
    double rc_prev = 0.0;
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      if (wv == 0) {
        // Pivot c.  Critical path: v[c] -> pivot -> row c -> v[c+1] (the next
        // pivot column), using only this lane's values and one readlane.
        // Row c-1's update of the later columns (LDS reads issued first,
        // consumed after the pivot) runs beside it.
        double rr[CB];
        if (c >= 1) {
#pragma unroll
          for (int i = 0; i < CB; ++i)
            if (i > c) rr[i] = Dl[(c - 1) * CB + i];
        }
        __builtin_amdgcn_sched_barrier(0);
        double p = readlane_d(v[c], c);
        // straight-line code only (branches here push v[] into scratch): a
        // failed pivot is recorded in a bit mask and replaced by a harmless 1
        const bool bad = !(p > 0.0) || !isfinite(p);
        failmask |= bad ? (1u << c) : 0u;
        p = bad ? 1.0 : p;
        // 1/sqrt(p) from the hardware estimate and two Newton steps
        double inv = __builtin_amdgcn_rsq(p);
        {
          const double h = 0.5 * p;
          double e = __builtin_fma(-h * inv, inv, 0.5);
          inv = __builtin_fma(inv, e, inv);
          e = __builtin_fma(-h * inv, inv, 0.5);
          inv = __builtin_fma(inv, e, inv);
        }
        const double dc = p * inv;
        const double rc = (j == c) ? dc : (j > c ? v[c] * inv : 0.0);
        v[c] = rc;
        // every lane holds the same pivot; lanes 32..63 mirror lanes 0..31
        Dl[c * CB + j] = rc;
        dinv[c] = inv;
        if (c + 1 < CB) v[c + 1] -= readlane_d(rc, c + 1) * rc;  // R(c, c+1) R(c, j)
        if (c >= 1) {
#pragma unroll
          for (int i = 0; i < CB; ++i)
            if (i > c) v[i] -= rr[i] * rc_prev;  // R(c-1, i) R(c-1, j)
        }
        rc_prev = rc;
      }
      if ((c & 3) == 3) {
        __syncthreads();  // rows c-3..c of R and their 1 / r_cc visible to the panel waves
        if (wv != 0 && w > 0) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int t = c - 3 + u;
            const double x = v[t] * dinv[t];
            v[t] = x;
#pragma unroll
            for (int l = 0; l < CB; ++l)
              if (l > t) v[l] -= Dl[t * CB + l] * x;
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    if (wv == 0 && lane == 0) sh_fail = failmask ? J0 + __builtin_ctz(failmask) + 1 : 0;
    if (wv != 0 && w > 0) {
      if (pt < wpad) {
#pragma unroll
        for (int t = 0; t < CB; ++t) Ps[t * wps + pt] = v[t];
      }
    }
    __syncthreads();
    if (STAMP && dbg && tid == 0) dbg[3 * (J0 / CB) + 1] = __builtin_amdgcn_s_memrealtime();
    if (sh_fail) break;
    // diagonal block to R (+ statistics)
    for (int idx = tid; idx < CB * CB; idx += CHOL_THREADS) {
      const int t = idx & (CB - 1), l = idx >> 5;
      if (t < jb && l < jb && t <= l) {
        const double val = Dl[t * CB + l];
        R[(J0 + t) + (long)(J0 + l) * k] = val;
        if (t == l) {
          dmax = fmax(dmax, fabs(val));
          dmin = fmin(dmin, fabs(val));
          dev = fmax(dev, fabs(val - 1.0));
        } else {
          dev = fmax(dev, fabs(val));
        }
      }
    }
    if (w <= 0) continue;
    // solved panel to R (coalesced: 32 consecutive threads write one column)
    for (int idx = tid; idx < CB * w; idx += CHOL_THREADS) {
      const int t = idx & (CB - 1), c = idx >> 5;
      const double val = Ps[t * wps + c];
      R[(J0 + t) + (long)(J0 + jb + c) * k] = val;
      dev = fmax(dev, fabs(val));
    }
    if (STAMP && dbg && tid == 0) dbg[3 * (J0 / CB) + 2] = __builtin_amdgcn_s_memrealtime();
    // trailing SYRK on the upper 16x16 tiles (as in k_chol3)
    const int nt = wpad >> 4;
    const int ntiles = nt * (nt + 1) / 2;
    for (int base = wv; base < ntiles; base += 4 * NW) {
      int ti[4], tj[4];
      bool ok[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int tile = base + q * NW;
        ok[q] = tile < ntiles;
        int cc = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
        while (cc * (cc + 1) / 2 > tile) --cc;
        while ((cc + 1) * (cc + 2) / 2 <= tile) ++cc;
        tj[q] = cc;
        ti[q] = tile - cc * (cc + 1) / 2;
      }
      const int i_l = lane & 15, j_l = lane >> 4;
      double old[4][4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int i = ti[q] * 16 + i_l, jj = tj[q] * 16 + j_l + 4 * u;
          old[q][u] = (ok[q] && i < w && jj < w && i <= jj) ? R[(J0 + jb + i) + (long)(J0 + jb + jj) * k] : 0.0;
        }
      v4d_f acc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = v4d_f{0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < CB; kk += 4) {
        const int t = kk + (lane >> 4);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          acc[q] = MFf<double>::op(Ps[t * wps + tj[q] * 16 + i_l], Ps[t * wps + ti[q] * 16 + i_l], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int i = ti[q] * 16 + i_l, jj = tj[q] * 16 + j_l + 4 * u;
          if (ok[q] && i < w && jj < w && i <= jj) R[(J0 + jb + i) + (long)(J0 + jb + jj) * k] = old[q][u] - acc[q][u];
        }
    }
  }
  __syncthreads();
  if (STAMP && dbg && tid == 0) dbg[3 * ((k + CB - 1) / CB)] = __builtin_amdgcn_s_memrealtime();
  const int failed = sh_fail;
  dmax = block_max(dmax, red);
  dmin = -block_max(-dmin, red);
  dev = block_max(dev, red);
  if (tid == 0) {
    flag[0] = failed;
    stats[0] = failed ? INFINITY : (dmin > 0.0 ? dmax / dmin : INFINITY);
    stats[1] = failed ? INFINITY : dev;
  }
}

// ---------------------------------------------------------------------------
// Wave 0's factorisation of the 32 x 32 diagonal block in Dl (column-major,
// the updated block B): lanes 0..31 hold one column each; lanes 32..63 run
// the same eliminations on the columns of the identity, so that they end
// with W = R_JJ^{-T} in Wt (Wt[x * TRS + y] = W(y, x)), which the panel is
// multiplied by instead of a 32-step substitution per panel column.  R_JJ
// overwrites Dl row by row.  Returns the mask of failed pivots.
__device__ __forceinline__ unsigned chol_pivot_chain(double *Dl, double *Wt, int lane) {
  const int j = lane & (CB - 1);
  const bool idl = lane >= CB;
  double *const wrow = idl ? Wt + j * TRS : Dl + j;
  double v[CB];
  {
    // (the identity columns are rebuilt in every block: as loop
    // invariants they would stay live across the whole kernel)
    int jj = j;
    asm volatile("" : "+v"(jj));
#pragma unroll
    for (int i = 0; i < CB; ++i) v[i] = idl ? (i == jj ? 1.0 : 0.0) : Dl[j * CB + i];
  }
  unsigned failmask = 0;
  double rc_prev = 0.0;
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    // Pivot c.  Critical path: v[c] -> pivot -> row c -> v[c+1] (the next
    // pivot column), using only this lane's values and one readlane.
    // Row c-1's update of the later columns (LDS reads issued first,
    // consumed after the pivot) runs beside it.
    double rr[CB];
    if (c >= 1) {
      // the row's address in one VGPR, the entries at immediate offsets
      // from it (as a compile-time constant the compiler would keep one
      // scalar base per pair of entries, spilled and reloaded each step)
      int rb = (c - 1) * CB;
      asm volatile("" : "+v"(rb));
      const double *rowp = Dl + rb;
#pragma unroll
      for (int i = 0; i < CB; ++i)
        if (i > c) rr[i] = rowp[i];
    }
    __builtin_amdgcn_sched_barrier(0);
    double p = readlane_d(v[c], c);
    // straight-line code only (branches here push v[] into scratch): a
    // failed pivot is recorded in a bit mask and replaced by a harmless 1
    const bool bad = !(p > 0.0) || !isfinite(p);
    failmask |= bad ? (1u << c) : 0u;
    p = bad ? 1.0 : p;
    // 1/sqrt(p) from the hardware estimate and two Newton steps
    double inv = __builtin_amdgcn_rsq(p);
    {
      const double h = 0.5 * p;
      double e = __builtin_fma(-h * inv, inv, 0.5);
      inv = __builtin_fma(inv, e, inv);
      e = __builtin_fma(-h * inv, inv, 0.5);
      inv = __builtin_fma(inv, e, inv);
    }
    // the entry by per-lane selects, no branches: with exec-mask branches
    // here the row update below could not be scheduled into the latency
    // of the rsq/Newton chain (it would sit in a later basic block)
    const double dc = p * inv, sc = v[c] * inv;
    const bool keep = idl || j > c;
    double rc = keep ? sc : 0.0;
    rc = (!idl && j == c) ? dc : rc;
    v[c] = rc;
    // row c of R_JJ (lanes 0..31) / entry (c, j) of W (lanes 32..63)
    // (one LDS address per lane: identity lanes write W, the others R_JJ;
    // the stride made opaque in each step: hoisted out of the block loop,
    // the 32 per-step addresses would be spilled to scratch)
    int wst = idl ? 1 : CB;
    asm volatile("" : "+v"(wst));
    *(wrow + c * wst) = rc;
    if (c + 1 < CB) v[c + 1] -= readlane_d(rc, c + 1) * rc;  // R(c, c+1) x_j
    if (c >= 1) {
#pragma unroll
      for (int i = 0; i < CB; ++i)
        if (i > c) v[i] -= rr[i] * rc_prev;  // R(c-1, i) x_j
    }
    rc_prev = rc;
    __builtin_amdgcn_sched_barrier(0);
  }
  return failmask;
}

// ---------------------------------------------------------------------------
// k_chol6: left-looking Cholesky with the block-row update off the critical path.
// Block row J is brought up to date from the factored rows above it, then
// its diagonal block is factorised in registers of wave 0 and its panel
// solved as one MFMA product with R_JJ^{-T} (k_chol4's pivot loop and fused
// panel).  While wave 0 runs the pivot chain of block J, the update waves
// compute the update of block row J+1 from the rows above block J,
// C = R(0:J0, J1:J1+32)^T R(0:J0, J1:k) (J1 = J0 + 32), by FP64 MFMA with
// the operands read straight from memory: R's strict lower triangle holds
// R's rows transposed while the kernel runs (R(t, c) also at (c, t) for
// every panel entry), so a 16-column row segment is 128 contiguous bytes.
// Block J's own rows, the last chunk of that update, are added from the
// panel still in LDS once it is solved.  The strict lower triangle is
// zeroed at the end.  The same waves prefetch block row J+1's entries of G.
// Wave 4, which shares a SIMD with wave 0, takes no part in the update: its
// FP64 MFMAs would compete with the pivot chain's FP64 arithmetic.  Tiles
// of the update and of the B-write: update wave u = 0..5 (waves 1-3, 5-7),
// tile = u + 6q, row tile u & 1, column tile u / 2 + 3q (at most 28 tiles
// for k <= 256; MAXU = 5 per wave).
// ---------------------------------------------------------------------------
template <int MAXT, int MAXU, int US, bool STAMP>
__global__ __launch_bounds__(CHOL_THREADS) void k_chol6(int k_, const double *__restrict__ G_, long ldg_,
                                                        double shift_rel, double *__restrict__ R_,
                                                        int *flag_, double *stats_,
                                                        unsigned long long *dbg, double *Wout_, double *fro_,
                                                        const CholJob *__restrict__ jobs) {
  // jobs != nullptr: workgroup b factorises the matrix of jobs[b] (batched)
  const CholJob *jb = jobs ? jobs + blockIdx.x : nullptr;
  const int k = jb ? jb->k : k_;
  const double *__restrict__ G = jb ? jb->G : G_;
  const long ldg = jb ? (long)jb->ldg : ldg_;
  double *__restrict__ R = jb ? jb->R : R_;
  int *flag = jb ? jb->flag : flag_;
  double *stats = jb ? jb->stats : stats_;
  double *Wout = jb ? jb->Wout : Wout_;
  double *fro = jb ? jb->fro : fro_;
  extern __shared__ __align__(16) char smem[];  // 16-byte aligned: FP64 LDS accesses must not straddle
  double *Dl = reinterpret_cast<double *>(smem);  // CB x CB, Dl[t * CB + l] = R(J0 + t, J0 + l)
  double *Ps = Dl + CB * CB;  // CB x wps, Ps[t * wps + c] = R(J0 + t, J0 + CB + c)
  __shared__ double red[CHOL_THREADS / 64];
  __shared__ double Wt[CB * TRS];  // Wt[x * TRS + y] = W(y, x), W = R_JJ^{-T}
  __shared__ int sh_fail;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform per wave: scalar branches
  constexpr int NW = CHOL_THREADS / 64, NU = NW - 2;       // NU update waves (1-3, 5-7)
  constexpr int wps = US;  // panel row stride (odd, > the widest panel)
  const bool upd = wv != 0 && wv != 4;
  const int u0 = wv - 1 - (wv > 4 ? 1 : 0);
  double tr = 0.0;
  for (int j = tid; j < k; j += CHOL_THREADS) tr += G[j + (long)j * ldg];
  tr = block_sum(tr, red);
  const double shift = shift_rel * tr / (k > 0 ? k : 1);
  if (tid == 0) sh_fail = 0;
  double dmax = 0.0, dmin = INFINITY, dev = 0.0, sq = 0.0;  // sq: ||R||_F^2 (this thread's share)
  // geometry of block row Jx: its rows jbx, the panel width wx, the tiles
  auto geom = [&](int Jx, int &jbx, int &wx, int &wpadx, int &ntx) {
    jbx = min(CB, k - Jx);
    wx = k - Jx - jbx;  // > 0 only when jbx == CB
    wpadx = (wx + 15) & ~15;
    ntx = 2 * (2 + (wpadx >> 4));
  };
  // block row Jx's entries of G (upper triangle) for this wave's tiles
  double gv[MAXU][4];
  auto load_g = [&](int Jx, int ln) {
    int jbx, wx, wpadx, ntx;
    geom(Jx, jbx, wx, wpadx, ntx);
#pragma unroll
    for (int q = 0; q < MAXU; ++q) {
      const int tile = u0 + NU * q, ti = tile & 1, tj = tile >> 1;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = ti * 16 + (ln & 15), c = tj * 16 + (ln >> 4) + 4 * u;
        const bool need = tile < ntx && r < jbx && (c < CB ? (c < jbx && r <= c) : c - CB < wx);
        gv[q][u] = need ? G[(Jx + r) + (long)(Jx + c) * ldg] : 0.0;
      }
    }
  };
  v4d_f acc[MAXU];  // C of the next block row to be factorised (this wave's tiles)
  {
    // block row 0 has no update: B = G rows + shift, copied by all threads
    // (k <= 256: at most 32 x 256 entries, 16 per thread; all loads issue
    // before the first store, from a clamped and so valid address)
    int jb, w, wpad, nt;
    geom(0, jb, w, wpad, nt);
    constexpr int NCP = CB * 256 / CHOL_THREADS;
    double cv[NCP];
#pragma unroll
    for (int e = 0; e < NCP; ++e) {
      const int idx = tid + CHOL_THREADS * e, r = idx & (CB - 1), c = idx >> 5;
      cv[e] = G[min(r, k - 1) + (long)min(c, k - 1) * ldg];
    }
#pragma unroll
    for (int e = 0; e < NCP; ++e) {
      const int idx = tid + CHOL_THREADS * e, r = idx & (CB - 1), c = idx >> 5;
      if (c < CB)
        Dl[c * CB + r] = (r < jb && c < jb) ? (r <= c ? cv[e] + (r == c ? shift : 0.0) : 0.0) : (r == c ? 1.0 : 0.0);
      else if (c - CB < wpad)
        Ps[r * wps + (c - CB)] = (c - CB < w) ? cv[e] : 0.0;
    }
  }
  for (int J0 = 0; J0 < k; J0 += CB) {
    int jb, w, wpad, ntiles;
    geom(J0, jb, w, wpad, ntiles);
    __syncthreads();  // the previous block's reads of Dl, Ps are done
    if (STAMP && dbg && tid == 0) dbg[3 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
    // per-lane index math of this block from an opaque copy of the lane
    // index: left to LICM, the addresses of every tile would be hoisted out
    // of the block loop and spilled
    int ln = lane;
    asm volatile("" : "+v"(ln));
    if (J0 > 0 && upd) {
      // B = G rows + shift - C into Dl (diagonal block, column-major,
      // identity past jb) and Ps (panel rows, zero past w);
      // acc[q][v] = C(ti*16 + lane%16, tj*16 + lane/16 + 4v) (ntiles <= 28)
      double sh = shift;  // opaque: its selects would be hoisted and spilled
      asm volatile("" : "+v"(sh));
#pragma unroll
      for (int q = 0; q < MAXU; ++q) {
        const int tile = u0 + NU * q, ti = tile & 1, tj = tile >> 1;
        if (tile < ntiles) {  // uniform
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int r = ti * 16 + (ln & 15), c = tj * 16 + (ln >> 4) + 4 * u;
            if (c < CB) {
              Dl[c * CB + r] = (r < jb && c < jb) ? (r <= c ? gv[q][u] + (r == c ? sh : 0.0) - acc[q][u] : 0.0)
                                                  : (r == c ? 1.0 : 0.0);
            } else if (c - CB < wpad) {
              Ps[r * wps + (c - CB)] = (c - CB < w) ? gv[q][u] - acc[q][u] : 0.0;
            }
          }
        }
      }
    }
    __syncthreads();
    const int J1 = J0 + CB;
    int jb1, w1, wpad1, nt1;
    geom(J1, jb1, w1, wpad1, nt1);
    if (wv == 0) {
      const unsigned failmask = chol_pivot_chain(Dl, Wt, lane);
      if (lane == 0) sh_fail = failmask ? J0 + __builtin_ctz(failmask) + 1 : 0;
#pragma unroll
      for (int q = 0; q < MAXU; ++q) acc[q] = v4d_f{0, 0, 0, 0};
    } else if (upd) {
#pragma unroll
      for (int q = 0; q < MAXU; ++q) acc[q] = v4d_f{0, 0, 0, 0};
      if (J1 < k) {
        // update of block row J1 from rows 0..J0-1 (final since earlier
        // blocks), from the transposed rows in R's lower triangle; columns
        // past k (tile padding) read column k-1 instead, a finite value the
        // B-write discards
        if (J0 > 0) {
#pragma unroll
          for (int q = 0; q < MAXU; ++q) {
            const int tile = u0 + NU * q, ti = tile & 1, tj = tile >> 1;
            if (tile < nt1) {  // uniform per wave: one MFMA chain per tile
              const int ca = min(J1 + ti * 16 + (ln & 15), k - 1), cb = min(J1 + tj * 16 + (ln & 15), k - 1);
              const double *pa = R + ca + (long)(ln >> 4) * k;
              const double *pb = R + cb + (long)(ln >> 4) * k;
              // operands of one 32-row chunk (8 MFMA steps) in registers, two
              // buffers: the next chunk's loads are in flight while this
              // chunk's MFMAs issue (the chunk offset opaque, so that the
              // addresses are formed at each load instead of hoisted)
              double xa[8], xb[8], ya[8], yb[8];
              auto ld = [&](double *oa, double *ob, int kc) {
                long off = (long)kc * k;
                asm volatile("" : "+v"(off));
#pragma unroll
                for (int s8 = 0; s8 < 8; ++s8) {
                  oa[s8] = pa[off + (long)(4 * s8) * k];
                  ob[s8] = pb[off + (long)(4 * s8) * k];
                }
              };
              auto mm = [&](const double *oa, const double *ob) {
#pragma unroll
                for (int s8 = 0; s8 < 8; ++s8) acc[q] = MFf<double>::op(ob[s8], oa[s8], acc[q]);
              };
              ld(xa, xb, 0);
              for (int kc = 0; kc < J0; kc += 2 * CB) {
                if (kc + CB < J0) ld(ya, yb, kc + CB);
                mm(xa, xb);
                if (kc + 2 * CB < J0) ld(xa, xb, kc + 2 * CB);
                if (kc + CB < J0) mm(ya, yb);
              }
            }
          }
        }
        load_g(J1, ln);
      }
    }
    __syncthreads();  // R_JJ in Dl, W in Wt
    if (Wout)  // W_J for the triangular solves with this factor
      for (int idx = tid; idx < CB * TRS; idx += CHOL_THREADS) Wout[(long)(J0 / CB) * (CB * TRS) + idx] = Wt[idx];
    if (!sh_fail && w > 0) {
      // panel rows of R: P = W B (B: the updated panel rows in Ps), FP64 MFMA
      // on 16x16 tiles: tile = wv + 8q, ti = wv & 1, tj = wv / 2 + 4q
      const int ntp = 2 * (wpad >> 4);
      const int ti = wv & 1, tj0 = wv >> 1;
      int lp = lane;  // opaque: the tile bases are formed here, not hoisted and spilled
      asm volatile("" : "+v"(lp));
      v4d_f pacc[MAXT];
#pragma unroll
      for (int q = 0; q < MAXT; ++q) {
        pacc[q] = v4d_f{0, 0, 0, 0};
        const int tj = tj0 + 4 * q;
        if (wv + NW * q < ntp) {  // uniform per wave
#pragma unroll
          for (int kk = 0; kk < CB; kk += 4) {
            const int t = kk + (lp >> 4);
            pacc[q] = MFf<double>::op(Ps[t * wps + tj * 16 + (lp & 15)], Wt[t * TRS + ti * 16 + (lp & 15)], pacc[q]);
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < MAXT; ++q) {
        const int tj = tj0 + 4 * q;
        if (wv + NW * q < ntp) {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            Ps[(ti * 16 + (lp & 15)) * wps + tj * 16 + (lp >> 4) + 4 * u] = pacc[q][u];
        }
      }
    }
    __syncthreads();
    if (STAMP && dbg && tid == 0) dbg[3 * (J0 / CB) + 1] = __builtin_amdgcn_s_memrealtime();
    if (sh_fail) break;
    // block J's rows: the last chunk of block row J1's update, from the
    // panel in LDS (Ps columns 0.. are block row J1's columns); before the
    // stores, which the other waves issue meanwhile and which nothing waits
    // for until the next barrier
    if (w > 0 && upd) {
#pragma unroll
      for (int q = 0; q < MAXU; ++q) {
        const int tile = u0 + NU * q, ti = tile & 1, tj = tile >> 1;
        if (tile < nt1) {
#pragma unroll
          for (int kk = 0; kk < CB; kk += 4) {
            const int t = kk + (ln >> 4);
            acc[q] = MFf<double>::op(Ps[t * wps + tj * 16 + (ln & 15)], Ps[t * wps + ti * 16 + (ln & 15)], acc[q]);
          }
        }
      }
    }
    // diagonal block to R (+ statistics)
    for (int idx = tid; idx < CB * CB; idx += CHOL_THREADS) {
      const int t = idx & (CB - 1), l = idx >> 5;
      if (t < jb && l < jb && t <= l) {
        const double val = Dl[t * CB + l];
        R[(J0 + t) + (long)(J0 + l) * k] = val;
        sq += val * val;
        if (t == l) {
          dmax = fmax(dmax, fabs(val));
          dmin = fmin(dmin, fabs(val));
          dev = fmax(dev, fabs(val - 1.0));
        } else {
          dev = fmax(dev, fabs(val));
        }
      }
    }
    if (w > 0) {
      // solved panel to R (coalesced: 32 consecutive threads write one column)
      for (int idx = tid; idx < CB * w; idx += CHOL_THREADS) {
        const int t = idx & (CB - 1), c = idx >> 5;
        const double val = Ps[t * wps + c];
        R[(J0 + t) + (long)(J0 + jb + c) * k] = val;
        sq += val * val;
        dev = fmax(dev, fabs(val));
      }
      // and transposed into the strict lower triangle (row J0 + t of R as
      // part of column J0 + t), which the later updates read
      for (int t = wv; t < CB; t += NW)
        for (int c = lane; c < w; c += 64) R[(J0 + jb + c) + (long)(J0 + t) * k] = Ps[t * wps + c];
    }
    if (STAMP && dbg && tid == 0) dbg[3 * (J0 / CB) + 2] = __builtin_amdgcn_s_memrealtime();
  }
  __syncthreads();
  if (STAMP && dbg && tid == 0) dbg[3 * ((k + CB - 1) / CB)] = __builtin_amdgcn_s_memrealtime();
  // R's strict lower triangle is zero on return
  for (int j = wv; j < k; j += NW)
    for (int i = j + 1 + lane; i < k; i += 64) R[i + (long)j * k] = 0.0;
  const int failed = sh_fail;
  dmax = block_max(dmax, red);
  dmin = -block_max(-dmin, red);
  dev = block_max(dev, red);
  if (fro) sq = block_sum(sq, red);
  if (tid == 0) {
    flag[0] = failed;
    stats[0] = failed ? INFINITY : (dmin > 0.0 ? dmax / dmin : INFINITY);
    stats[1] = failed ? INFINITY : dev;
    if (fro) fro[0] = sqrt(sq);
  }
}

constexpr int CHOL_MAXK = 512;

bool chol_factor(hipStream_t s, int64_t k, const double *G, int64_t ldg, double shift_rel,
                 double *R, int *flag, double *stats, unsigned long long *dbg, double *Wout, double *fro) {
  bool wrote_w = false;
  TTG_CHECK(k >= 1 && k <= CHOL_MAXK, TTG_BAD_ARG, "Cholesky kernel supports 1 <= k <= 512");
  const int64_t wmax = std::max<int64_t>(k - CB, 0);
  const int64_t wps = ((wmax + 15) & ~(int64_t)15) + 1;
  const size_t lds = ((size_t)CB * CB + (size_t)CB * wps) * sizeof(double);
  ProfScope prof(KC_CHOL, s, (double)k * k * k / 3.0, 8.0 * 2.0 * (double)k * k);
  if (((wmax + 15) & ~(int64_t)15) > CHOL_THREADS - 64) {  // panels beyond 448 columns
    auto fn = dbg ? k_chol3<true> : k_chol3<false>;
    TTG_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(fn, dim3(1), dim3(CHOL_THREADS), lds, s, (int)k, G, (long)ldg, shift_rel, R, flag, stats, dbg);
  } else if (k > 256) {
    auto fn = dbg ? k_chol4<true> : k_chol4<false>;
    TTG_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(fn, dim3(1), dim3(CHOL_THREADS), lds, s, (int)k, G, (long)ldg, shift_rel, R, flag, stats, dbg,
                       (const CholJob *)nullptr);
  } else {
    // k <= 256: at most 2 * 16 output tiles (4 per wave), 32 x 224 staged
    // elements (14 per thread), staging and panel row stride 225; Dl + Ps
    // at that stride (8224 doubles) hold the staging buffer and the reads
    // of the tiles past a block (up to index 7230)
    constexpr int MAXT = 4, MAXU = 5, US = 225;
    const size_t lds6 = std::max(lds, (size_t)(CB * CB + CB * US) * sizeof(double));
    auto fn = dbg ? k_chol6<MAXT, MAXU, US, true> : k_chol6<MAXT, MAXU, US, false>;
    TTG_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds6));
    hipLaunchKernelGGL(fn, dim3(1), dim3(CHOL_THREADS), lds6, s, (int)k, G, (long)ldg,
                       shift_rel, R, flag, stats, dbg, Wout, fro, (const CholJob *)nullptr);
    wrote_w = true;
  }
  TTG_HIP(hipGetLastError());
  return wrote_w;
}

// Batched k_chol6: one workgroup per job, every job with k <= 256 (no shift).
void chol_batched(hipStream_t s, const CholJob *jobs_dev, int njobs, int64_t kmax) {
  if (njobs <= 0) return;
  TTG_CHECK(kmax >= 1 && kmax <= 256, TTG_BAD_ARG, "batched Cholesky supports k <= 256");
  constexpr int MAXT = 4, MAXU = 5, US = 225;
  const int64_t wmax = std::max<int64_t>(kmax - CB, 0);
  const int64_t wps = ((wmax + 15) & ~(int64_t)15) + 1;
  const size_t lds = ((size_t)CB * CB + (size_t)CB * wps) * sizeof(double);
  const size_t lds6 = std::max(lds, (size_t)(CB * CB + CB * US) * sizeof(double));
  ProfScope prof(KC_CHOL, s, (double)njobs * kmax * kmax * kmax / 3.0, 8.0 * 2.0 * njobs * kmax * kmax);
  TTG_HIP(hipFuncSetAttribute((const void *)k_chol6<MAXT, MAXU, US, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds6));
  hipLaunchKernelGGL((k_chol6<MAXT, MAXU, US, false>), dim3((unsigned)njobs), dim3(CHOL_THREADS), lds6, s, 0,
                     (const double *)nullptr, 0L, 0.0, (double *)nullptr, (int *)nullptr, (double *)nullptr,
                     (unsigned long long *)nullptr, (double *)nullptr, (double *)nullptr, jobs_dev);
  TTG_HIP(hipGetLastError());
}

// Batched Cholesky for 256 < k <= 480 (k_chol4, one workgroup per job).
void chol4_batched(hipStream_t s, const CholJob *jobs_dev, int njobs, int64_t kmax) {
  if (njobs <= 0) return;
  TTG_CHECK(kmax >= 1 && kmax <= 480, TTG_BAD_ARG, "batched Cholesky (k_chol4) supports k <= 480");
  const int64_t wmax = std::max<int64_t>(kmax - CB, 0);
  const int64_t wps = ((wmax + 15) & ~(int64_t)15) + 1;
  const size_t lds = ((size_t)CB * CB + (size_t)CB * wps) * sizeof(double);
  ProfScope prof(KC_CHOL, s, (double)njobs * kmax * kmax * kmax / 3.0, 8.0 * 2.0 * njobs * kmax * kmax);
  TTG_HIP(hipFuncSetAttribute((const void *)k_chol4<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_chol4<false>, dim3((unsigned)njobs), dim3(CHOL_THREADS), lds, s, 0, (const double *)nullptr,
                     0L, 0.0, (double *)nullptr, (int *)nullptr, (double *)nullptr, (unsigned long long *)nullptr,
                     jobs_dev);
  TTG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// TRSM: X (m x k) = A R^{-1}
// ---------------------------------------------------------------------------
template <class T, bool STAMP>
__global__ __launch_bounds__(256) void k_trsm(int m_, int k_, const T *__restrict__ A_, long lda_,
                                              const double *__restrict__ R_, T *__restrict__ X_,
                                              long ldx_, unsigned long long *dbg,
                                              const TrsmJob *__restrict__ jobs = nullptr) {
  // jobs != nullptr (T = double): blockIdx.y selects the job (batched); row
  // blocks past a job's m leave at once (the whole workgroup)
  const TrsmJob *jb = jobs ? jobs + blockIdx.y : nullptr;
  const int m = jb ? jb->m : m_;
  const int k = jb ? jb->k : k_;
  const T *__restrict__ A = jb ? reinterpret_cast<const T *>(jb->A) : A_;
  const long lda = jb ? (long)jb->lda : lda_;
  const double *__restrict__ R = jb ? jb->R : R_;
  T *__restrict__ X = jb ? reinterpret_cast<T *>(jb->X) : X_;
  const long ldx = jb ? (long)jb->ldx : ldx_;
  if (jb && (int)blockIdx.x * CB >= m) return;
  extern __shared__ __align__(16) char smem[];  // 16-byte aligned: FP64 LDS accesses must not straddle
  typedef typename MFf<T>::V V;
  const int kp = (k + CB - 1) / CB * CB;
  // k <= 256: the whole column panel R(0:J0, J0:J0+32) is staged at once
  // (one batch of loads per block column); larger k: 32 x 32 chunks
  const bool panel = kp <= 256;
  T *Xs = reinterpret_cast<T *>(smem);   // 32 x kp, column-major (Xs[col*32 + row])
  T *Rs = Xs + (size_t)CB * kp;           // R rows t of the panel, row-major (Rs[t*32 + n])
  double *Ds = reinterpret_cast<double *>(Rs + (panel ? kp : CB) * CB);  // R_JJ, row-major (Ds[l*32 + j])
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tr = wv & 1, tc = wv >> 1;
  const int r0 = blockIdx.x * CB;
  const int srow = tid >> 3, sq = tid & 7;  // substitution: 8 lanes per row
  const bool rec = STAMP && dbg && blockIdx.x == 0 && tid == 0;
  if (rec) dbg[0] = __builtin_amdgcn_s_memrealtime();
  for (int idx = tid; idx < CB * kp; idx += blockDim.x) {
    const int row = idx % CB, col = idx / CB;
    const int gr = r0 + row;
    Xs[idx] = (gr < m && col < k) ? A[gr + (long)col * lda] : T(0);
  }
  if (rec) dbg[1] = __builtin_amdgcn_s_memrealtime();
  for (int J0 = 0; J0 < k; J0 += CB) {
    const int jb = min(CB, k - J0);
    if (rec) dbg[2 + 3 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
    V acc = {0, 0, 0, 0};
    if (panel) {
      if (J0 > 0) {
        // R(t, J0 + n) for t < J0, n < 32: contiguous in t; all loads in flight at once
        constexpr int MAXL = 256 * CB / 256;
        T v[MAXL];
        const int tot = J0 * CB;
#pragma unroll
        for (int e = 0; e < MAXL; ++e) {
          const int idx = tid + e * 256;
          const int t = idx % J0, n = idx / J0;
          v[e] = (idx < tot && n < jb) ? (T)R[t + (long)(J0 + n) * k] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < MAXL; ++e) {
          const int idx = tid + e * 256;
          if (idx < tot) Rs[(idx % J0) * CB + idx / J0] = v[e];
        }
        __syncthreads();
        if (rec) dbg[3 + 3 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
        for (int kc = 0; kc < J0; kc += CB) {
#pragma unroll
          for (int kk = 0; kk < CB; kk += 4) {
            const int t = kk + (lane >> 4);
            const T a = Xs[(kc + t) * CB + tr * 16 + (lane & 15)];
            const T b = Rs[(kc + t) * CB + tc * 16 + (lane & 15)];
            acc = MFf<T>::op(a, b, acc);
          }
        }
      }
    } else {
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
    if (rec) dbg[4 + 3 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
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
  if (rec) dbg[2 + 3 * (kp / CB)] = __builtin_amdgcn_s_memrealtime();
  for (int idx = tid; idx < CB * k; idx += blockDim.x) {
    const int row = idx % CB, col = idx / CB;
    const int gr = r0 + row;
    if (gr < m) X[gr + (long)col * ldx] = Xs[col * CB + row];
  }
  if (rec) {
    __builtin_amdgcn_s_waitcnt(0);
    dbg[3 + 3 * (kp / CB)] = __builtin_amdgcn_s_memrealtime();
  }
}

// ---------------------------------------------------------------------------
// k_trsm3: X = A R^{-1} (FP64, k <= 256) when the Cholesky kernel has left the
// inverses W_J = R_JJ^{-T} of the diagonal blocks (Winv, TRS-strided, W_J(c, s)
// at [J][s * TRS + c]).  32 rows per workgroup, 4 waves, one 16 x 16 tile of
// each 32 x 32 block per wave.  Per block J, all on FP64 MFMA:
//   B   = A_J - X_{<J} R_{<J,J}
//   X0  = B W^T                      (= B R_JJ^{-1} to the accuracy of W)
//   X_J = X0 + (B - X0 R_JJ) W^T     (one refinement step: the error of the
//                                     explicit block inverse, O(kappa(R_JJ) u),
//                                     is squared away, so the residual
//                                     ||X R - A|| stays at rounding level)
// instead of a 32-step substitution chain.  The next block's A_J, R_JJ, W_J
// and panel R(0:J0, J) are loaded into registers while the current block
// multiplies.  Odd LDS row stride (33) throughout.  A == nullptr: A = I.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr size_t trsm3_lds(int kp) {
  return ((size_t)TRS * kp + (size_t)(kp > CB ? kp - CB : 0) * TRS + 4 * (size_t)CB * TRS) * sizeof(double);
}

__global__ __launch_bounds__(256) void k_trsm3(int m_, int k_, const double *A_, long lda_,
                                               const double *__restrict__ R_, const double *__restrict__ Winv_,
                                               double *__restrict__ X_, long ldx_,
                                               const TrsmJob *__restrict__ jobs) {
  extern __shared__ __align__(16) char smem[];
  // jobs != nullptr: blockIdx.y selects the job (batched); row blocks past a
  // job's m exit at once (the whole workgroup)
  const TrsmJob *jb = jobs ? jobs + blockIdx.y : nullptr;
  const int m = jb ? jb->m : m_;
  const int k = jb ? jb->k : k_;
  const double *A = jb ? jb->A : A_;
  const long lda = jb ? (long)jb->lda : lda_;
  const double *__restrict__ R = jb ? jb->R : R_;
  const double *__restrict__ Winv = jb ? jb->Winv : Winv_;
  double *__restrict__ X = jb ? jb->X : X_;
  const long ldx = jb ? (long)jb->ldx : ldx_;
  if ((int)blockIdx.x * CB >= m) return;
  typedef v4d_f V;
  const int kp = (k + CB - 1) / CB * CB;
  double *Xs = reinterpret_cast<double *>(smem);  // Xs[col * TRS + row]: A_J, then B_J, then X_J
  double *Rs = Xs + (size_t)TRS * kp;              // Rs[t * TRS + n] = R(t, J0 + n), t < J0
  double *Ds = Rs + (size_t)(kp > CB ? kp - CB : 0) * TRS;  // Ds[a * TRS + b] = R(J0 + a, J0 + b)
  double *Wb = Ds + CB * TRS;                      // Wb[s * TRS + c] = W_J(c, s)
  double *X0s = Wb + CB * TRS;                     // X0s[c * TRS + r] = X0(r, c)
  double *Rb = X0s + CB * TRS;                     // Rb[c * TRS + r] = (B - X0 R_JJ)(r, c)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tr = wv & 1, tc = wv >> 1;
  const int il = lane & 15;
  const int r0 = blockIdx.x * CB;
  constexpr int MAXP = 4 * (256 / CB - 1);  // panel elements per thread (J0 <= 224)
  double pa[4], pd[4], pw[5], pr[MAXP];
  auto fetch = [&](int J0) {
    const int jb = min(CB, k - J0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int lo = idx & (CB - 1), hi = idx >> 5;
      pa[e] = (r0 + lo < m && hi < jb) ? (A ? A[(r0 + lo) + (long)(J0 + hi) * lda] : (r0 + lo == J0 + hi ? 1.0 : 0.0))
                                       : 0.0;
      pd[e] = (lo < jb && hi < jb) ? R[(J0 + lo) + (long)(J0 + hi) * k] : (lo == hi ? 1.0 : 0.0);
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const int idx = tid + 256 * e;
      pw[e] = idx < CB * TRS ? Winv[(long)(J0 / CB) * (CB * TRS) + idx] : 0.0;
    }
    const int tot = J0 * CB;
#pragma unroll
    for (int e = 0; e < MAXP; ++e) {
      const int idx = tid + 256 * e;
      const int t = ((idx >> 10) << 5) + (idx & (CB - 1)), n = (idx >> 5) & (CB - 1);
      pr[e] = (idx < tot && n < jb) ? R[t + (long)(J0 + n) * k] : 0.0;
    }
  };
  // A = I (an inverse): row block r0 of X = R^{-1} is zero left of column
  // block r0 (X_J = (0 - X_<J R_<J,J) W_J = 0 for J < r0 by induction), so
  // those column blocks are written as zeros and skipped
  const int jstart = A ? 0 : min(r0, kp);
  for (int idx = tid; idx < CB * jstart; idx += 256) {
    const int row = idx & (CB - 1), col = idx >> 5;
    if (r0 + row < m && col < k) X[(r0 + row) + (long)col * ldx] = 0.0;
  }
  fetch(jstart);
  for (int J0 = jstart; J0 < k; J0 += CB) {
    const int jb = min(CB, k - J0);
    __syncthreads();  // the previous block is done with Ds, Wb, Rs
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int idx = tid + 256 * e;
      const int lo = idx & (CB - 1), hi = idx >> 5;
      Xs[(J0 + hi) * TRS + lo] = pa[e];
      Ds[lo * TRS + hi] = pd[e];
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const int idx = tid + 256 * e;
      if (idx < CB * TRS) Wb[idx] = pw[e];
    }
    {
      const int tot = J0 * CB;
#pragma unroll
      for (int e = 0; e < MAXP; ++e) {
        const int idx = tid + 256 * e;
        const int t = ((idx >> 10) << 5) + (idx & (CB - 1)), n = (idx >> 5) & (CB - 1);
        if (idx < tot) Rs[t * TRS + n] = pr[e];
      }
    }
    __syncthreads();
    if (J0 + CB < k) fetch(J0 + CB);
    // acc[v] = C(tr*16 + lane%16, tc*16 + lane/16 + 4v); Bv = this lane's entries of B
    V acc = {0, 0, 0, 0};
    if (J0 > jstart) {
      for (int kc = jstart; kc < J0; kc += CB) {  // X's columns left of jstart are zero
#pragma unroll
        for (int kk = 0; kk < CB; kk += 4) {
          const int t = kc + kk + (lane >> 4);
          acc = MFf<double>::op(Rs[t * TRS + tc * 16 + il], Xs[t * TRS + tr * 16 + il], acc);
        }
      }
    }
    double Bv[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = tr * 16 + il, col = tc * 16 + (lane >> 4) + 4 * v;
      Bv[v] = Xs[(J0 + col) * TRS + row] - acc[v];
    }
    // (each lane rewrites only the entries it read; the update read columns < J0)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = tr * 16 + il, col = tc * 16 + (lane >> 4) + 4 * v;
      Xs[(J0 + col) * TRS + row] = Bv[v];
    }
    __syncthreads();
    // X0 = B W^T
    V x0 = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < CB; kk += 4) {
      const int s2 = kk + (lane >> 4);
      x0 = MFf<double>::op(Wb[s2 * TRS + tc * 16 + il], Xs[(J0 + s2) * TRS + tr * 16 + il], x0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) X0s[(tc * 16 + (lane >> 4) + 4 * v) * TRS + tr * 16 + il] = x0[v];
    __syncthreads();
    // residual B - X0 R_JJ
    V pr2 = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < CB; kk += 4) {
      const int s2 = kk + (lane >> 4);
      pr2 = MFf<double>::op(Ds[s2 * TRS + tc * 16 + il], X0s[s2 * TRS + tr * 16 + il], pr2);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) Rb[(tc * 16 + (lane >> 4) + 4 * v) * TRS + tr * 16 + il] = Bv[v] - pr2[v];
    __syncthreads();
    // X_J = X0 + (B - X0 R_JJ) W^T
    V dx = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < CB; kk += 4) {
      const int s2 = kk + (lane >> 4);
      dx = MFf<double>::op(Wb[s2 * TRS + tc * 16 + il], Rb[s2 * TRS + tr * 16 + il], dx);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = tr * 16 + il, col = tc * 16 + (lane >> 4) + 4 * v;
      const double xv = x0[v] + dx[v];
      Xs[(J0 + col) * TRS + row] = xv;
      if (r0 + row < m && col < jb) X[(r0 + row) + (long)(J0 + col) * ldx] = xv;
    }
  }
}

// ---------------------------------------------------------------------------
// k_trsm2: X = A R^{-1} for k <= 256, 32 rows per workgroup, 4 waves.
// Per 32-column block J:
//   B_J = A_J - X_{<J} R_{<J,J}    (all waves: FP64 / FP32 MFMA, 2 x 2 tiles)
//   X_J R_JJ = B_J                 (wave 0: one row per lane, right-looking
//                                   substitution in FP64 registers, rows of
//                                   R_JJ read from LDS, no cross-lane traffic;
//                                   X_J is stored to X at once)
// while waves 1..3 stage block J+1 (A_{J+1} into Xs, R_{J+1,J+1} into the
// other Ds buffer, the panel R(0:J0+32, J+1) into Rs, which the MFMA of
// block J no longer reads).  Staged blocks use an odd LDS row stride (33):
// column and row accesses are conflict-free.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr size_t trsm2_lds(int kp, size_t tsz) {
  return (size_t)TRS * kp * tsz + (size_t)(kp > CB ? kp - CB : 0) * TRS * tsz + 2 * (size_t)CB * TRS * 8;
}

template <class T, bool STAMP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_trsm2(int m, int k, const T *__restrict__ A, long lda,
                                               const double *__restrict__ R, T *__restrict__ X,
                                               long ldx, unsigned long long *dbg) {
  extern __shared__ __align__(16) char smem[];
  typedef typename MFf<T>::V V;
  const int kp = (k + CB - 1) / CB * CB;
  T *Xs = reinterpret_cast<T *>(smem);  // Xs[col * TRS + row]: A_J, then B_J, then X_J
  T *Rs = Xs + (size_t)TRS * kp;         // Rs[t * TRS + n] = R(t, J0 + n), t < J0
  // Ds[buf][a * TRS + b] = R(J0 + a, J0 + b) (identity past jb), buf = block parity
  double *Ds0 = reinterpret_cast<double *>(smem + trsm2_lds(kp, sizeof(T)) - 2 * (size_t)CB * TRS * 8);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int tr = wv & 1, tc = wv >> 1;
  const int r0 = blockIdx.x * CB;
  const bool rec = STAMP && dbg && blockIdx.x == 0 && tid == 0;
  if (rec) dbg[0] = __builtin_amdgcn_s_memrealtime();
  // stage A_J and R_JJ of block J0 with NTHR threads from thread FIRST on
  // (every load in flight before the first LDS store)
  auto stage_diag = [&](int J0, auto first_c, auto nthr_c) {
    constexpr int FIRST = decltype(first_c)::value, NTHR = decltype(nthr_c)::value;
    constexpr int NPT = (CB * CB + NTHR - 1) / NTHR;
    const int jb = min(CB, k - J0);
    double *Ds = Ds0 + (size_t)((J0 / CB) & 1) * CB * TRS;
    T va[NPT];
    double vd[NPT];
#pragma unroll
    for (int e = 0; e < NPT; ++e) {
      const int idx = tid - FIRST + NTHR * e;
      const int lo = idx & (CB - 1), hi = idx >> 5;
      const bool in = idx < CB * CB;
      va[e] = (in && r0 + lo < m && hi < jb)
                  ? (A ? A[(r0 + lo) + (long)(J0 + hi) * lda] : T(r0 + lo == J0 + hi ? 1 : 0))
                  : T(0);
      vd[e] = (in && lo < jb && hi < jb) ? R[(J0 + lo) + (long)(J0 + hi) * k] : (lo == hi ? 1.0 : 0.0);
    }
#pragma unroll
    for (int e = 0; e < NPT; ++e) {
      const int idx = tid - FIRST + NTHR * e;
      const int lo = idx & (CB - 1), hi = idx >> 5;
      if (idx < CB * CB) {
        Xs[(J0 + hi) * TRS + lo] = va[e];
        Ds[lo * TRS + hi] = vd[e];
      }
    }
  };
  stage_diag(0, std::integral_constant<int, 0>{}, std::integral_constant<int, 256>{});
  if (rec) dbg[1] = __builtin_amdgcn_s_memrealtime();
  __syncthreads();
  for (int J0 = 0; J0 < k; J0 += CB) {
    const int jb = min(CB, k - J0);
    if (rec) dbg[2 + 8 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
    if (J0 > 0) {
      V acc = {0, 0, 0, 0};
      for (int kc = 0; kc < J0; kc += CB) {
#pragma unroll
        for (int kk = 0; kk < CB; kk += 4) {
          const int t = kc + kk + (lane >> 4);
          acc = MFf<T>::op(Xs[t * TRS + tr * 16 + (lane & 15)], Rs[t * TRS + tc * 16 + (lane & 15)], acc);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tr * 16 + MFf<T>::row(lane, v), col = tc * 16 + (lane & 15);
        Xs[(J0 + col) * TRS + row] -= acc[v];
      }
    }
    __syncthreads();  // B_J complete; Rs free
    if (rec) dbg[3 + 8 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
    if (wv == 0) {
      // lanes 32..63 repeat rows 0..31 (identical values; only lanes < 32 store to X)
      const double *Ds = Ds0 + (size_t)((J0 / CB) & 1) * CB * TRS;
      const int row = lane & (CB - 1);
      const double rinv = 1.0 / Ds[row * TRS + row];
      double b[CB];
#pragma unroll
      for (int l = 0; l < CB; ++l) b[l] = (double)Xs[(J0 + l) * TRS + row];
      if (rec) {
        __builtin_amdgcn_s_waitcnt(0);
        dbg[4 + 8 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
      }
      // row j + 1 of R_JJ is read while row j is applied
      double dc[CB], dn[CB];
#pragma unroll
      for (int l = 1; l < CB; ++l) dc[l] = Ds[l];
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        if (j + 1 < CB) {
#pragma unroll
          for (int l = j + 2; l < CB; ++l) dn[l] = Ds[(j + 1) * TRS + l];
        }
        const double x = b[j] * readlane_d(rinv, j);
        b[j] = x;
#pragma unroll
        for (int l = j + 1; l < CB; ++l) b[l] -= x * dc[l];
#pragma unroll
        for (int l = j + 2; l < CB; ++l) dc[l] = dn[l];
      }
      if (rec) dbg[5 + 8 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime() + (unsigned long long)(b[CB - 1] == 12345.0);
#pragma unroll
      for (int l = 0; l < CB; ++l) Xs[(J0 + l) * TRS + row] = (T)b[l];
      if (lane < CB && r0 + row < m) {
#pragma unroll
        for (int l = 0; l < CB; ++l)
          if (l < jb) X[(r0 + row) + (long)(J0 + l) * ldx] = (T)b[l];
      }
      if (rec) dbg[6 + 8 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
    } else if (J0 + CB < k) {
      const int J1 = J0 + CB, jb1 = min(CB, k - J1);
      stage_diag(J1, std::integral_constant<int, 64>{}, std::integral_constant<int, 192>{});
      // panel R(t, J1 + n), t < J1: 32 consecutive threads read 32 consecutive t
      // (k <= 256: at most 7168 elements, one batch of loads)
      const int tot = J1 * CB;
      constexpr int CH = (7 * CB * CB + 191) / 192;
      for (int base = tid - 64; base < tot; base += 192 * CH) {
        T v[CH];
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          const int idx = base + 192 * e;
          const int t = ((idx >> 10) << 5) + (idx & (CB - 1)), n = (idx >> 5) & (CB - 1);
          v[e] = (idx < tot && n < jb1) ? (T)R[t + (long)(J1 + n) * k] : T(0);
        }
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          const int idx = base + 192 * e;
          const int t = ((idx >> 10) << 5) + (idx & (CB - 1)), n = (idx >> 5) & (CB - 1);
          if (idx < tot) Rs[t * TRS + n] = v[e];
        }
      }
    }
    __syncthreads();  // X_J in Xs; block J+1 staged
    if (rec) dbg[7 + 8 * (J0 / CB)] = __builtin_amdgcn_s_memrealtime();
  }
  if (rec) dbg[2 + 8 * (kp / CB)] = __builtin_amdgcn_s_memrealtime();
}

// X = I + sign * U (U == nullptr: X = I)
__global__ void k_eye_plus(int k, double *__restrict__ X, const double *__restrict__ U, double sign) {
  const long total = (long)k * k;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long c = idx / k, r = idx - c * k;
    X[idx] = (r == c ? 1.0 : 0.0) + (U ? sign * U[idx] : 0.0);
  }
}
void eye_plus(hipStream_t s, int64_t k, double *X, const double *U, double sign) {
  const long total = k * k;
  const int blocks = (int)std::min<long>((total + 255) / 256, 1024);
  hipLaunchKernelGGL(k_eye_plus, dim3(blocks), dim3(256), 0, s, (int)k, X, U, sign);
  TTG_HIP(hipGetLastError());
}

template <class T>
void trsm_right_upper(hipStream_t s, int64_t m, int64_t k, const T *A, int64_t lda, const double *R,
                      T *X, int64_t ldx, unsigned long long *dbg, const double *Winv) {
  if (m <= 0 || k <= 0) return;
  const int64_t kp = (k + CB - 1) / CB * CB;
  ProfScope prof(KC_TRTRI, s, (double)m * k * k, (double)sizeof(T) * 2.0 * m * k + 8.0 * k * k);
  if constexpr (std::is_same<T, double>::value) {
    if (Winv && kp <= 256 && !dbg) {
      const size_t lds3 = trsm3_lds((int)kp);
      TTG_HIP(hipFuncSetAttribute((const void *)k_trsm3, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds3));
      hipLaunchKernelGGL(k_trsm3, dim3((unsigned)((m + CB - 1) / CB)), dim3(256), lds3, s, (int)m, (int)k, A, (long)lda,
                         R, Winv, X, (long)ldx, (const TrsmJob *)nullptr);
      TTG_HIP(hipGetLastError());
      return;
    }
  }
  TTG_CHECK(A || kp <= 256, TTG_BAD_ARG, "identity right-hand side needs k <= 256");
  if (kp <= 256) {
    const size_t lds2 = trsm2_lds((int)kp, sizeof(T));
    auto fn = dbg ? k_trsm2<T, true> : k_trsm2<T, false>;
    TTG_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2));
    hipLaunchKernelGGL(fn, dim3((unsigned)((m + CB - 1) / CB)), dim3(256), lds2, s, (int)m, (int)k,
                       A, (long)lda, R, X, (long)ldx, dbg);
    TTG_HIP(hipGetLastError());
    return;
  }
  const size_t lds = ((size_t)CB * kp + (kp <= 256 ? kp : CB) * CB) * sizeof(T) + CB * CB * sizeof(double);
  TTG_CHECK(lds <= 160 * 1024, TTG_BAD_ARG, "rank too large for the triangular solve");
  auto fn = dbg ? k_trsm<T, true> : k_trsm<T, false>;
  TTG_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(fn, dim3((unsigned)((m + CB - 1) / CB)), dim3(256), lds, s, (int)m, (int)k,
                     A, (long)lda, R, X, (long)ldx, dbg, (const TrsmJob *)nullptr);
  TTG_HIP(hipGetLastError());
}
// Batched generic k_trsm (FP64, every job's A given; any k whose staging fits
// LDS): grid (row blocks of the tallest job) x njobs.
void trsm_batched(hipStream_t s, const TrsmJob *jobs_dev, int njobs, int64_t mmax, int64_t kmax) {
  if (njobs <= 0 || mmax <= 0) return;
  const int64_t kp = (kmax + CB - 1) / CB * CB;
  const size_t lds = ((size_t)CB * kp + (kp <= 256 ? kp : CB) * CB) * sizeof(double) + CB * CB * sizeof(double);
  TTG_CHECK(lds <= 160 * 1024, TTG_BAD_ARG, "rank too large for the triangular solve");
  ProfScope prof(KC_TRTRI, s, (double)njobs * mmax * kmax * kmax, 8.0 * njobs * (2.0 * mmax * kmax + kmax * kmax));
  TTG_HIP(hipFuncSetAttribute((const void *)k_trsm<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  hipLaunchKernelGGL((k_trsm<double, false>), dim3((unsigned)((mmax + CB - 1) / CB), (unsigned)njobs), dim3(256), lds,
                     s, 0, 0, (const double *)nullptr, 0L, (const double *)nullptr, (double *)nullptr, 0L,
                     (unsigned long long *)nullptr, jobs_dev);
  TTG_HIP(hipGetLastError());
}
// Batched k_trsm3 (k <= 256 for every job, Winv from k_chol6): grid
// (row blocks of the tallest job) x njobs.
void trsm3_batched(hipStream_t s, const TrsmJob *jobs_dev, int njobs, int64_t mmax, int64_t kmax) {
  if (njobs <= 0 || mmax <= 0) return;
  const int64_t kp = (kmax + CB - 1) / CB * CB;
  TTG_CHECK(kp <= 256, TTG_BAD_ARG, "batched triangular solve supports k <= 256");
  const size_t lds3 = trsm3_lds((int)kp);
  ProfScope prof(KC_TRTRI, s, (double)njobs * mmax * kmax * kmax, 8.0 * njobs * (2.0 * mmax * kmax + kmax * kmax));
  TTG_HIP(hipFuncSetAttribute((const void *)k_trsm3, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds3));
  hipLaunchKernelGGL(k_trsm3, dim3((unsigned)((mmax + CB - 1) / CB), (unsigned)njobs), dim3(256), lds3, s, 0, 0,
                     (const double *)nullptr, 0L, (const double *)nullptr, (const double *)nullptr, (double *)nullptr,
                     0L, jobs_dev);
  TTG_HIP(hipGetLastError());
}

template void trsm_right_upper<double>(hipStream_t, int64_t, int64_t, const double *, int64_t, const double *, double *, int64_t, unsigned long long *, const double *);
template void trsm_right_upper<float>(hipStream_t, int64_t, int64_t, const float *, int64_t, const double *, float *, int64_t, unsigned long long *, const double *);

}  // namespace ttg
