// Persistent batched rounding: one workgroup rounds one whole TT tensor.
//
// A batch of many small-to-medium TT tensors (uneven d, n, r) is rounded by a
// single launch of k_round_persist: NT-thread workgroups, one per slot of
// device workspace, take tensors from a device work counter in longest-first
// order (the host sorts the batch by estimated QR work) until the batch is
// exhausted.  Each workgroup runs @tt_tensor/round.m for its tensor entirely
// on chip, with no host round trip:
//   * sweep 1 (round.m:35-53): Householder QR of the (r_i n_i) x r_{i+1}
//     unfolding (robust for the rank-deficient unfoldings of tensors built by
//     addition), nrm(i+1) = ||R||_F, R /= nrm, next core <- R * core_{i+1}
//     (FP64 MFMA), ranks capped at min(r_i n_i, r_{i+1});
//   * sweep 2 (round.m:59-89): SVD of the r_i x (n_i r_{i+1}) unfolding as a
//     Householder QR of its tall orientation and a one-sided cyclic Jacobi
//     SVD of the k x k R factor (k = min of the two dimensions), singular
//     values sorted on chip and truncated by my_chop2 (core/my_chop2.m:18-34)
//     with ep = eps/sqrt(d-1); core_i <- V', core_{i-1} <- core_{i-1} U S;
//   * normalisation and balancing (round.m:90-109), packing with the final
//     ranks.
// Working cores live in the caller's output region (its capacity is the
// QR-capped size, an upper bound of every intermediate core); per-slot scratch
// holds the QR operand, the Q factor, the Jacobi matrix [S; J] and small
// factors.  Every scratch use is checked against the slot's capacities; a
// tensor that would not fit is abandoned with an error status, never written
// out of bounds.
#include "ttg_internal.h"
#include "ttg_device.h"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace ttg {

typedef double v4d_p __attribute__((ext_vector_type(4)));

constexpr int PMAX_D = 64;        // tensors with more cores use the stream path
constexpr int PMAX_K = 512;       // largest bond rank

// Block-wide sum with a fixed combination order (deterministic).  `red` is
// used by this function only; its trailing barrier makes the next call's
// writes safe, so no leading barrier is needed.
template <int NT>
__device__ double p_block_sum(double v, double *red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

// Sum over a group of G consecutive lanes (G = 8: lanes 8j..8j+7, G = 16: a
// DPP row, G = 64: the wave), every lane of the group getting the
// bitwise-identical result; all 64 lanes must execute it.
template <int G> __device__ __forceinline__ double group_sum(double v) {
  if constexpr (G == 64) {
    return wave_sum(v);
  } else {
    v += dpp_d<0xB1>(v);   // quad_perm xor 1
    v += dpp_d<0x4E>(v);   // quad_perm xor 2
    v += dpp_d<0x141>(v);  // row_half_mirror: the 8-lane half-row
    if constexpr (G == 16) v += dpp_d<0x140>(v);  // row_mirror: the 16-lane row
    return v;
  }
}

template <int NT>
__device__ void p_copy(double *__restrict__ dst, const double *__restrict__ src, long count) {
  for (long i = threadIdx.x; i < count; i += NT) dst[i] = src[i];
  __syncthreads();
}

// dst (cols x rows) = src (rows x cols)^T, both column-major
template <int NT>
__device__ void p_transpose(double *__restrict__ dst, const double *__restrict__ src, long rows, long cols) {
  const long total = rows * cols;
  for (long idx = threadIdx.x; idx < total; idx += NT) {
    const long c = idx / rows, i = idx - c * rows;
    dst[c + i * cols] = src[idx];
  }
  __syncthreads();
}

// C (M x N, ldc) = alpha op(A) (M x K) op(B) (K x N) + beta C with FP64 MFMA
// (beta == 0: C is not read).  64 x 64 blocks of C; per 16-deep K step the
// op(A) and op(B) tiles are loaded cooperatively (coalesced along the
// contiguous index of each operand) into LDS (`lds`, 2560 doubles), the next
// step's tiles are prefetched into registers while the waves multiply; the
// waves form the transposed product (op(B)^T op(A)^T) so that 16 lanes store
// 16 consecutive rows of C.
constexpr int PG_BM = 64, PG_BN = 64, PG_BK = 16, PG_LD = 80;  // LD = 80: lanes 16-31 hit the other 32 banks
template <int NT>
__device__ __noinline__ void p_gemm(bool ta, bool tb, long M, long N, long K, const double *A, long lda,
                       const double *B, long ldb, double *C, long ldc, double *lds, double alpha = 1.0,
                       double beta = 0.0) {
  constexpr int NW = NT / 64, WN = NW / 2, NTW = (PG_BN / WN) / 16;  // n-tiles per wave
  constexpr int LA = PG_BM * PG_BK / NT, LB = PG_BK * PG_BN / NT;   // elements per thread
  double *As = lds, *Bs = lds + PG_BK * PG_LD;  // As[k][m], Bs[k][n]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = wv & 1, wn = wv >> 1;
  const long bm = (M + PG_BM - 1) / PG_BM, bn = (N + PG_BN - 1) / PG_BN;
  for (long blk = 0; blk < bm * bn; ++blk) {
    const long m0 = (blk % bm) * PG_BM, n0 = (blk / bm) * PG_BN;
    v4d_p acc[2][NTW];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < NTW; ++j) acc[i][j] = v4d_p{0, 0, 0, 0};
    double ra[LA], rb[LB];
    auto fetch = [&](long k0) {
#pragma unroll
      for (int e = 0; e < LA; ++e) {
        const int idx = tid + e * NT;
        const int mi = ta ? idx / PG_BK : idx % PG_BM, ki = ta ? idx % PG_BK : idx / PG_BM;
        const long gm = m0 + mi, gk = k0 + ki;
        ra[e] = (gm < M && gk < K) ? (ta ? A[gk + gm * lda] : A[gm + gk * lda]) : 0.0;
      }
#pragma unroll
      for (int e = 0; e < LB; ++e) {
        const int idx = tid + e * NT;
        const int ni = tb ? idx % PG_BN : idx / PG_BK, ki = tb ? idx / PG_BN : idx % PG_BK;
        const long gn = n0 + ni, gk = k0 + ki;
        rb[e] = (gn < N && gk < K) ? (tb ? B[gn + gk * ldb] : B[gk + gn * ldb]) : 0.0;
      }
    };
    auto stash = [&]() {
#pragma unroll
      for (int e = 0; e < LA; ++e) {
        const int idx = tid + e * NT;
        const int mi = ta ? idx / PG_BK : idx % PG_BM, ki = ta ? idx % PG_BK : idx / PG_BM;
        As[ki * PG_LD + mi] = ra[e];
      }
#pragma unroll
      for (int e = 0; e < LB; ++e) {
        const int idx = tid + e * NT;
        const int ni = tb ? idx % PG_BN : idx / PG_BK, ki = tb ? idx / PG_BN : idx % PG_BK;
        Bs[ki * PG_LD + ni] = rb[e];
      }
    };
    fetch(0);
    for (long k0 = 0; k0 < K; k0 += PG_BK) {
      __syncthreads();
      stash();
      __syncthreads();
      if (k0 + PG_BK < K) fetch(k0 + PG_BK);
#pragma unroll
      for (int kk = 0; kk < PG_BK; kk += 4) {
        const int kr = kk + lk;
        double av[2], bv[NTW];
#pragma unroll
        for (int i = 0; i < 2; ++i) av[i] = As[kr * PG_LD + wm * 32 + i * 16 + li];
#pragma unroll
        for (int j = 0; j < NTW; ++j) bv[j] = Bs[kr * PG_LD + wn * (16 * NTW) + j * 16 + li];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < NTW; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j], av[i], acc[i][j], 0, 0, 0);
      }
    }
    // acc[i][j][v] of lane l: C(m0 + wm*32 + i*16 + l%16, n0 + wn*16*NTW + j*16 + l/16 + 4v)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const long gm = m0 + wm * 32 + i * 16 + li;
      if (gm >= M) continue;
#pragma unroll
      for (int j = 0; j < NTW; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const long gn = n0 + wn * (16 * NTW) + j * 16 + lk + 4 * v;
          if (gn < N) {
            double *cp = C + gm + gn * ldc;
            *cp = (beta == 0.0) ? alpha * acc[i][j][v] : __builtin_fma(alpha, acc[i][j][v], beta * *cp);
          }
        }
    }
  }
  __syncthreads();
}

// Householder QR of A (m x k, column-major, ld m), in place: R in the upper
// trapezoid, the vectors v_j (v_j[j] = 1 implicit) below the diagonal, tau[j]
// (dgeqr2).  Then Q (m x kk, ld m, kk = min(m, k)) explicitly (dorg2r:
// Q = H_0 ... H_{kk-1} [I; 0], applied backwards).  One wave per column in
// the updates.
// H_j = I - t v v^T (v = [1; A(j+1:m, j)]) applied to the columns c0 .. k-1
// of X (ld m), rows j..m-1: each column by a group of G lanes (64 / G
// columns per wave; the group loop is uniform over the wave).
template <int NT, int G>
__device__ __forceinline__ void p_apply_reflector(const double *__restrict__ vj, double tj, long j, long m,
                                                  double *__restrict__ X, long c0, long k) {
  constexpr int PW = 64 / G;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gl = lane % G, ps = lane / G;
  for (long cb = c0 + (long)wv * PW; cb < k; cb += (long)(NT / 64) * PW) {
    const long c = cb + ps;
    const bool act = c < k;
    double *xc = X + (act ? c : c0) * m;
    double w = (act && gl == 0) ? xc[j] : 0.0;
    if (act)
      for (long i = j + 1 + gl; i < m; i += G) w = __builtin_fma(vj[i], xc[i], w);
    w = group_sum<G>(w) * tj;
    if (act) {
      if (gl == 0) xc[j] -= w;
      for (long i = j + 1 + gl; i < m; i += G) xc[i] = __builtin_fma(-w, vj[i], xc[i]);
    }
  }
}

template <int NT>
__device__ __noinline__ void p_householder(double *__restrict__ A, long m, long k, double *__restrict__ tau,
                              double *__restrict__ Q, double *red, double *bcast) {
  // columns of up to 256 rows take 16 lanes each (4 per wave), longer ones a wave
  const bool grp = m <= 256;
  const long kk = m < k ? m : k;
  for (long j = 0; j < kk; ++j) {
    double *aj = A + j * m;
    double s = 0.0;
    for (long i = j + 1 + threadIdx.x; i < m; i += NT) s = __builtin_fma(aj[i], aj[i], s);
    s = p_block_sum<NT>(s, red);
    const double al = aj[j];
    double tj = 0.0, scale = 0.0, beta = al;
    if (s > 0.0) {
      const double nrm = sqrt(__builtin_fma(al, al, s));
      beta = al >= 0.0 ? -nrm : nrm;
      tj = (beta - al) / beta;
      scale = 1.0 / (al - beta);
    }
    if (s > 0.0)
      for (long i = j + 1 + threadIdx.x; i < m; i += NT) aj[i] *= scale;
    __syncthreads();
    if (threadIdx.x == 0) {
      aj[j] = beta;
      tau[j] = tj;
    }
    // H_j on the trailing columns j+1 .. k-1
    if (tj != 0.0) {
      if (grp) p_apply_reflector<NT, 16>(aj, tj, j, m, A, j + 1, k);
      else p_apply_reflector<NT, 64>(aj, tj, j, m, A, j + 1, k);
    }
    __syncthreads();
  }
  (void)bcast;
  // Q = [I; 0] then H_j, j = kk-1 .. 0, on rows j.., columns j..kk-1
  for (long idx = threadIdx.x; idx < m * kk; idx += NT) {
    const long c = idx / m, i = idx - c * m;
    Q[idx] = (i == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (long j = kk - 1; j >= 0; --j) {
    const double tj = tau[j];
    const double *vj = A + j * m;
    if (tj != 0.0) {
      if (grp) p_apply_reflector<NT, 16>(vj, tj, j, m, Q, j, kk);
      else p_apply_reflector<NT, 64>(vj, tj, j, m, Q, j, kk);
    }
    __syncthreads();
  }
}

// The same factorisation with one workgroup barrier per column (m <= 256):
// the reflector of column j+1 is formed right after H_j reaches that column,
// by the 16-lane group that applied H_j to it (the first group of wave 0),
// while the other groups still apply H_j to the columns after it; the next
// step starts after the barrier.  (p_householder forms each reflector with
// a block-wide reduction between barriers: three or four per column.)  Same
// output as p_householder.
__device__ __forceinline__ void p_wave_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }
template <int NT>
__device__ __noinline__ void p_householder_la(double *__restrict__ A, long m, long k, double *__restrict__ tau,
                                              double *__restrict__ Q) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gl = lane & 15, ps = lane >> 4;
  const long kk = m < k ? m : k;
  // reflector j from column j (rows j..m-1) by lanes 0..15 of wave 0 (the
  // whole wave runs the group sum)
  auto prep = [&](long j) {
    double *aj = A + j * m;
    double s = 0.0;
    if (ps == 0)
      for (long i = j + 1 + gl; i < m; i += 16) s = __builtin_fma(aj[i], aj[i], s);
    s = group_sum<16>(s);
    if (ps == 0) {
      const double al = aj[j];
      double tj = 0.0, scale = 0.0, beta = al;
      if (s > 0.0) {
        const double nrm = sqrt(__builtin_fma(al, al, s));
        beta = al >= 0.0 ? -nrm : nrm;
        tj = (beta - al) / beta;
        scale = 1.0 / (al - beta);
        for (long i = j + 1 + gl; i < m; i += 16) aj[i] *= scale;
      }
      p_wave_sync();  // every lane of the group has read aj[j]
      if (gl == 0) {
        aj[j] = beta;
        tau[j] = tj;
      }
    }
  };
  if (wv == 0) prep(0);
  __syncthreads();
  for (long j = 0; j < kk; ++j) {
    const double tj = tau[j];
    if (tj != 0.0) p_apply_reflector<NT, 16>(A + j * m, tj, j, m, A, j + 1, k);
    if (j + 1 < kk && wv == 0) {
      p_wave_sync();  // column j+1 updated (its rows by several lanes of the group)
      prep(j + 1);
    }
    __syncthreads();
  }
  // Q = [I; 0] then H_j, j = kk-1 .. 0, on rows j.., columns j..kk-1
  for (long idx = threadIdx.x; idx < m * kk; idx += NT) {
    const long c = idx / m, i = idx - c * m;
    Q[idx] = (i == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (long j = kk - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj != 0.0) p_apply_reflector<NT, 16>(A + j * m, tj, j, m, Q, j, kk);
    __syncthreads();
  }
}

// Blocked Householder QR (dgeqrf/dorgqr structure, compact WY), same output
// as p_householder: 32-column panels factorised column by column (updates
// confined to the panel), then with the panel's explicit reflector block V
// (Vs, unit diagonal, zeros above; ld m - j0) and triangular factor T
// (dlarft, forward columnwise, from V^T V) the trailing columns get
// A2 <- A2 - V (T^T (V^T A2)) as three MFMA GEMMs; Q = H_0 ... H_{kk-1} [I; 0]
// is accumulated backwards panel by panel, Q2 <- Q2 - V (T (V^T Q2)).
// Scratch: Vs ((m - j0) x 32 <= m x k), W and W2 (32 x k each), Ts (32 x 32
// per panel), lds (2560 + 2 * 32 * 33 doubles).
constexpr int PH = 32;
template <int NT>
__device__ void p_hh_panel_vt(double *A, long m, long j0, int jb, const double *tau, double *Vs, double *Ts,
                              double *lds) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long rows = m - j0;
  for (long idx = threadIdx.x; idx < rows * jb; idx += NT) {
    const long c = idx / rows, i = idx - c * rows;  // row j0 + i of column j0 + c
    Vs[idx] = (i < c) ? 0.0 : (i == c ? 1.0 : A[j0 + i + (j0 + c) * m]);
  }
  __syncthreads();
  double *Gv = lds + 2560, *Tl = Gv + PH * (PH + 1);
  p_gemm<NT>(true, false, jb, jb, rows, Vs, rows, Vs, rows, Gv, PH + 1, lds);
  // T(i,i) = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i)  (one wave, lane r owns row r)
  if (wv == 0) {
    const int r = lane;
    for (int i = 0; i < jb; ++i) {
      const double ti = tau[j0 + i];
      double t = 0.0;
      if (r < i) {
        for (int l = r; l < i; ++l) t = __builtin_fma(Tl[r + l * (PH + 1)], Gv[l + i * (PH + 1)], t);
        t *= -ti;
      } else if (r == i) {
        t = ti;
      }
      if (r < PH) Tl[r + i * (PH + 1)] = (r <= i) ? t : 0.0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < PH * PH; idx += NT) {
    const int r = idx % PH, c = idx / PH;
    Ts[idx] = (r < jb && c < jb) ? Tl[r + c * (PH + 1)] : 0.0;
  }
  __syncthreads();
}

template <int NT>
__device__ __noinline__ void p_householder_blocked(double *__restrict__ A, long m, long k, double *__restrict__ tau,
                                      double *__restrict__ Q, double *red, double *Vs, double *W, double *W2,
                                      double *Tall, double *lds) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long kk = m < k ? m : k;
  for (long j0 = 0; j0 < kk; j0 += PH) {
    const int jb = (int)((kk - j0) < PH ? (kk - j0) : PH);
    for (long j = j0; j < j0 + jb; ++j) {
      double *aj = A + j * m;
      double s = 0.0;
      for (long i = j + 1 + threadIdx.x; i < m; i += NT) s = __builtin_fma(aj[i], aj[i], s);
      s = p_block_sum<NT>(s, red);
      const double al = aj[j];
      double tj = 0.0, scale = 0.0, beta = al;
      if (s > 0.0) {
        const double nrm = sqrt(__builtin_fma(al, al, s));
        beta = al >= 0.0 ? -nrm : nrm;
        tj = (beta - al) / beta;
        scale = 1.0 / (al - beta);
        for (long i = j + 1 + threadIdx.x; i < m; i += NT) aj[i] *= scale;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        aj[j] = beta;
        tau[j] = tj;
      }
      if (tj != 0.0) {
        for (long c = j + 1 + wv; c < j0 + jb; c += NT / 64) {
          double *ac = A + c * m;
          double w = (lane == 0) ? ac[j] : 0.0;
          for (long i = j + 1 + lane; i < m; i += 64) w = __builtin_fma(aj[i], ac[i], w);
          w = wave_sum(w) * tj;
          if (lane == 0) ac[j] -= w;
          for (long i = j + 1 + lane; i < m; i += 64) ac[i] = __builtin_fma(-w, aj[i], ac[i]);
        }
      }
      __syncthreads();
    }
    double *Ts = Tall + (j0 / PH) * PH * PH;
    p_hh_panel_vt<NT>(A, m, j0, jb, tau, Vs, Ts, lds);
    const long n2 = k - j0 - jb, rows = m - j0;
    if (n2 > 0) {
      double *A2 = A + j0 + (j0 + jb) * m;
      p_gemm<NT>(true, false, jb, n2, rows, Vs, rows, A2, m, W, jb, lds);
      p_gemm<NT>(true, false, jb, n2, jb, Ts, PH, W, jb, W2, jb, lds);
      p_gemm<NT>(false, false, rows, n2, jb, Vs, rows, W2, jb, A2, m, lds, -1.0, 1.0);
    }
  }
  for (long idx = threadIdx.x; idx < m * kk; idx += NT) {
    const long c = idx / m, i = idx - c * m;
    Q[idx] = (i == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (long j0 = ((kk - 1) / PH) * PH; j0 >= 0; j0 -= PH) {
    const int jb = (int)((kk - j0) < PH ? (kk - j0) : PH);
    const long rows = m - j0, nc = kk - j0;
    const double *Ts = Tall + (j0 / PH) * PH * PH;
    for (long idx = threadIdx.x; idx < rows * jb; idx += NT) {
      const long c = idx / rows, i = idx - c * rows;
      Vs[idx] = (i < c) ? 0.0 : (i == c ? 1.0 : A[j0 + i + (j0 + c) * m]);
    }
    __syncthreads();
    double *Q2 = Q + j0 + j0 * m;
    p_gemm<NT>(true, false, jb, nc, rows, Vs, rows, Q2, m, W, jb, lds);
    p_gemm<NT>(false, false, jb, nc, jb, Ts, PH, W, jb, W2, jb, lds);
    p_gemm<NT>(false, false, rows, nc, jb, Vs, rows, W2, jb, Q2, m, lds, -1.0, 1.0);
  }
}

// Rotation of columns with squared norms al, be and inner product ga (Rutishauser).
__device__ __forceinline__ void p_jrot(double al, double be, double ga, double &cs, double &sn) {
  const double dd = be - al, g2 = 2.0 * ga;
  const double t = (dd >= 0.0 ? g2 : -g2) / (fabs(dd) + sqrt(__builtin_fma(dd, dd, g2 * g2)));
  cs = 1.0 / sqrt(__builtin_fma(t, t, 1.0));
  sn = cs * t;
}

// One-sided cyclic Jacobi on Y = [S; J] (2k x k, ld 2k; S k x k, J = I on
// entry): round-robin parallel ordering, k/2 disjoint pairs per round, dots
// over S's k rows, rotation if |g| > tol sqrt(a b) and neither column is
// numerically zero (||s||^2 <= k (2u)^2 ||S||_F^2); a sweep none of whose
// rotations had cos^2 > 1e-16 ends the iteration.  Returns the number of
// sweeps (max_sweeps + 1 when not converged).  Each pair is handled by a
// group of G lanes (64 / G pairs per wave, so that the k/2 pairs of a round
// take as few passes as possible for small k: G = 8 for k <= 16, 16 for
// k <= 64, the whole wave above).
template <int NT, int G>
__device__ __noinline__ int p_jacobi_g(double *Y, int k, int max_sweeps, double *red, int *cnt) {
  constexpr int PW = 64 / G;  // pairs per wave
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gl = lane % G, ps = lane / G;
  const long ld = 2L * k;
  double fro = 0.0;
  for (long idx = threadIdx.x; idx < (long)k * k; idx += NT) {
    const long c = idx / k, i = idx - c * k;
    const double v = Y[i + c * ld];
    fro = __builtin_fma(v, v, fro);
  }
  fro = p_block_sum<NT>(fro, red);
  const double small2 = (double)k * 4.0 * DBL_EPSILON * DBL_EPSILON * fro;
  const double tol = sqrt((double)k) * DBL_EPSILON, tol2 = tol * tol;
  const int P = k + (k & 1);  // players (one bye per round when k is odd)
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (threadIdx.x == 0) cnt[0] = 0;
    __syncthreads();
    int mybig = 0;
    for (int rnd = 0; rnd < P - 1; ++rnd) {
      // the group loop is uniform over the wave (every lane runs the sums)
      for (int q0 = wv * PW; q0 < P / 2; q0 += (NT / 64) * PW) {
        const int q = q0 + ps;
        int a = P, b = P;
        if (q < P / 2) {
          if (q == 0) {
            a = P - 1;
            b = rnd;
          } else {
            a = (rnd + q) % (P - 1);
            b = (rnd - q + P - 1) % (P - 1);
          }
        }
        const bool act = a < k && b < k;
        double *ya = Y + (long)(act ? a : 0) * ld, *yb = Y + (long)(act ? b : 0) * ld;
        double al = 0.0, be = 0.0, ga = 0.0;
        if (act) {
          for (int i = gl; i < k; i += G) {
            const double x = ya[i], y = yb[i];
            al = __builtin_fma(x, x, al);
            be = __builtin_fma(y, y, be);
            ga = __builtin_fma(x, y, ga);
          }
        }
        al = group_sum<G>(al);
        be = group_sum<G>(be);
        ga = group_sum<G>(ga);
        const double g2 = ga * ga, ab = al * be;
        const bool rot = act && g2 > tol2 * ab && ab >= DBL_MIN && al > small2 && be > small2;
        if (rot) {
          if (gl == 0) mybig += (g2 > 1e-16 * ab) ? 1 : 0;
          double cs, sn;
          p_jrot(al, be, ga, cs, sn);
          for (long i = gl; i < ld; i += G) {
            const double x = ya[i], y = yb[i];
            ya[i] = cs * x - sn * y;
            yb[i] = sn * x + cs * y;
          }
        }
      }
      __syncthreads();
    }
    if (gl == 0 && mybig > 0) atomicAdd(cnt, mybig);
    __syncthreads();
    const int big = cnt[0];
    __syncthreads();
    if (big == 0) return sweep + 1;
  }
  return max_sweeps + 1;
}
template <int NT>
__device__ int p_jacobi(double *Y, int k, int max_sweeps, double *red, int *cnt) {
  if (k <= 16) return p_jacobi_g<NT, 8>(Y, k, max_sweeps, red, cnt);
  if (k <= 64) return p_jacobi_g<NT, 16>(Y, k, max_sweeps, red, cnt);
  return p_jacobi_g<NT, 64>(Y, k, max_sweeps, red, cnt);
}

// The same iteration for a Y = [S; J] (global, ld 2k) too large for LDS:
// block-cyclic ordering.  Columns form blocks of bw; the blocks play a
// round-robin tournament, and each block pair is staged in LDS (2 bw
// columns of 2k rows), rotated there and written back.  In the first block
// round of a sweep every pair inside the two blocks is visited (2 bw - 1
// rounds), later only the bw x bw cross pairs (bw rounds), so a sweep
// visits every column pair once.  Same rotation rule and stopping test as
// p_jacobi.
template <int NT>
__device__ __noinline__ int p_jacobi_blocks(double *Yg, int k, int bw, double *L, int max_sweeps, double *red, int *cnt) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long ld = 2L * k, ldl = 2L * k + 1;  // odd LDS column stride
  double fro = 0.0;
  for (long idx = threadIdx.x; idx < (long)k * k; idx += NT) {
    const long c = idx / k, i = idx - c * k;
    const double v = Yg[i + c * ld];
    fro = __builtin_fma(v, v, fro);
  }
  fro = p_block_sum<NT>(fro, red);
  const double small2 = (double)k * 4.0 * DBL_EPSILON * DBL_EPSILON * fro;
  const double tol = sqrt((double)k) * DBL_EPSILON, tol2 = tol * tol;
  const int nb = (k + bw - 1) / bw, PB = nb + (nb & 1), LC = 2 * bw;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (threadIdx.x == 0) cnt[0] = 0;
    __syncthreads();
    int mybig = 0;
    for (int bt = 0; bt < PB - 1; ++bt) {
      for (int bq = 0; bq < PB / 2; ++bq) {
        int ba, bb;
        if (bq == 0) {
          ba = PB - 1;
          bb = bt;
        } else {
          ba = (bt + bq) % (PB - 1);
          bb = (bt - bq + PB - 1) % (PB - 1);
        }
        const bool va = ba < nb, vb = bb < nb;
        if (!(va && vb) && !(bt == 0 && (va || vb))) continue;  // a lone block only needs its inner pairs
        // local column l < bw: column ba*bw + l; l >= bw: column bb*bw + l - bw (-1: absent)
        auto gcol = [&](int l) -> int {
          const int blk = l < bw ? ba : bb, off = l < bw ? l : l - bw;
          const bool v = l < bw ? va : vb;
          const int c = blk * bw + off;
          return (v && c < k) ? c : -1;
        };
        for (long idx = threadIdx.x; idx < (long)LC * ld; idx += NT) {
          const int l = (int)(idx / ld);
          const long i = idx - (long)l * ld;
          const int c = gcol(l);
          L[l * ldl + i] = c >= 0 ? Yg[i + (long)c * ld] : 0.0;
        }
        __syncthreads();
        const int rounds = (bt == 0) ? LC - 1 : bw;
        for (int t = 0; t < rounds; ++t) {
          for (int q = wv; q < bw; q += NT / 64) {
            int a, b;
            if (bt == 0) {
              if (q == 0) {
                a = LC - 1;
                b = t;
              } else {
                a = (t + q) % (LC - 1);
                b = (t - q + LC - 1) % (LC - 1);
              }
            } else {
              a = q;
              b = bw + ((q + t) % bw);
            }
            if (gcol(a) < 0 || gcol(b) < 0) continue;
            double *ya = L + (long)a * ldl, *yb = L + (long)b * ldl;
            double al = 0.0, be = 0.0, ga = 0.0;
            for (int i = lane; i < k; i += 64) {
              const double x = ya[i], y = yb[i];
              al = __builtin_fma(x, x, al);
              be = __builtin_fma(y, y, be);
              ga = __builtin_fma(x, y, ga);
            }
            al = wave_sum(al);
            be = wave_sum(be);
            ga = wave_sum(ga);
            const double g2 = ga * ga, ab = al * be;
            const bool rot = g2 > tol2 * ab && ab >= DBL_MIN && al > small2 && be > small2;
            if (!rot) continue;
            mybig += (g2 > 1e-16 * ab) ? 1 : 0;
            double cs, sn;
            p_jrot(al, be, ga, cs, sn);
            for (long i = lane; i < ld; i += 64) {
              const double x = ya[i], y = yb[i];
              ya[i] = cs * x - sn * y;
              yb[i] = sn * x + cs * y;
            }
          }
          __syncthreads();
        }
        for (long idx = threadIdx.x; idx < (long)LC * ld; idx += NT) {
          const int l = (int)(idx / ld);
          const long i = idx - (long)l * ld;
          const int c = gcol(l);
          if (c >= 0) Yg[i + (long)c * ld] = L[l * ldl + i];
        }
        __syncthreads();
      }
    }
    if (lane == 0 && mybig > 0) atomicAdd(cnt, mybig);
    __syncthreads();
    const int big = cnt[0];
    __syncthreads();
    if (big == 0) return sweep + 1;
  }
  return max_sweeps + 1;
}

// Column norms of Y's top k rows, sorted descending (ties by index), and
// my_chop2(sv, ||sv|| ep) capped at rmax (core/my_chop2.m:18-34); perm maps
// sorted position -> column.  Returns r1.
template <int NT>
__device__ int p_sort_chop(const double *Y, long ldy, int k, double ep, double *sig, double *sv,
                           int *perm, int *bc) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = wv; j < k; j += NT / 64) {
    const double *col = Y + (long)j * ldy;
    double s = 0.0;
    for (int i = lane; i < k; i += 64) s = __builtin_fma(col[i], col[i], s);
    s = wave_sum(s);
    if (lane == 0) sig[j] = sqrt(s);
  }
  for (int j = threadIdx.x; j < k; j += NT) perm[j] = j;  // stays a valid index even for NaN norms
  __syncthreads();
  for (int j = threadIdx.x; j < k; j += NT) {
    const double v = sig[j];
    int pos = 0;
    for (int i = 0; i < k; ++i) {
      const double u = sig[i];
      pos += (u > v) || (u == v && i < j);
    }
    sv[pos] = v;
    perm[pos] = j;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ss = 0.0;
    for (int j = 0; j < k; ++j) ss += sv[j] * sv[j];
    const double nrm = sqrt(ss);
    const double e = nrm * ep;
    int r;
    if (nrm == 0.0) {
      r = 1;
    } else if (!(e > 0.0)) {
      r = k;
    } else {
      const double e2 = e * e;
      double c = 0.0;
      int K = 0;
      for (int j = k - 1; j >= 0; --j) {
        c += sv[j] * sv[j];
        if (c < e2) K = k - j;
      }
      r = k - K;
    }
    bc[0] = r;
  }
  __syncthreads();
  const int r1 = bc[0];
  __syncthreads();
  return r1;
}

// Per-slot scratch (doubles): Wa, Wq, Vs (core-sized), Y (2 kmax x kmax),
// Ra, Rb (kmax^2), tau (kmax), W, W2 (32 kmax), T factors (32 kmax + 1024).
struct PSlot {
  long core, kmax;
  __host__ __device__ long elems() const {
    return 3 * core + 4 * kmax * kmax + kmax + 3 * 32 * kmax + 1024 + 64;
  }
};

template <int NT>
__global__ __launch_bounds__(NT) void k_round_persist(int count, const PTensorDesc *__restrict__ tens,
                                                      const int *__restrict__ order,
                                                      const long *__restrict__ nall,
                                                      const long *__restrict__ rall,
                                                      const double *__restrict__ in, double *__restrict__ out,
                                                      long *__restrict__ r_out, double *__restrict__ ws,
                                                      PSlot slot, int *ctr, int *tstat, double eps,
                                                      int max_sweeps, long ycap) {
  __shared__ double red[NT / 64 + 2];
  __shared__ double sig[PMAX_K], sv[PMAX_K], nrm[PMAX_D + 2];
  __shared__ int perm[PMAX_K];
  __shared__ long rc[PMAX_D + 1], nn[PMAX_D], wofs[PMAX_D + 1];
  __shared__ int sh_t, sh_bc[2], sh_fail;
  extern __shared__ __align__(16) double ylds[];
  double *Wa = ws + (long)blockIdx.x * slot.elems();
  double *Wq = Wa + slot.core;
  double *Yg = Wq + slot.core;
  double *Ra = Yg + 2 * slot.kmax * slot.kmax;
  double *Rb = Ra + slot.kmax * slot.kmax;
  double *tau = Rb + slot.kmax * slot.kmax;
  double *Vs = tau + slot.kmax;
  double *Wg = Vs + slot.core;
  double *Wg2 = Wg + 32 * slot.kmax;
  double *Tall = Wg2 + 32 * slot.kmax;
  for (;;) {
    if (threadIdx.x == 0) sh_t = atomicAdd(ctr, 1);
    __syncthreads();
    const int ti = sh_t;
    __syncthreads();
    if (ti >= count) break;
    const int t = order[ti];
    const PTensorDesc td = tens[t];
    const int d = td.d;
    const double *xin = in + td.in_ofs;
    double *xout = out + td.out_ofs;
    int sweeps_tot = 0, sweeps_max = 0, status = 0;
    // ranks, modes, offsets of the input cores and of the (QR-capped) working slots
    if (threadIdx.x == 0) {
      sh_fail = 0;
      long o = 0;
      for (int i = 0; i < d; ++i) nn[i] = nall[td.nofs + i];
      for (int i = 0; i <= d; ++i) rc[i] = rall[td.rofs + i];
      // capped ranks (round.m:48) give the working layout
      long rq = rc[0];
      for (int i = 0; i < d; ++i) {
        const long rn = (i + 1 < d) ? ((rq * nn[i] < rc[i + 1]) ? rq * nn[i] : rc[i + 1]) : rc[d];
        wofs[i] = o;
        o += rq * nn[i] * rn;
        rq = rn;
      }
      wofs[d] = o;
    }
    __syncthreads();
    // non-finite input is an error for the tensor (like round_dev)
    {
      long insize = 0;
      for (int i = 0; i < d; ++i) insize += rc[i] * nn[i] * rc[i + 1];
      int bad = 0;
      for (long i = threadIdx.x; i < insize; i += NT) bad |= !isfinite(xin[i]);
      if (bad) atomicOr(&sh_fail, 2);
      __syncthreads();
    }
    if (sh_fail) {
      status = sh_fail;
    } else if (d == 1) {
      const long sz = rc[0] * nn[0] * rc[1];
      if (threadIdx.x == 0) nrm[0] = 0.0;
      p_copy<NT>(xout, xin, sz);
    } else {
      // ---- sweep 1: L -> R Householder QR (round.m:35-53) ------------------
      {
        const long sz0 = rc[0] * nn[0] * rc[1];
        if (sz0 > slot.core) atomicOr(&sh_fail, 4);
        __syncthreads();
        if (!sh_fail) p_copy<NT>(Wa, xin, sz0);
      }
      long in_ofs_next = rc[0] * nn[0] * rc[1];
      for (int i = 0; i + 1 < d && !sh_fail; ++i) {
        const long m = rc[i] * nn[i], k = rc[i + 1], kk = m < k ? m : k;
        const long N1 = nn[i + 1] * rc[i + 2];
        if (m * kk > slot.core || kk * N1 > slot.core || kk * k > slot.kmax * slot.kmax || kk > slot.kmax) {
          if (threadIdx.x == 0) sh_fail = 4;
          __syncthreads();
          break;
        }
        // Wa holds core i as (m x k); Q -> working slot i (m x kk), R -> Ra (kk x k).
        // Small bonds are factorised in LDS (each column step is a chain of
        // dependent reductions, so memory latency dominates).
        const double *Af = Wa;
        if (k <= 64 && m * (k + kk) <= ycap) {
          p_copy<NT>(ylds, Wa, m * k);
          if (m <= 256) p_householder_la<NT>(ylds, m, k, tau, ylds + m * k);
          else p_householder<NT>(ylds, m, k, tau, ylds + m * k, red, nullptr);
          p_copy<NT>(xout + wofs[i], ylds + m * k, m * kk);
          Af = ylds;
        } else if (k <= 64) {
          p_householder<NT>(Wa, m, k, tau, xout + wofs[i], red, nullptr);
        } else {
          p_householder_blocked<NT>(Wa, m, k, tau, xout + wofs[i], red, Vs, Wg, Wg2, Tall, ylds);
        }
        for (long idx = threadIdx.x; idx < kk * k; idx += NT) {
          const long c = idx / kk, r = idx - c * kk;
          Ra[idx] = (r <= c) ? Af[r + c * m] : 0.0;
        }
        __syncthreads();
        double s = 0.0;
        for (long idx = threadIdx.x; idx < kk * k; idx += NT) s = __builtin_fma(Ra[idx], Ra[idx], s);
        s = p_block_sum<NT>(s, red);
        const double nv = sqrt(s);
        if (threadIdx.x == 0) nrm[i + 1] = nv;
        if (nv != 0.0) {
          const double inv = 1.0 / nv;
          for (long idx = threadIdx.x; idx < kk * k; idx += NT) Ra[idx] *= inv;
        }
        __syncthreads();
        // core_{i+1} <- R core_{i+1} (input core i+1 is k x N1), into Wa
        p_gemm<NT>(false, false, kk, N1, k, Ra, kk, xin + in_ofs_next, k, Wa, kk, ylds);
        in_ofs_next += k * N1;
        if (threadIdx.x == 0) rc[i + 1] = kk;
        __syncthreads();
      }
      if (!sh_fail) {
        // the last core of sweep 1 -> its working slot
        const long szl = rc[d - 1] * nn[d - 1] * rc[d];
        p_copy<NT>(xout + wofs[d - 1], Wa, szl);
      }
      // ---- sweep 2: R -> L truncated SVDs (round.m:59-89) -------------------
      const double ep = eps / sqrt((double)(d - 1));
      for (int i = d - 1; i >= 1 && !sh_fail; --i) {
        const long kw = rc[i], N = nn[i] * rc[i + 1];
        const bool wide = N >= kw;
        const long kj = wide ? kw : N, mt = wide ? N : kw;
        double *Wc = xout + wofs[i];
        if (mt * kj > slot.core || kj > slot.kmax || kj > PMAX_K) {
          if (threadIdx.x == 0) sh_fail = 4;
          __syncthreads();
          break;
        }
        // tall orientation B (mt x kj) in Wa, QR: R in Wa's upper part, Q -> Wq
        if (wide) p_transpose<NT>(Wa, Wc, kw, N);
        else p_copy<NT>(Wa, Wc, kw * N);
        // R (kj x kj) -> Ra, in LDS for small bonds as in sweep 1
        if (kj <= 64 && mt * 2 * kj <= ycap) {
          p_copy<NT>(ylds, Wa, mt * kj);
          if (mt <= 256) p_householder_la<NT>(ylds, mt, kj, tau, ylds + mt * kj);
          else p_householder<NT>(ylds, mt, kj, tau, ylds + mt * kj, red, nullptr);
          p_copy<NT>(Wq, ylds + mt * kj, mt * kj);
          for (long idx = threadIdx.x; idx < kj * kj; idx += NT) {
            const long c = idx / kj, r = idx - c * kj;
            Ra[idx] = (r <= c) ? ylds[r + c * mt] : 0.0;
          }
        } else {
          if (kj <= 64) p_householder<NT>(Wa, mt, kj, tau, Wq, red, nullptr);
          else p_householder_blocked<NT>(Wa, mt, kj, tau, Wq, red, Vs, Wg, Wg2, Tall, ylds);
          for (long idx = threadIdx.x; idx < kj * kj; idx += NT) {
            const long c = idx / kj, r = idx - c * kj;
            Ra[idx] = (r <= c) ? Wa[r + c * mt] : 0.0;
          }
        }
        __syncthreads();
        // Y = [S; I] with S = R^T (wide) or R (tall): in LDS when it fits,
        // else in the slot, rotated block pair by block pair through LDS
        const long ldy = 2 * kj;
        const bool in_lds = ldy * kj <= ycap;
        int bw = 0;
        if (!in_lds) bw = (32 * (ldy + 1) <= ycap) ? 16 : ((16 * (ldy + 1) <= ycap) ? 8 : 0);
        double *Y = in_lds ? ylds : Yg;
        for (long idx = threadIdx.x; idx < ldy * kj; idx += NT) {
          const long c = idx / ldy, r = idx - c * ldy;
          double v;
          if (r < kj) {
            const long rr = wide ? c : r, cc = wide ? r : c;  // S(r, c) = R(c, r) when wide
            v = Ra[rr + cc * kj];
          } else {
            v = (r - kj == c) ? 1.0 : 0.0;
          }
          Y[idx] = v;
        }
        __syncthreads();
        const int sw = bw ? p_jacobi_blocks<NT>(Y, (int)kj, bw, ylds, max_sweeps, red, sh_bc)
                          : p_jacobi<NT>(Y, (int)kj, max_sweeps, red, sh_bc);
        if (sw > max_sweeps) status |= 1;
        sweeps_tot += sw > max_sweeps ? max_sweeps : sw;
        sweeps_max = sw > sweeps_max ? sw : sweeps_max;
        const int r1 = p_sort_chop<NT>(Y, ldy, (int)kj, ep, sig, sv, perm, sh_bc);
        // U S and V from Y = [S J; J] (column perm[c] is the c-th singular triple)
        if (wide) {
          // W^T = Q R, R^T J = Y_top: U S = Y_top[:, perm] (kw x r1) -> Rb,
          // V = Q J[:, perm]: core_i = J_r^T Q^T (r1 x N)
          for (long idx = threadIdx.x; idx < kw * r1; idx += NT) {
            const long c = idx / kw, r = idx - c * kw;
            const long src = r + (long)perm[c] * ldy;
            Rb[idx] = Y[src];
            Ra[idx] = Y[src + kj];  // J_r (kw x r1)
          }
          __syncthreads();
          p_gemm<NT>(true, true, r1, N, kw, Ra, kw, Wq, N, Wc, r1, ylds);
        } else {
          // W = Q R, R J = Y_top: U S = Q Y_top[:, perm] (kw x r1), V = J[:, perm]
          for (long idx = threadIdx.x; idx < N * r1; idx += NT) {
            const long c = idx / N, r = idx - c * N;
            const long src = r + (long)perm[c] * ldy;
            Ra[idx] = Y[src];            // Y_r (N x r1)
            Wa[c + r * r1] = Y[src + kj];  // core_i = J_r^T (r1 x N)
          }
          __syncthreads();
          p_gemm<NT>(false, false, kw, r1, N, Wq, kw, Ra, N, Rb, kw, ylds);
          p_copy<NT>(Wc, Wa, r1 * N);
        }
        // core_{i-1} <- core_{i-1} (U S)   (round.m:81-83), via Wa
        const long mp = rc[i - 1] * nn[i - 1];
        p_gemm<NT>(false, false, mp, r1, kw, xout + wofs[i - 1], mp, Rb, kw, Wa, mp, ylds);
        p_copy<NT>(xout + wofs[i - 1], Wa, mp * r1);
        if (threadIdx.x == 0) rc[i] = r1;
        __syncthreads();
      }
    }
    if (sh_fail) status |= sh_fail;
    if (!(status & ~1)) {
      // ---- nrm(1), balance, pack (round.m:90-109) ---------------------------
      const long s0 = rc[0] * nn[0] * rc[1];
      double s = 0.0;
      for (long idx = threadIdx.x; idx < s0; idx += NT) s = __builtin_fma(xout[idx], xout[idx], s);
      s = p_block_sum<NT>(s, red);
      double sc0 = 1.0, sc = 1.0;
      if (d > 1) {
        nrm[0] = sqrt(s);  // every thread holds the same value
        double lsum = 0.0;
        for (int i = 0; i < d; ++i) lsum += log(fabs(nrm[i]));
        const double nrm0 = exp(lsum / d);
        sc0 = (nrm[0] != 0.0) ? nrm0 / nrm[0] : nrm0;
        sc = nrm0;
      }
      __syncthreads();
      // packed offsets never exceed the working ones: copy forward in chunks
      long o = 0;
      for (int i = 0; i < d; ++i) {
        const long sz = rc[i] * nn[i] * rc[i + 1];
        const double f = (i == 0) ? sc0 : sc;
        const double *src = xout + wofs[i];
        double *dst = xout + o;
        for (long base = 0; base < sz; base += NT) {
          const long idx = base + threadIdx.x;
          const double v = idx < sz ? src[idx] : 0.0;
          __syncthreads();
          if (idx < sz) dst[idx] = f * v;
          __syncthreads();
        }
        o += sz;
      }
      for (int i = threadIdx.x; i <= d; i += NT) r_out[td.rofs + i] = rc[i];
    }
    if (threadIdx.x == 0) {
      tstat[4 * t + 0] = status;
      tstat[4 * t + 1] = sweeps_tot;
      tstat[4 * t + 2] = sweeps_max;
    }
    __syncthreads();
  }
}

__global__ void k_persist_reset(int *ctr) {
  if (threadIdx.x == 0) ctr[0] = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool persist_supported(int64_t d, const int64_t *n, const int64_t *r) {
  if (d < 1 || d > PMAX_D) return false;
  std::vector<int64_t> rc(r, r + d + 1);
  for (int64_t i = 0; i + 1 < d; ++i) rc[i + 1] = std::min(rc[i] * n[i], rc[i + 1]);
  for (int64_t i = 0; i <= d; ++i)
    if (r[i] > PMAX_K) return false;
  for (int64_t i = 0; i < d; ++i)
    if (std::min(rc[i], n[i] * rc[i + 1]) > PMAX_K) return false;
  return true;
}

void persist_prepare(int nt, int64_t ycap) {
  const size_t lds = (size_t)ycap * sizeof(double);
  if (nt == 512)
    TTG_HIP(hipFuncSetAttribute((const void *)k_round_persist<512>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  else
    TTG_HIP(hipFuncSetAttribute((const void *)k_round_persist<256>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

void persist_launch(hipStream_t s, int nt, int slots, int count, const PTensorDesc *tens,
                    const int *order, const int64_t *nall, const int64_t *rall, const double *in,
                    double *out, int64_t *r_out, double *ws, int64_t slot_core, int64_t slot_kmax, int *ctr,
                    int *tstat, double eps, int max_sweeps, int64_t ycap) {
  PSlot sl{(long)slot_core, (long)slot_kmax};
  hipLaunchKernelGGL(k_persist_reset, dim3(1), dim3(64), 0, s, ctr);
  TTG_HIP(hipGetLastError());
  const size_t lds = (size_t)ycap * sizeof(double);
  if (nt == 512) {
    hipLaunchKernelGGL((k_round_persist<512>), dim3(slots), dim3(512), lds, s, count, tens, order,
                       (const long *)nall, (const long *)rall, in, out, (long *)r_out, ws, sl, ctr, tstat, eps,
                       max_sweeps, (long)ycap);
  } else {
    hipLaunchKernelGGL((k_round_persist<256>), dim3(slots), dim3(256), lds, s, count, tens, order,
                       (const long *)nall, (const long *)rall, in, out, (long *)r_out, ws, sl, ctr, tstat, eps,
                       max_sweeps, (long)ycap);
  }
  TTG_HIP(hipGetLastError());
}

int64_t persist_slot_elems(int64_t slot_core, int64_t slot_kmax) {
  return PSlot{(long)slot_core, (long)slot_kmax}.elems();
}

}  // namespace ttg
