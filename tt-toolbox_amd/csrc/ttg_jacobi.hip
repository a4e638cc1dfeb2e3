// One-sided block Jacobi SVD (Hestenes rotations, Rutishauser formulas) on gfx950.
//
// Problem: Y = [S; J] (2k rows x kpad columns, FP64, column-major); S is the
// k x k matrix being orthogonalised, J accumulates the right rotations
// (starts as the identity).  At convergence S J0 = Y_top has orthogonal
// columns (norms = singular values) and J is orthogonal.
//
// Parallel schedule: 2p column blocks of b columns, p cooperating
// workgroups; each outer step gives workgroup w one block pair (round-robin
// "circle" schedule), 2p-1 outer steps make one sweep.  Inside a step:
//   * the 2b columns of S (not J) are loaded into LDS;
//   * scalar rotations, b disjoint pairs per round, one wave per pair, the
//     pair held in registers.  In the first outer step of a sweep every pair
//     inside the two blocks is visited (2b-1 rounds); later steps visit only
//     the b*b cross pairs (b rounds), so a sweep visits every pair once in
//     k-1 rounds, a cyclic Jacobi sweep in a parallel ordering.  Dots
//     alpha, beta, gamma of the current columns (no Gram updates, so
//     accuracy is that of one-sided Jacobi), rotation if
//     |gamma| > tol*sqrt(alpha*beta);
//   * the rotations are also accumulated into V (2b x 2b, LDS) and, if any
//     happened, J's 2b columns are updated once as J <- J V with FP64 MFMA.
// A sweep without rotations ends the iteration.  The accumulated J drifts
// from orthogonality by O(u) per update; one Newton-Schulz polar step
// M = (3I - J^T J)/2, [S J; J] <- [S J; J] M afterwards makes the drift O(u^2)
// and keeps Y_top J^T = S^T to working accuracy.
#include "ttg_internal.h"
#include "ttg_device.h"
#include <cmath>
#include <cfloat>
#include <algorithm>

namespace ttg {

typedef double v4d_j __attribute__((ext_vector_type(4)));

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

// Every wave of the workgroup waits until its own stores have reached L2
// (s_waitcnt vmcnt(0); a workgroup barrier alone does not wait for them), so
// that one thread's agent-scope release after the barrier (the L2 write-back)
// publishes the whole workgroup's stores device-wide.
__device__ __forceinline__ void wave_stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Software grid barrier for a cooperative launch: bar[0] arrivals, bar[1]
// generation.  Bounded spin: after ~4 s a timeout flag is raised and every
// workgroup leaves the kernel at its next check.
__device__ bool grid_sync(unsigned *bar, unsigned nblocks, unsigned &gen, int *err) {
  __shared__ int sh_ok;
  wave_stores_done();
  __syncthreads();
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
        __builtin_amdgcn_s_sleep(1);
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

// 1/x and 1/sqrt(x) from the hardware estimates (about 23 correct bits) and
// two Newton steps each: full double precision for the normal, positive
// arguments that occur in a rotation, without the special-case handling of a
// general division or square root.
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(y, e, y);
}
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double e = __builtin_fma(-h * y, y, 0.5);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-h * y, y, 0.5);
  return __builtin_fma(y, e, y);
}

// Rotation that makes columns with squared norms al, be and inner product ga
// orthogonal (Rutishauser): with d = be - al, t = sign(d) 2ga / (|d| + sqrt(d^2
// + 4ga^2)), the smaller root of t^2 + 2 zeta t - 1 = 0, zeta = d / (2ga); t
// only needs ~46 bits (one Newton step on each estimate), c = 1/sqrt(1+t^2)
// is refined to full precision so that c^2 + s^2 = 1 to rounding.
__device__ __forceinline__ void jrot(double al, double be, double ga, double &cs, double &sn,
                                     double &tn) {
  const double d = be - al, g2 = 2.0 * ga;
  const double q = __builtin_fma(d, d, g2 * g2);
  double y = __builtin_amdgcn_rsq(q);
  y = y * __builtin_fma(-0.5 * q * y, y, 1.5);
  const double den = fabs(d) + q * y;
  double inv = __builtin_amdgcn_rcp(den);
  inv = inv * __builtin_fma(-den, inv, 2.0);
  const double t = (d >= 0.0 ? g2 : -g2) * inv;
  cs = rsq_nr(__builtin_fma(t, t, 1.0));
  sn = cs * t;
  tn = t;
}

// Layout of the barrier/statistics buffer: [0] arrivals, [1] generation,
// [2, 2+ms) rotations per sweep, then per sweep the number of rotations whose
// cos^2 exceeded the stopping threshold.
static __host__ __device__ inline int jacobi_mc_offset(int max_sweeps) { return (2 + max_sweeps + 1) & ~1; }
// then one double: ||S||_F^2 (8-byte aligned: the offset is even)
static __host__ __device__ inline int jacobi_fro_offset(int max_sweeps) { return jacobi_mc_offset(max_sweeps) + 2 * max_sweeps; }
// then the J workers' barrier (2 words), the V publication flags (2 x 64, by
// step parity) and their acknowledgements (64), see k_jacobi4 with SJ = true
static __host__ __device__ inline int jacobi_jw_offset(int max_sweeps) { return jacobi_fro_offset(max_sweeps) + 2; }
constexpr int JW_MAXP = 64;
int jacobi_bar_words(int max_sweeps) { return jacobi_jw_offset(max_sweeps) + 2 + 3 * JW_MAXP; }

// Thread 0: wait until *p >= target (acquire).  Bounded like grid_sync:
// gives up when the error flag is raised or after ~4 s (raises it: 3).
__device__ bool spin_ge(unsigned *p, unsigned target, int *err) {
  const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
  while (__hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(1);
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if (__builtin_amdgcn_s_memrealtime() - t0 > 400000000ull) {
      atomicExch(err, 3);
      return false;
    }
  }
  return true;
}
constexpr unsigned JW_DONE = 0xFFFFFFFFu;

// ||S||_F^2 of the top k x k block of Y (one workgroup, fixed summation order)
__global__ __launch_bounds__(1024) void k_fro2(int k, const double *__restrict__ Y, long ldy, double *out) {
  __shared__ double red[16];
  double sum = 0.0;
  const long total = (long)k * k;
  for (long idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const long c = idx / k, i = idx - c * k;
    const double v = Y[i + c * ldy];
    sum = __builtin_fma(v, v, sum);
  }
  sum = block_sum(sum, red);
  if (threadIdx.x == 0) out[0] = sum;
}

// X(:, pair) <- X(:, pair) * V for the 16-row bands rt = w, w + nw, ... of a
// k-row matrix (FP64 MFMA).  SRC_LDS: the operand rows come from the LDS
// copy Xl (column stride ldx) and go to the global columns of the pair in Xg;
// otherwise they are read from and written back to Xg.  Up to NBW bands are
// processed together so that 2*NBW independent accumulator chains hide the
// MFMA latency.
template <int LC, int NBW, bool SRC_LDS>
__device__ __forceinline__ void jacobi_pair_mul(int k, int w, int nw, const double *Xl, int ldx,
                                                double *__restrict__ Xg, long ldg, const double *V,
                                                int bA, int bB, int lane) {
  constexpr int B = LC / 2, NCT = LC / 16, NK = LC / 4;
  const int nband = (k + 15) >> 4;
  for (int rt0 = w; rt0 < nband; rt0 += NBW * nw) {
    double a[NBW][NK];
#pragma unroll
    for (int bb = 0; bb < NBW; ++bb) {
      const int rt = rt0 + bb * nw;
      const int i = rt * 16 + (lane & 15);
#pragma unroll
      for (int kk = 0; kk < NK; ++kk) {
        const int l = 4 * kk + (lane >> 4);
        if (SRC_LDS) {
          a[bb][kk] = (rt < nband) ? Xl[l * ldx + i] : 0.0;
        } else {
          const int gl = (l < B) ? bA * B + l : bB * B + (l - B);
          a[bb][kk] = (rt < nband && i < k) ? Xg[i + (long)gl * ldg] : 0.0;
        }
      }
    }
    v4d_j acc[NBW][NCT];
#pragma unroll
    for (int bb = 0; bb < NBW; ++bb)
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) acc[bb][ct] = v4d_j{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < NK; ++kk) {
      const int l = 4 * kk + (lane >> 4);
      double bv[NCT];
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) bv[ct] = V[(ct * 16 + (lane & 15)) * LC + l];
      // transposed product (X V)^T = V^T X^T: each lane then holds 4 columns of
      // one row, so 16 consecutive lanes store 16 consecutive rows (128 B)
#pragma unroll
      for (int bb = 0; bb < NBW; ++bb)
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
          acc[bb][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[ct], a[bb][kk], acc[bb][ct], 0, 0, 0);
    }
#pragma unroll
    for (int bb = 0; bb < NBW; ++bb) {
      const int rt = rt0 + bb * nw;
      const int row = rt * 16 + (lane & 15);
      if (rt >= nband || row >= k) continue;
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int c = ct * 16 + (lane >> 4) + 4 * v;
          const int gcol = (c < B) ? bA * B + c : bB * B + (c - B);
          Xg[row + (long)gcol * ldg] = acc[bb][ct][v];
        }
      }
    }
  }
}

// Block-Gram one-sided Jacobi.  KL: S rows are padded to 16*KL (>= k);
// LC = 2b local columns (16 or 32); NT threads, of which the first B/4 waves
// rotate (four column pairs each) and all form the Gram and the products.
// One outer step, for the block pair (bA, bB) of this workgroup:
//   1. the LC columns of S are staged in LDS and their Gram G = S_p^T S_p is
//      formed with FP64 MFMA (upper 16x16 tiles, K split over the waves,
//      four accumulator chains per wave);
//   2. the rotation rounds work on G two-sided (G <- J^T G J: a rotation
//      touches two rows and two columns of G instead of two length-k
//      columns); a 16-lane group owns one pair per round, with a column
//      phase and a row phase separated by workgroup barriers; rotations
//      accumulate in V (LC x LC).  The angle of pair (p, q) comes from G_pp,
//      G_qq, G_pq, the dot products a one-sided sweep would compute, so the
//      iteration is the one-sided Jacobi method; G is formed afresh from S
//      every step, so rounding in the two-sided updates never accumulates
//      beyond one step;
//   3. S_p <- S_p V and J_p <- J_p V (FP64 MFMA from LDS) go back to HBM for
//      the next owners of the blocks.  The J columns are loaded into
//      registers at the start of the step, so their latency hides behind the
//      Gram and the rounds, and are staged in LDS (over the Gram partials)
//      only after the rounds.
// Convergence: a sweep without rotations, or one none of whose rotations had
// cos^2 > stop_cos2 (they leave cosines of O(cos^2) behind).
template <int KL, int LC, int NT, bool SJ, bool STAMP>
__global__ __launch_bounds__(NT) void k_jacobi4(int k, int p, double *__restrict__ Sg, long lds,
                                                double *__restrict__ Jg, long ldj, double tol,
                                                double stop_cos2, int max_sweeps, int *sweeps_out,
                                                unsigned *bar, int *err, unsigned long long *dbg,
                                                double *__restrict__ Vx, const int *skip) {
  constexpr int B = LC / 2;
  constexpr int ROWS = 16 * KL;            // padded row count of S
  constexpr int LDC = ROWS + 1;            // odd column stride: conflict-free MFMA operand reads
  constexpr int LDG = LC + 1;              // odd: row and column accesses of G conflict-free
  constexpr int NW = NT / 64;
  constexpr int WR = B / 4;                // rotating waves
  constexpr int NT16 = LC / 16;
  constexpr int NTILE = NT16 * (NT16 + 1) / 2;
  constexpr int KS = NW / NTILE;           // K split of each Gram tile
  constexpr int KR = ROWS / KS;            // rows per Gram job
  constexpr int TPC = NT / LC;
  constexpr int NJ = (LDC + TPC - 1) / TPC;
  static_assert(B % 4 == 0 && WR <= NW, "block pair does not fit the workgroup");
  static_assert(NT % LC == 0 && KR % 16 == 0 && KS >= 1, "bad Jacobi tiling");
  extern __shared__ __align__(16) char smem[];  // 16-byte aligned: FP64 LDS accesses must not straddle
  double *Sl = reinterpret_cast<double *>(smem);  // LC x LDC
  double *Gl = Sl + LC * LDC;                     // LC x LDG, symmetric
  double *Vl = Gl + LC * LDG;                     // LC x LC
  double *Jl = Vl + LC * LC;                      // LC x LDC (J columns of the pair)
  double *Pt = Jl;                                // Gram partial tiles (NTILE * KS * 256), dead before Jl is written
  __shared__ int sh_steprot, sh_rot, sh_big, sh_total, sh_bigtot, sh_rot_ok;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int grp = tid >> 4, li = tid & 15;
  const int cc = tid / TPC, i0 = tid - cc * TPC;
  const int w = blockIdx.x, nblk = 2 * p;
  if (skip && *skip) return;  // the caller proved the SVD unnecessary (no truncation)
  unsigned *jbar = bar + jacobi_jw_offset(max_sweeps);
  unsigned *vflag = jbar + 2, *vack = vflag + 2 * JW_MAXP;  // vflag[parity * JW_MAXP + w]
  if (SJ && w >= p) {
    // J worker wj: applies S worker wj's step rotations V to the J columns of
    // the same block pair, one step behind, on its own CU.  V of global step g
    // arrives through Vx slot g mod 2 with the release flag 2g + (any
    // rotation) in flag slot g mod 2 (a slot is rewritten only after the J
    // worker acknowledged it); the J workers keep their own grid barrier so
    // that a step's J columns are complete before the next owners read them.
    const int wj = w - p;
    unsigned jgen = 0;
    __shared__ unsigned sh_flag;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
      bool done = false;
      for (int step = 0; step < nblk - 1; ++step) {
        const unsigned g = (unsigned)(sweep * (nblk - 1) + step + 1);
        if (tid == 0) {
          unsigned *fl = &vflag[(g & 1) * JW_MAXP + wj];
          const bool ok = spin_ge(fl, 2 * g, err);
          sh_flag = ok ? __hip_atomic_load(fl, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : JW_DONE;
        }
        __syncthreads();
        const unsigned f = sh_flag;
        if (f == JW_DONE) {
          done = true;
          break;
        }
        const double *vsrc = Vx + ((long)(g & 1) * p + wj) * (LC * LC);
        for (int idx = tid; idx < LC * LC; idx += NT) Vl[idx] = vsrc[idx];
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&vack[wj], g, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (f & 1u) {
          int bA, bB;
          circle_pair(nblk, step, wj, bA, bB);
          const int gcol_cc = (cc < B) ? bA * B + cc : bB * B + (cc - B);
          const double *jsrc = Jg + (long)gcol_cc * ldj;
          double jv[NJ];
#pragma unroll
          for (int j = 0; j < NJ; ++j) {
            const int i = i0 + TPC * j;
            jv[j] = (i < k) ? jsrc[i] : 0.0;
          }
#pragma unroll
          for (int j = 0; j < NJ; ++j)
            if (LDC % TPC == 0 || i0 + TPC * j < LDC) Jl[cc * LDC + i0 + TPC * j] = jv[j];
          __syncthreads();
          jacobi_pair_mul<LC, 2, true>(k, wv, NW, Jl, LDC, Jg, ldj, Vl, bA, bB, lane);
        }
        if (!grid_sync(jbar, p, jgen, err)) return;
      }
      if (done) break;
    }
    return;
  }
  unsigned *rotcnt = bar + 2;
  unsigned *bigcnt = bar + jacobi_mc_offset(max_sweeps);
  const double tol2 = tol * tol;
  // columns with ||s||^2 below k (2u)^2 ||S||_F^2 are numerically zero: they
  // are not rotated (their directions are rounding noise; rotating them only
  // stalls convergence on rank-deficient input)
  const double small2 = (double)k * 4.0 * DBL_EPSILON * DBL_EPSILON *
                        *reinterpret_cast<const double *>(bar + jacobi_fro_offset(max_sweeps));
  unsigned gen = 0, last_g = 0;
  bool converged = false;
  int sweep;
  for (sweep = 0; sweep < max_sweeps && !converged; ++sweep) {
    if (tid == 0) {
      sh_rot = 0;
      sh_big = 0;
    }
    for (int step = 0; step < nblk - 1; ++step) {
      int bA, bB;
      circle_pair(nblk, step, w, bA, bB);
      const int gcol_cc = (cc < B) ? bA * B + cc : bB * B + (cc - B);
      double jv[NJ];  // J columns of the pair, in flight until after the rounds
      {
        const double *src = Sg + (long)gcol_cc * lds;
        const double *jsrc = Jg + (long)gcol_cc * ldj;
        double v[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int i = i0 + TPC * j;
          v[j] = (i < k) ? src[i] : 0.0;
        }
        if (!SJ) {
#pragma unroll
          for (int j = 0; j < NJ; ++j) {
            const int i = i0 + TPC * j;
            jv[j] = (i < k) ? jsrc[i] : 0.0;
          }
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          if (LDC % TPC == 0 || i0 + TPC * j < LDC) Sl[cc * LDC + i0 + TPC * j] = v[j];
      }
      for (int idx = tid; idx < LC * LC; idx += NT) Vl[idx] = ((idx % LC) == (idx / LC)) ? 1.0 : 0.0;
      if (tid == 0) sh_steprot = 0;
      __syncthreads();
      const bool rec = STAMP && dbg && sweep == 0 && w == 0 && tid == 0;
      if (rec) dbg[step * 4 + 0] = __builtin_amdgcn_s_memrealtime();
      // 1. Gram: wave j < NTILE * KS forms rows [(j % KS) KR, +KR) of upper tile
      //    j / KS = (ti, tj) (column order); partial tiles are summed in LDS
      if (wv < NTILE * KS) {
        const int tile = wv / KS, ks = wv - tile * KS;
        int tj = 0, ti = tile;
        while (ti > tj) {
          ti -= tj + 1;
          ++tj;
        }
        v4d_j acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = v4d_j{0, 0, 0, 0};
        const double *ca = Sl + (ti * 16 + (lane & 15)) * LDC + ks * KR + (lane >> 4);
        const double *cb = Sl + (tj * 16 + (lane & 15)) * LDC + ks * KR + (lane >> 4);
#pragma unroll 4
        for (int r = 0; r < KR; r += 16) {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[r + 4 * q], cb[r + 4 * q], acc[q], 0, 0, 0);
        }
        double *pt = Pt + (tile * KS + ks) * 256;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          pt[(lane & 15) * 16 + (lane >> 4) + 4 * v] = (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]);
      }
      __syncthreads();
      for (int idx = tid; idx < LC * LC; idx += NT) {
        const int i = idx % LC, j = idx / LC;
        const int a = i <= j ? i : j, c = i <= j ? j : i;  // upper entry (a <= c)
        const int tile = (c >> 4) * ((c >> 4) + 1) / 2 + (a >> 4);
        double g = 0.0;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) g += Pt[(tile * KS + ks) * 256 + (c & 15) * 16 + (a & 15)];
        Gl[j * LDG + i] = g;
      }
      __syncthreads();
      if (rec) dbg[step * 4 + 1] = __builtin_amdgcn_s_memrealtime();
      // 2. rotation rounds on G
      const int rounds = (step == 0) ? LC - 1 : B;
      int myrot = 0, mybig = 0;
      const bool rw = wv < WR;
      for (int t = 0; t < rounds; ++t) {
        if (!rw) {  // non-rotating waves only keep the barrier count
          __syncthreads();
          __syncthreads();
          continue;
        }
        int pa, pb;
        if (step == 0) {
          circle_pair(LC, t, grp, pa, pb);
        } else {
          pa = grp;
          const int q = grp + t;
          pb = B + (q >= B ? q - B : q);
        }
        const double al = Gl[pa * LDG + pa], be = Gl[pb * LDG + pb], ga = Gl[pb * LDG + pa];
        const double g2 = ga * ga, ab = al * be;
        const bool rot = g2 > tol2 * ab && ab >= DBL_MIN && al > small2 && be > small2;
        double cs, sn, tn;
        jrot(al, be, ga, cs, sn, tn);
        cs = rot ? cs : 1.0;
        sn = rot ? sn : 0.0;
        mybig += (rot && g2 > stop_cos2 * ab) ? 1 : 0;
        myrot += rot ? 1 : 0;
        // column phase: G(:, pa), G(:, pb), V(:, pa), V(:, pb) -- all loads
        // first, so one LDS latency covers the whole update
        constexpr int NR = LC / 16;
        {
          double gp[NR], gq[NR], vp[NR], vq[NR];
#pragma unroll
          for (int u = 0; u < NR; ++u) {
            const int r = li + 16 * u;
            gp[u] = Gl[pa * LDG + r];
            gq[u] = Gl[pb * LDG + r];
            vp[u] = Vl[pa * LC + r];
            vq[u] = Vl[pb * LC + r];
          }
#pragma unroll
          for (int u = 0; u < NR; ++u) {
            const int r = li + 16 * u;
            Gl[pa * LDG + r] = cs * gp[u] - sn * gq[u];
            Gl[pb * LDG + r] = sn * gp[u] + cs * gq[u];
            Vl[pa * LC + r] = cs * vp[u] - sn * vq[u];
            Vl[pb * LC + r] = sn * vp[u] + cs * vq[u];
          }
        }
        __syncthreads();
        // row phase: G(pa, :), G(pb, :), then the exact 2x2 block
        {
          double gp[NR], gq[NR];
#pragma unroll
          for (int u = 0; u < NR; ++u) {
            const int r = li + 16 * u;
            gp[u] = Gl[r * LDG + pa];
            gq[u] = Gl[r * LDG + pb];
          }
#pragma unroll
          for (int u = 0; u < NR; ++u) {
            const int r = li + 16 * u;
            Gl[r * LDG + pa] = cs * gp[u] - sn * gq[u];
            Gl[r * LDG + pb] = sn * gp[u] + cs * gq[u];
          }
        }
        if (li == 0 && rot) {
          Gl[pa * LDG + pa] = al - tn * ga;
          Gl[pb * LDG + pb] = be + tn * ga;
          Gl[pb * LDG + pa] = 0.0;
          Gl[pa * LDG + pb] = 0.0;
        }
        __syncthreads();
      }
      if (li == 0 && myrot > 0) atomicAdd(&sh_steprot, myrot);
      if (li == 0 && mybig > 0) atomicAdd(&sh_big, mybig);
      if (!SJ) {
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          if (LDC % TPC == 0 || i0 + TPC * j < LDC) Jl[cc * LDC + i0 + TPC * j] = jv[j];
      }
      __syncthreads();
      const int rot = sh_steprot;
      if (SJ) {
        // publish V for the J worker: wait until it took the V of two steps
        // ago from this parity's slot, copy, release the flag
        const unsigned g = (unsigned)(sweep * (nblk - 1) + step + 1);
        if (tid == 0) sh_rot_ok = (g <= 2) ? 1 : (spin_ge(&vack[w], g - 2, err) ? 1 : 0);
        __syncthreads();
        if (!sh_rot_ok) return;
        double *vdst = Vx + ((long)(g & 1) * p + w) * (LC * LC);
        for (int idx = tid; idx < LC * LC; idx += NT) vdst[idx] = Vl[idx];
        wave_stores_done();  // every wave's part of V, published by thread 0's release below
        __syncthreads();
        if (tid == 0)
          __hip_atomic_store(&vflag[(g & 1) * JW_MAXP + w], 2 * g + (rot > 0 ? 1u : 0u), __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_AGENT);
        last_g = g;
      }
      if (rec) dbg[step * 4 + 2] = __builtin_amdgcn_s_memrealtime();
      // 3. S_p <- S_p V and J_p <- J_p V, straight to HBM
      if (rot > 0) {
        jacobi_pair_mul<LC, 2, true>(k, wv, NW, Sl, LDC, Sg, lds, Vl, bA, bB, lane);
        if (!SJ) jacobi_pair_mul<LC, 2, true>(k, wv, NW, Jl, LDC, Jg, ldj, Vl, bA, bB, lane);
      }
      if (tid == 0) sh_rot += rot;
      if (rec) dbg[step * 4 + 3] = __builtin_amdgcn_s_memrealtime();
      if (p > 1) {
        if (!grid_sync(bar, p, gen, err)) return;
      } else {
        __syncthreads();
      }
    }
    // sweep totals: rotations, and rotations above the stopping cosine
    __syncthreads();
    if (p > 1) {
      if (tid == 0) {
        atomicAdd(&rotcnt[sweep], (unsigned)sh_rot);
        atomicAdd(&bigcnt[sweep], (unsigned)sh_big);
      }
      if (!grid_sync(bar, p, gen, err)) return;
      if (tid == 0) {
        sh_total = (int)__hip_atomic_load(&rotcnt[sweep], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        sh_bigtot = (int)__hip_atomic_load(&bigcnt[sweep], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else if (tid == 0) {
      sh_total = sh_rot;
      sh_bigtot = sh_big;
    }
    __syncthreads();
    converged = (sh_total == 0) || (sh_bigtot == 0);
  }
  if (SJ && tid == 0) {
    // the J worker must have taken the last V before its flag slots say done
    if (last_g == 0 || spin_ge(&vack[w], last_g, err)) {
      __hip_atomic_store(&vflag[w], JW_DONE, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&vflag[JW_MAXP + w], JW_DONE, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (w == 0 && tid == 0) {
    sweeps_out[0] += sweep;
    if (!converged) {
      // the soft flag also ends the J workers' waits: raise it only once every
      // J worker has taken the last V (its remaining update has no wait)
      if (SJ && last_g > 0)
        for (int q = 0; q < p; ++q) spin_ge(&vack[q], last_g, err);
      atomicExch(err, 1);
    }
  }
}

// Y = [S'; I] with S' = S or S^T (k x k), zero-padded to kpad columns
__global__ void k_jacobi_setup(int k, int kpad, const double *__restrict__ S, int transS,
                               double *__restrict__ Y) {
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

// Local column count 2b is 16 or 32 (MFMA tiles); above k = 32 the matrix is
// split over p cooperating workgroups, with 16-column blocks pairs above
// k = 256 (the S block pair must fit in LDS).
JacobiPlan jacobi_plan(int64_t k) {
  JacobiPlan pl;
  pl.k = k;
  pl.rows = 2 * k;
  if (k <= 16) {
    pl.b = 8;
    pl.p = 1;
  } else if (k <= 256) {
    pl.b = 16;
    pl.p = (int)((k + 31) / 32);
  } else {
    pl.b = 8;
    pl.p = (int)((k + 15) / 16);
  }
  pl.kpad = 2 * (int64_t)pl.b * pl.p;
  return pl;
}
size_t jacobi_ws_bytes(const JacobiPlan &pl) { return (size_t)pl.rows * pl.kpad * sizeof(double); }

template <int KL, int LC, int NT>
static void launch_jacobi_t(hipStream_t s, const JacobiPlan &pl, double *Sg, long lds, double *Jg,
                            long ldj, double tol, double stop_cos2, int max_sweeps, int *sweeps_dev,
                            unsigned *bar, int *err, unsigned long long *dbg, double *Vx, const int *skip) {
  constexpr int NT16 = LC / 16, NTILE = NT16 * (NT16 + 1) / 2, KS = (NT / 64) / NTILE;
  static_assert((size_t)NTILE * KS * 256 <= (size_t)LC * (16 * KL + 1), "Gram partials must fit in the J area");
  const size_t lsz = (2 * (size_t)LC * (16 * KL + 1) + (size_t)LC * (LC + 1) + (size_t)LC * LC) * sizeof(double);
  TTG_CHECK(lsz <= 160 * 1024 - 256, TTG_BAD_ARG, "Jacobi block does not fit in LDS");
  TTG_CHECK(pl.k <= 16 * KL && 2 * pl.b == LC, TTG_BAD_ARG, "Jacobi plan/kernel mismatch");
  int k = (int)pl.k, p = pl.p, ms = max_sweeps;
  if (p > 1) {
    // J columns updated by p separate workgroups (split J) when V can be exchanged
    const bool sj = Vx != nullptr && p <= JW_MAXP;
    const void *fn = sj ? (dbg ? (const void *)k_jacobi4<KL, LC, NT, true, true>
                               : (const void *)k_jacobi4<KL, LC, NT, true, false>)
                        : (dbg ? (const void *)k_jacobi4<KL, LC, NT, false, true>
                               : (const void *)k_jacobi4<KL, LC, NT, false, false>);
    TTG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lsz));
    void *args[] = {&k, &p, &Sg, &lds, &Jg, &ldj, &tol, &stop_cos2, &ms, &sweeps_dev, &bar, &err, &dbg, &Vx, &skip};
    TTG_HIP(hipLaunchCooperativeKernel(fn, dim3(sj ? 2 * p : p), dim3(NT), args, lsz, s));
  } else {
    auto fn = dbg ? k_jacobi4<KL, LC, NT, false, true> : k_jacobi4<KL, LC, NT, false, false>;
    TTG_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lsz));
    hipLaunchKernelGGL(fn, dim3(1), dim3(NT), lsz, s, k, p, Sg, lds, Jg, ldj, tol,
                       stop_cos2, ms, sweeps_dev, bar, err, dbg, Vx, skip);
    TTG_HIP(hipGetLastError());
  }
}

static void launch_jacobi(hipStream_t s, const JacobiPlan &pl, double *Sg, long lds, double *Jg, long ldj,
                          double tol, double stop_cos2, int max_sweeps, int *sweeps_dev, unsigned *bar,
                          int *err, unsigned long long *dbg, double *Vx, const int *skip) {
  TTG_HIP(hipMemsetAsync(bar, 0, sizeof(unsigned) * jacobi_bar_words(max_sweeps), s));
  hipLaunchKernelGGL(k_fro2, dim3(1), dim3(1024), 0, s, (int)pl.k, Sg, lds,
                     reinterpret_cast<double *>(bar + jacobi_fro_offset(max_sweeps)));
  TTG_HIP(hipGetLastError());
  const int64_t k = pl.k;
  const int lc = 2 * pl.b;
#define TTG_JAC(KL, LC, NT) launch_jacobi_t<KL, LC, NT>(s, pl, Sg, lds, Jg, ldj, tol, stop_cos2, max_sweeps, sweeps_dev, bar, err, dbg, Vx, skip)
  if (k <= 16) {
    TTG_CHECK(lc == 16, TTG_BAD_ARG, "unsupported Jacobi block size");
    TTG_JAC(4, 16, 256);
  } else if (k <= 256) {
    TTG_CHECK(lc == 32, TTG_BAD_ARG, "unsupported Jacobi block size");
    if (k <= 64) TTG_JAC(4, 32, 512);
    else if (k <= 128) TTG_JAC(8, 32, 512);
    else TTG_JAC(16, 32, 512);
  } else {
    TTG_CHECK(lc == 16 && k <= 512, TTG_BAD_ARG, "Jacobi SVD supports k <= 512");
    TTG_JAC(32, 16, 256);
  }
#undef TTG_JAC
}

__global__ void k_polar_m(int n, double *M) {
  const long total = (long)n * n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    M[idx] = (i == j ? 1.5 : 0.0) - 0.5 * M[idx];
  }
}

size_t jacobi_polar_ws_elems(const JacobiPlan &pl) {
  return (size_t)pl.kpad * pl.kpad + (size_t)pl.rows * pl.kpad;
}

void jacobi_svd(hipStream_t s, const JacobiPlan &pl, const double *S, bool transS, double *Y,
                int *sweeps_dev, unsigned *barrier_dev, int *err_dev, int max_sweeps, double *ws,
                unsigned long long *dbg, double stop_cos2, const int *skip) {
  const int k = (int)pl.k, kpad = (int)pl.kpad;
  const long total = (long)pl.rows * kpad;
  const int blocks = (int)std::min<long>((total + 255) / 256, 2048);
  hipLaunchKernelGGL(k_jacobi_setup, dim3(blocks), dim3(256), 0, s, k, kpad, S, transS ? 1 : 0, Y);
  TTG_HIP(hipGetLastError());
  TTG_HIP(hipMemsetAsync(err_dev, 0, sizeof(int), s));
  TTG_HIP(hipMemsetAsync(sweeps_dev, 0, sizeof(int), s));
  // algorithmic flops of the small SVD (R-SVD of the k x k factor, GVL Fig. 8.6.1: 20k^3)
  ProfScope prof(KC_JACOBI, s, 20.0 * (double)k * k * k, 8.0 * 4.0 * (double)k * k);
  const double tol = sqrt((double)std::max(k, 1)) * DBL_EPSILON;
  // ws (the polar-step scratch, idle during the sweeps) carries the split-J
  // exchange of V: 2 parities x p workgroups x (2b)^2 doubles
  double *Vx = (2 * (size_t)pl.p * (2 * pl.b) * (2 * pl.b) <= jacobi_polar_ws_elems(pl)) ? ws : nullptr;
  launch_jacobi(s, pl, Y, (long)pl.rows, Y + k, (long)pl.rows, tol, stop_cos2, max_sweeps, sweeps_dev,
                barrier_dev, err_dev, dbg, Vx, skip);
  // polar correction of J (and of S J with it)
  double *M = ws, *Y2 = ws + (size_t)kpad * kpad;
  gemm<double>(s, true, false, kpad, kpad, k, 1.0, Y + k, pl.rows, Y + k, pl.rows, 0.0, M, kpad,
               nullptr, 0, 256);
  hipLaunchKernelGGL(k_polar_m, dim3((unsigned)std::min<long>(((long)kpad * kpad + 255) / 256, 1024)),
                     dim3(256), 0, s, kpad, M);
  TTG_HIP(hipGetLastError());
  gemm<double>(s, false, false, pl.rows, kpad, kpad, 1.0, Y, pl.rows, M, kpad, 0.0, Y2, pl.rows,
               nullptr, 0, 256);
  TTG_HIP(hipMemcpyAsync(Y, Y2, (size_t)pl.rows * kpad * sizeof(double), hipMemcpyDeviceToDevice, s));
}

}  // namespace ttg
