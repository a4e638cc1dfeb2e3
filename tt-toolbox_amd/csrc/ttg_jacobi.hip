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

// Software grid barrier for a cooperative launch: bar[0] arrivals, bar[1]
// generation.  Bounded spin: after ~4 s a timeout flag is raised and every
// workgroup leaves the kernel at its next check.
__device__ bool grid_sync(unsigned *bar, unsigned nblocks, unsigned &gen, int *err) {
  __shared__ int sh_ok;
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

// Sums of three values over each 16-lane row; afterwards every lane of the
// row holds the bitwise-identical row total (each step adds a lane to its
// mirror partner, and FP addition is commutative).
__device__ __forceinline__ void row16_sum3(double &a, double &b, double &c) {
  a += dpp_d<0xB1>(a); b += dpp_d<0xB1>(b); c += dpp_d<0xB1>(c);     // quad_perm [1,0,3,2]
  a += dpp_d<0x4E>(a); b += dpp_d<0x4E>(b); c += dpp_d<0x4E>(c);     // quad_perm [2,3,0,1]
  a += dpp_d<0x141>(a); b += dpp_d<0x141>(b); c += dpp_d<0x141>(c);  // row_half_mirror
  a += dpp_d<0x140>(a); b += dpp_d<0x140>(b); c += dpp_d<0x140>(c);  // row_mirror
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
__device__ __forceinline__ void jrot(double al, double be, double ga, double &cs, double &sn) {
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
}

// Workgroup barrier that orders LDS only: waits for this wave's LDS traffic
// (lgkmcnt = 0) but not for its outstanding global loads, so a wave can keep
// loads in flight across it.  Encoding of s_waitcnt on gfx9: vmcnt = 63,
// expcnt = 7, lgkmcnt = 0.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Barrier among the rotating waves only, through an LDS counter (the helper
// waves of the workgroup never arrive here).  All rotating waves execute the
// same number of rounds, so the count is always reached; the spin cap only
// turns a logic error into a reported failure instead of a hang.
__device__ __forceinline__ void rot_sync(unsigned *cnt, unsigned target, int *err) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if ((threadIdx.x & 63) == 0) atomicAdd(cnt, 1u);
  unsigned spins = 0;
  while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target) {
    if (++spins > (1u << 24)) {
      atomicExch(err, 3);
      break;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Layout of the barrier/statistics buffer: [0] arrivals, [1] generation,
// [2, 2+ms) rotations per sweep, then per sweep the number of rotations whose
// cos^2 exceeded the stopping threshold.
static __host__ __device__ inline int jacobi_mc_offset(int max_sweeps) { return (2 + max_sweeps + 1) & ~1; }
int jacobi_bar_words(int max_sweeps) { return jacobi_mc_offset(max_sweeps) + 2 * max_sweeps; }

// J(:, pair) <- J(:, pair) * V for the 16-row band rt (FP64 MFMA, one wave),
// split into the load of the band's J entries and the multiply-store.
template <int LC>
__device__ __forceinline__ void jacobi_j_band_load(int rt, int k, const double *__restrict__ Jg, long ldj,
                                                   int bA, int bB, int lane, double *a) {
  constexpr int B = LC / 2;
  const int i = rt * 16 + (lane & 15);
#pragma unroll
  for (int kk = 0; kk < LC / 4; ++kk) {
    const int l = 4 * kk + (lane >> 4);
    const int gl = (l < B) ? bA * B + l : bB * B + (l - B);
    a[kk] = (i < k) ? Jg[i + (long)gl * ldj] : 0.0;
  }
}
template <int LC>
__device__ __forceinline__ void jacobi_j_band_mul(int rt, int k, double *__restrict__ Jg, long ldj,
                                                  const double *V, int bA, int bB, int lane,
                                                  const double *a) {
  constexpr int B = LC / 2, NCT = LC / 16;
  v4d_j acc[NCT];
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) acc[ct] = v4d_j{0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < LC / 4; ++kk) {
    const int l = 4 * kk + (lane >> 4);
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct)
      acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], V[l + (ct * 16 + (lane & 15)) * LC], acc[ct], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);  // keep the V loads from all being hoisted (register pressure)
  }
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = rt * 16 + (lane >> 4) + 4 * v;
      const int c = ct * 16 + (lane & 15);
      const int gcol = (c < B) ? bA * B + c : bB * B + (c - B);
      if (row < k) Jg[row + (long)gcol * ldj] = acc[ct][v];
    }
  }
}
template <int LC>
__device__ __forceinline__ void jacobi_j_band(int rt, int k, double *__restrict__ Jg, long ldj,
                                              const double *V, int bA, int bB, int lane) {
  double a[LC / 4];
  jacobi_j_band_load<LC>(rt, k, Jg, ldj, bA, bB, lane, a);
  jacobi_j_band_mul<LC>(rt, k, Jg, ldj, V, bA, bB, lane, a);
}

// KL: rows per lane of a 16-lane group (k <= 16*KL); LC = 2b local columns
// (multiple of 16, <= 64); NT threads.  Each 16-lane group owns one column
// pair per round, so a wave rotates four pairs at once and the dot products
// need only row-local DPP sums.  Waves not needed for the B = LC/2 pairs of a
// round apply the previous step's accumulated rotations to J meanwhile (the
// J columns of that pair are not touched by anyone else until the next grid
// barrier, which orders them before their next owner's update).
// Convergence: a sweep without rotations, or a sweep none of whose rotations
// had cos^2 > stop_cos2 (they leave cosines of O(cos^2) behind).
template <int KL, int LC, int NT>
__global__ __launch_bounds__(NT) void k_jacobi3(int k, int p, double *__restrict__ Sg, long lds,
                                                double *__restrict__ Jg, long ldj, double tol,
                                                double stop_cos2, int max_sweeps, int *sweeps_out,
                                                unsigned *bar, int *err, unsigned long long *dbg) {
  constexpr int B = LC / 2;
  constexpr int LDC = 16 * KL + 16;  // column stride: 128 B mod 256 B, groups alternate bank halves
  constexpr int NW = NT / 64;
  constexpr int WR = B / 4;
  constexpr int NH = NW - WR;
  static_assert(B % 4 == 0 && WR < NW, "block pair does not fit the workgroup");
  constexpr int TPC = NT / LC;                 // threads per column for loads and stores
  constexpr int NJ = (LDC + TPC - 1) / TPC;    // rows per thread
  static_assert(NT % LC == 0, "threads must divide evenly over the columns");
  extern __shared__ __align__(16) char smem[];  // 16-byte aligned: FP64 LDS accesses must not straddle
  double *Sl = reinterpret_cast<double *>(smem);  // LC x LDC
  double *Vl = Sl + LC * LDC;                     // LC x LC, column-major (this step)
  double *Vp = Vl + LC * LC;                      // LC x LC (previous step, pending for J)
  __shared__ int sh_steprot, sh_rot;
  __shared__ unsigned sh_rsync;
  __shared__ int sh_big, sh_total, sh_bigtot;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int grp = tid >> 4, li = tid & 15;
  const int cc = tid / TPC, i0 = tid - cc * TPC;
  const int w = blockIdx.x, nblk = 2 * p;
  const int nband = (k + 15) >> 4;
  unsigned *rotcnt = bar + 2;
  unsigned *bigcnt = bar + jacobi_mc_offset(max_sweeps);
  const double tol2 = tol * tol;
  unsigned gen = 0;
  bool converged = false;
  int pend = 0, pA = 0, pB = 0;
  int sweep;
  for (sweep = 0; sweep < max_sweeps && !converged; ++sweep) {
    if (tid == 0) {
      sh_rot = 0;
      sh_big = 0;
    }
    int mybig = 0;  // rotations with cos^2 > stop_cos2 (this lane's groups)
    for (int step = 0; step < nblk - 1; ++step) {
      int bA, bB;
      circle_pair(nblk, step, w, bA, bB);
      {
        // thread -> (column c, rows i0 + TPC*j): invariants stay few, offsets immediate
        const int gcol = (cc < B) ? bA * B + cc : bB * B + (cc - B);
        const double *src = Sg + (long)gcol * lds;
        double v[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int i = i0 + TPC * j;
          v[j] = (i < k) ? src[i] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          if (LDC % TPC == 0 || i0 + TPC * j < LDC) Sl[cc * LDC + i0 + TPC * j] = v[j];
      }
      for (int idx = tid; idx < LC * LC; idx += NT) Vl[idx] = ((idx % LC) == (idx / LC)) ? 1.0 : 0.0;
      if (tid == 0) {
        sh_steprot = 0;
        sh_rsync = 0u;
      }
      __syncthreads();
      const bool rec = dbg && sweep == 0 && w == 0 && tid == 0;
      if (rec) dbg[step * 4 + 0] = __builtin_amdgcn_s_memrealtime();
      const int rounds = (step == 0) ? LC - 1 : B;
      int myrot = 0;
      if (wv < WR) {
        // Rotating waves, synchronised per round through an LDS counter so the
        // helper waves never hold them up.  After the first step of a sweep the
        // pairs are (g, B + (g + t) mod B): column g stays in registers for the
        // whole step and only the partner column moves through LDS.
        constexpr int NV = LC / 16;
        const bool xs = step > 0;
        double x[KL], vx[NV];
        if (xs) {
#pragma unroll
          for (int q = 0; q < KL; ++q) x[q] = Sl[grp * LDC + li + 16 * q];
#pragma unroll
          for (int r = 0; r < NV; ++r) vx[r] = Vl[grp * LC + li + 16 * r];
        }
        for (int t = 0; t < rounds; ++t) {
          int pa, pb;
          if (!xs) {
            circle_pair(LC, t, grp, pa, pb);
          } else {
            pa = grp;
            const int q = grp + t;
            pb = B + (q >= B ? q - B : q);
          }
          double *cp = Sl + pa * LDC + li;
          double *cq = Sl + pb * LDC + li;
          double y[KL], vy[NV];
          if (!xs) {
#pragma unroll
            for (int q = 0; q < KL; ++q) x[q] = cp[16 * q];
#pragma unroll
            for (int r = 0; r < NV; ++r) vx[r] = Vl[pa * LC + li + 16 * r];
          }
#pragma unroll
          for (int q = 0; q < KL; ++q) y[q] = cq[16 * q];
#pragma unroll
          for (int r = 0; r < NV; ++r) vy[r] = Vl[pb * LC + li + 16 * r];
          double al0 = 0.0, al1 = 0.0, be0 = 0.0, be1 = 0.0, ga0 = 0.0, ga1 = 0.0;
#pragma unroll
          for (int q = 0; q < KL; q += 2) {
            al0 = __builtin_fma(x[q], x[q], al0);
            be0 = __builtin_fma(y[q], y[q], be0);
            ga0 = __builtin_fma(x[q], y[q], ga0);
            al1 = __builtin_fma(x[q + 1], x[q + 1], al1);
            be1 = __builtin_fma(y[q + 1], y[q + 1], be1);
            ga1 = __builtin_fma(x[q + 1], y[q + 1], ga1);
          }
          double al = al0 + al1, be = be0 + be1, ga = ga0 + ga1;
          row16_sum3(al, be, ga);
          const double g2 = ga * ga, ab = al * be;
          const bool rot = g2 > tol2 * ab && ab >= DBL_MIN;
          // branch-free: a pair that does not rotate gets c = 1, s = 0 (exact identity)
          double cs, sn;
          jrot(al, be, ga, cs, sn);
          cs = rot ? cs : 1.0;
          sn = rot ? sn : 0.0;
          mybig += (rot && g2 > stop_cos2 * ab) ? 1 : 0;
          myrot += rot ? 1 : 0;
#pragma unroll
          for (int q = 0; q < KL; ++q) {
            const double xn = cs * x[q] - sn * y[q];
            y[q] = sn * x[q] + cs * y[q];
            x[q] = xn;
            cq[16 * q] = y[q];
          }
#pragma unroll
          for (int r = 0; r < NV; ++r) {
            const double vn = cs * vx[r] - sn * vy[r];
            vy[r] = sn * vx[r] + cs * vy[r];
            vx[r] = vn;
            Vl[pb * LC + li + 16 * r] = vy[r];
          }
          if (!xs) {
#pragma unroll
            for (int q = 0; q < KL; ++q) cp[16 * q] = x[q];
#pragma unroll
            for (int r = 0; r < NV; ++r) Vl[pa * LC + li + 16 * r] = vx[r];
          }
          rot_sync(&sh_rsync, (unsigned)(WR * (t + 1)), err);
        }
        if (xs) {
#pragma unroll
          for (int q = 0; q < KL; ++q) Sl[grp * LDC + li + 16 * q] = x[q];
#pragma unroll
          for (int r = 0; r < NV; ++r) Vl[grp * LC + li + 16 * r] = vx[r];
        }
      } else if (pend) {
        // helper waves: the previous step's rotations applied to J
        for (int rt = wv - WR; rt < nband; rt += NH) jacobi_j_band<LC>(rt, k, Jg, ldj, Vp, pA, pB, lane);
      }
      if (li == 0 && myrot > 0) atomicAdd(&sh_steprot, myrot);
      if (li == 0 && mybig > 0) atomicAdd(&sh_big, mybig);
      mybig = 0;
      __syncthreads();
      const int rot = sh_steprot;
      if (rec) dbg[step * 4 + 1] = __builtin_amdgcn_s_memrealtime();
      if (rot > 0) {
        {
          const int gcol = (cc < B) ? bA * B + cc : bB * B + (cc - B);
          double *dst = Sg + (long)gcol * lds;
#pragma unroll
          for (int j = 0; j < NJ; ++j) {
            const int i = i0 + TPC * j;
            if (i < k) dst[i] = Sl[cc * LDC + i];
          }
        }
        for (int idx = tid; idx < LC * LC; idx += NT) Vp[idx] = Vl[idx];
        pend = 1;
        pA = bA;
        pB = bB;
      } else {
        pend = 0;
      }
      if (tid == 0) sh_rot += rot;
      if (rec) dbg[step * 4 + 2] = __builtin_amdgcn_s_memrealtime();
      if (p > 1) {
        if (!grid_sync(bar, p, gen, err)) return;
      } else {
        __syncthreads();
      }
      if (rec) dbg[step * 4 + 3] = __builtin_amdgcn_s_memrealtime();
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
  if (pend) {
    for (int rt = wv; rt < nband; rt += NW) jacobi_j_band<LC>(rt, k, Jg, ldj, Vp, pA, pB, lane);
  }
  if (w == 0 && tid == 0) {
    sweeps_out[0] += sweep;
    if (!converged) atomicExch(err, 1);
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

// Local column count 2b is 16 or 32 (MFMA tiles of the J update); above k = 32
// the matrix is split over p = ceil(k/32) cooperating workgroups.
JacobiPlan jacobi_plan(int64_t k) {
  JacobiPlan pl;
  pl.k = k;
  pl.rows = 2 * k;
  if (k <= 32) {
    pl.b = k <= 16 ? 8 : 16;
    pl.p = 1;
  } else {
    pl.b = 16;
    pl.p = (int)((k + 31) / 32);
  }
  pl.kpad = 2 * (int64_t)pl.b * pl.p;
  return pl;
}
size_t jacobi_ws_bytes(const JacobiPlan &pl) { return (size_t)pl.rows * pl.kpad * sizeof(double); }

template <int KL, int LC, int NT>
static void launch_jacobi_t(hipStream_t s, const JacobiPlan &pl, double *Sg, long lds, double *Jg,
                            long ldj, double tol, double stop_cos2, int max_sweeps, int *sweeps_dev,
                            unsigned *bar, int *err, unsigned long long *dbg) {
  const size_t lsz = ((size_t)LC * (16 * KL + 16) + 2 * (size_t)LC * LC) * sizeof(double);
  TTG_CHECK(lsz <= 160 * 1024 - 256, TTG_BAD_ARG, "Jacobi block does not fit in LDS");
  TTG_CHECK(pl.k <= 16 * KL && 2 * pl.b == LC, TTG_BAD_ARG, "Jacobi plan/kernel mismatch");
  TTG_HIP(hipFuncSetAttribute((const void *)k_jacobi3<KL, LC, NT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lsz));
  int k = (int)pl.k, p = pl.p, ms = max_sweeps;
  if (p > 1) {
    void *args[] = {&k, &p, &Sg, &lds, &Jg, &ldj, &tol, &stop_cos2, &ms, &sweeps_dev, &bar, &err, &dbg};
    TTG_HIP(hipLaunchCooperativeKernel((const void *)k_jacobi3<KL, LC, NT>, dim3(p), dim3(NT), args, lsz, s));
  } else {
    hipLaunchKernelGGL((k_jacobi3<KL, LC, NT>), dim3(1), dim3(NT), lsz, s, k, p, Sg, lds, Jg, ldj, tol,
                       stop_cos2, ms, sweeps_dev, bar, err, dbg);
    TTG_HIP(hipGetLastError());
  }
}

static void launch_jacobi(hipStream_t s, const JacobiPlan &pl, double *Sg, long lds, double *Jg, long ldj,
                          double tol, double stop_cos2, int max_sweeps, int *sweeps_dev, unsigned *bar,
                          int *err, unsigned long long *dbg) {
  TTG_HIP(hipMemsetAsync(bar, 0, sizeof(unsigned) * jacobi_bar_words(max_sweeps), s));
  const int64_t k = pl.k;
  const int lc = 2 * pl.b;
#define TTG_JAC(KL, LC, NT) launch_jacobi_t<KL, LC, NT>(s, pl, Sg, lds, Jg, ldj, tol, stop_cos2, max_sweeps, sweeps_dev, bar, err, dbg)
  if (k <= 32) {
    switch (lc) {
      case 16: TTG_JAC(4, 16, 512); break;
      case 32: TTG_JAC(4, 32, 512); break;
      default: throw Error(TTG_BAD_ARG, "unsupported Jacobi block size");
    }
  } else {
    TTG_CHECK(lc == 32, TTG_BAD_ARG, "unsupported Jacobi block size");
    if (k <= 64) TTG_JAC(4, 32, 512);
    else if (k <= 128) TTG_JAC(8, 32, 512);
    else if (k <= 256) TTG_JAC(16, 32, 512);
    else if (k <= 512) TTG_JAC(32, 32, 512);
    else throw Error(TTG_BAD_ARG, "Jacobi SVD supports k <= 512");
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
                unsigned long long *dbg) {
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
  // stop after a sweep whose largest pair cosine is <= 1e-8: its rotations
  // leave cosines of O(1e-16) behind
  const double stop_cos2 = 1e-16;
  launch_jacobi(s, pl, Y, (long)pl.rows, Y + k, (long)pl.rows, tol, stop_cos2, max_sweeps, sweeps_dev,
                barrier_dev, err_dev, dbg);
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
