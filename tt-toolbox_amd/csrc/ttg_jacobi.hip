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

// Three independent wave sums, interleaved so their DPP chains overlap.
__device__ __forceinline__ void wave_sum3(double &a, double &b, double &c) {
  a += dpp_d<0xB1>(a); b += dpp_d<0xB1>(b); c += dpp_d<0xB1>(c);
  a += dpp_d<0x4E>(a); b += dpp_d<0x4E>(b); c += dpp_d<0x4E>(c);
  a += dpp_d<0x141>(a); b += dpp_d<0x141>(b); c += dpp_d<0x141>(c);
  a += dpp_d<0x140>(a); b += dpp_d<0x140>(b); c += dpp_d<0x140>(c);
  a = (readlane_d(a, 0) + readlane_d(a, 16)) + (readlane_d(a, 32) + readlane_d(a, 48));
  b = (readlane_d(b, 0) + readlane_d(b, 16)) + (readlane_d(b, 32) + readlane_d(b, 48));
  c = (readlane_d(c, 0) + readlane_d(c, 16)) + (readlane_d(c, 32) + readlane_d(c, 48));
}

// KL: rows of S per lane (k <= 64*KL); LC = 2b local columns (multiple of 16, <= 64).
template <int KL, int LC>
__global__ __launch_bounds__(1024) void k_jacobi2(int k, int p, double *__restrict__ Sg, long lds,
                                                  double *__restrict__ Jg, long ldj, double tol,
                                                  int max_sweeps, int *sweeps_out, unsigned *bar,
                                                  int *err, unsigned long long *dbg) {
  constexpr int B = LC / 2;
  extern __shared__ char smem[];
  double *Sl = reinterpret_cast<double *>(smem);  // LC x k, column-major
  double *Vl = Sl + (size_t)LC * k;               // LC x LC, column-major
  __shared__ int sh_steprot, sh_rot, sh_total;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  const int w = blockIdx.x;
  const int nblk = 2 * p;
  const int total = LC * k;
  unsigned gen = 0;
  bool converged = false;
  int sweep = 0;
  for (sweep = 0; sweep < max_sweeps && !converged; ++sweep) {
    if (tid == 0) sh_rot = 0;
    const int outer = nblk - 1;
    for (int step = 0; step < outer; ++step) {
      int bA, bB;
      circle_pair(nblk, step, w, bA, bB);
      // load the S columns of the block pair (batched, 8 independent loads per pass)
      for (int base = 0; base < total; base += 8 * 1024) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int idx = base + tid + j * 1024;
          v[j] = 0.0;
          if (idx < total) {
            const int c = idx / k, i = idx - c * k;
            const int gcol = (c < B) ? bA * B + c : bB * B + (c - B);
            v[j] = Sg[i + (long)gcol * lds];
          }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int idx = base + tid + j * 1024;
          if (idx < total) Sl[idx] = v[j];
        }
      }
      for (int idx = tid; idx < LC * LC; idx += blockDim.x) Vl[idx] = (idx % LC == idx / LC) ? 1.0 : 0.0;
      if (tid == 0) sh_steprot = 0;
      __syncthreads();
      const bool rec = dbg && sweep == 0 && w == 0 && tid == 0;
      if (rec) dbg[step * 4 + 0] = __builtin_amdgcn_s_memrealtime();
      const int rounds = (step == 0) ? LC - 1 : B;
      for (int t = 0; t < rounds; ++t) {
        for (int pr = wv; pr < B; pr += nw) {
          int pa, pb;
          if (step == 0) {
            circle_pair(LC, t, pr, pa, pb);
          } else {
            pa = pr;
            pb = B + (pr + t) % B;
          }
          double *cp = Sl + (long)pa * k;
          double *cq = Sl + (long)pb * k;
          double x[KL], y[KL];
          double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
          for (int q = 0; q < KL; ++q) {
            const int i = lane + 64 * q;
            const int ic = i < k ? i : k - 1;  // branch-free: clamp, then zero
            const double xv = cp[ic], yv = cq[ic];
            x[q] = i < k ? xv : 0.0;
            y[q] = i < k ? yv : 0.0;
          }
#pragma unroll
          for (int q = 0; q < KL; ++q) {
            al += x[q] * x[q];
            be += y[q] * y[q];
            ga += x[q] * y[q];
          }
          wave_sum3(al, be, ga);
          if (ga * ga > tol * tol * al * be) {
            const double zeta = (be - al) / (2.0 * ga);
            double tt;
            if (fabs(zeta) > 1e150) tt = 0.5 / zeta;
            else tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            const double cs = 1.0 / sqrt(1.0 + tt * tt), sn = cs * tt;
#pragma unroll
            for (int q = 0; q < KL; ++q) {
              const int i = lane + 64 * q;
              if (i < k) {
                cp[i] = cs * x[q] - sn * y[q];
                cq[i] = sn * x[q] + cs * y[q];
              }
            }
            if (lane < LC) {
              const double vp = Vl[pa * LC + lane], vq = Vl[pb * LC + lane];
              Vl[pa * LC + lane] = cs * vp - sn * vq;
              Vl[pb * LC + lane] = sn * vp + cs * vq;
            }
            if (lane == 0) atomicAdd(&sh_steprot, 1);
          }
        }
        __syncthreads();
      }
      const int rot = sh_steprot;
      if (rec) dbg[step * 4 + 1] = __builtin_amdgcn_s_memrealtime();
      if (rot > 0) {
        for (int base = 0; base < total; base += 8 * 1024) {
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int idx = base + tid + j * 1024;
            if (idx < total) {
              const int c = idx / k, i = idx - c * k;
              const int gcol = (c < B) ? bA * B + c : bB * B + (c - B);
              Sg[i + (long)gcol * lds] = Sl[idx];
            }
          }
        }
        // J(:, pair) <- J(:, pair) * V with FP64 MFMA; each wave owns 16-row bands
        constexpr int NCT = LC / 16;
        const int nrt = (k + 15) / 16;
        for (int rt = wv; rt < nrt; rt += nw) {
          const int i = rt * 16 + (lane & 15);
          double a[LC / 4];
#pragma unroll
          for (int kk = 0; kk < LC / 4; ++kk) {
            const int l = 4 * kk + (lane >> 4);
            const int gl = (l < B) ? bA * B + l : bB * B + (l - B);
            a[kk] = (i < k) ? Jg[i + (long)gl * ldj] : 0.0;
          }
          v4d_j acc[NCT];
#pragma unroll
          for (int ct = 0; ct < NCT; ++ct) acc[ct] = v4d_j{0, 0, 0, 0};
#pragma unroll
          for (int kk = 0; kk < LC / 4; ++kk) {
            const int l = 4 * kk + (lane >> 4);
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct)
              acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], Vl[l + (ct * 16 + (lane & 15)) * LC], acc[ct], 0, 0, 0);
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
    if (p > 1) {
      if (tid == 0) atomicAdd(&bar[2 + sweep], (unsigned)sh_rot);
      if (!grid_sync(bar, p, gen, err)) return;
      if (tid == 0) sh_total = (int)__hip_atomic_load(&bar[2 + sweep], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (tid == 0) sh_total = sh_rot;
    }
    __syncthreads();
    converged = (sh_total == 0);
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

// Local column count 2b is a multiple of 16 (MFMA tiles of the J update) and
// at most 64 (four accumulator tiles per wave).
JacobiPlan jacobi_plan(int64_t k) {
  JacobiPlan pl;
  pl.k = k;
  pl.rows = 2 * k;
  if (k <= 64) {
    pl.b = (int)(8 * ((std::max<int64_t>(k, 1) + 15) / 16));
    pl.p = 1;
  } else {
    pl.b = 16;
    pl.p = (int)((k + 31) / 32);
  }
  pl.kpad = 2 * (int64_t)pl.b * pl.p;
  return pl;
}
size_t jacobi_ws_bytes(const JacobiPlan &pl) { return (size_t)pl.rows * pl.kpad * sizeof(double); }

template <int KL, int LC>
static void launch_jacobi_t(hipStream_t s, const JacobiPlan &pl, double *Sg, long lds, double *Jg,
                            long ldj, double tol, int max_sweeps, int *sweeps_dev, unsigned *bar,
                            int *err, unsigned long long *dbg) {
  const size_t lsz = ((size_t)LC * pl.k + (size_t)LC * LC) * sizeof(double);
  TTG_CHECK(lsz <= 160 * 1024, TTG_BAD_ARG, "Jacobi block does not fit in LDS");
  TTG_HIP(hipFuncSetAttribute((const void *)k_jacobi2<KL, LC>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lsz));
  int k = (int)pl.k, p = pl.p, ms = max_sweeps;
  if (p > 1) {
    void *args[] = {&k, &p, &Sg, &lds, &Jg, &ldj, &tol, &ms, &sweeps_dev, &bar, &err, &dbg};
    TTG_HIP(hipLaunchCooperativeKernel((const void *)k_jacobi2<KL, LC>, dim3(p), dim3(1024), args, lsz, s));
  } else {
    hipLaunchKernelGGL((k_jacobi2<KL, LC>), dim3(1), dim3(1024), lsz, s, k, p, Sg, lds, Jg, ldj, tol, ms,
                       sweeps_dev, bar, err, dbg);
    TTG_HIP(hipGetLastError());
  }
}

template <int KL>
static void launch_jacobi_kl(hipStream_t s, const JacobiPlan &pl, double *Sg, long lds, double *Jg,
                             long ldj, double tol, int ms, int *sw, unsigned *bar, int *err,
                             unsigned long long *dbg) {
  switch (2 * pl.b) {
    case 16: launch_jacobi_t<KL, 16>(s, pl, Sg, lds, Jg, ldj, tol, ms, sw, bar, err, dbg); break;
    case 32: launch_jacobi_t<KL, 32>(s, pl, Sg, lds, Jg, ldj, tol, ms, sw, bar, err, dbg); break;
    case 48: launch_jacobi_t<KL, 48>(s, pl, Sg, lds, Jg, ldj, tol, ms, sw, bar, err, dbg); break;
    case 64: launch_jacobi_t<KL, 64>(s, pl, Sg, lds, Jg, ldj, tol, ms, sw, bar, err, dbg); break;
    default: throw Error(TTG_BAD_ARG, "unsupported Jacobi block size");
  }
}

static void launch_jacobi(hipStream_t s, const JacobiPlan &pl, double *Sg, long lds, double *Jg, long ldj,
                          double tol, int max_sweeps, int *sweeps_dev, unsigned *bar, int *err,
                          unsigned long long *dbg) {
  TTG_HIP(hipMemsetAsync(bar, 0, sizeof(unsigned) * (2 + max_sweeps), s));
  const int64_t k = pl.k;
  if (k <= 64) launch_jacobi_kl<1>(s, pl, Sg, lds, Jg, ldj, tol, max_sweeps, sweeps_dev, bar, err, dbg);
  else if (k <= 128) launch_jacobi_kl<2>(s, pl, Sg, lds, Jg, ldj, tol, max_sweeps, sweeps_dev, bar, err, dbg);
  else if (k <= 256) launch_jacobi_kl<4>(s, pl, Sg, lds, Jg, ldj, tol, max_sweeps, sweeps_dev, bar, err, dbg);
  else if (k <= 512) launch_jacobi_kl<8>(s, pl, Sg, lds, Jg, ldj, tol, max_sweeps, sweeps_dev, bar, err, dbg);
  else throw Error(TTG_BAD_ARG, "Jacobi SVD supports k <= 512");
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
  launch_jacobi(s, pl, Y, (long)pl.rows, Y + k, (long)pl.rows, tol, max_sweeps, sweeps_dev,
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
