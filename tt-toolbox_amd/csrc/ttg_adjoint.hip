// Adjoint kernels (gradients) — TT-Toolbox has no gradients; the formulas
// come from its interface products:
//   * completion loss 0.5*||P_Omega(X) - y||^2: entry = product of r x r
//     slices (@tt_tensor/subsref.m:116-122); per-sample left/right interface
//     vectors as in tt_regression/reg_als.m:22-39,66-92;
//   * bilinear form <y, A x>: left/right interface chains of
//     core/tt_mvdot.m:16-57 (and dot.m:92-111 for two vectors).
#include "ttg_internal.h"
#include "ttg_device.h"
#include <algorithm>
#include <cstdlib>

namespace ttg {

// ---------------------------------------------------------------------------
// completion
// ---------------------------------------------------------------------------
// Group of G threads per sample (G >= max rank, power of two, <= 256).
// Lbuf[k] (S x r_k): left interface of core k; the chain is stored for all k.
// Each step stages the sample's r_k x r_{k+1} slice G_k(:, i_k, :) in LDS
// (coalesced along r_k, rows padded by one against bank conflicts) and thread
// t forms entry t of the next interface from it.
// Cores [k0, k1): the chain starts from Lin (S x r_k0; null: the scalar 1,
// k0 = 0), stores L_k at Lbuf + lofs[k] when `store`, and ends either with
// the residuals and loss partials (k1 = d) or by writing L_k1 to Lout.
template <class T>
__global__ void k_completion_left(int d, const long *__restrict__ n, const long *__restrict__ r,
                                  const long *__restrict__ cofs, const long *__restrict__ lofs,
                                  const T *__restrict__ core, long S, const int *__restrict__ idx,
                                  const T *__restrict__ vals, int G, int slice_ld, T *__restrict__ Lbuf,
                                  T *__restrict__ res, double *__restrict__ lpart, int k0, int k1, int store,
                                  const T *__restrict__ Lin, T *__restrict__ Lout) {
  extern __shared__ __align__(16) char smem[];  // 16-byte aligned: FP64 LDS accesses must not straddle
  __shared__ double red[16];
  const int groups = blockDim.x / G;
  const int gid = threadIdx.x / G, t = threadIdx.x % G;
  T *buf = reinterpret_cast<T *>(smem) + (long)gid * (2 * G + (long)slice_ld * G);  // cur, nxt, slice
  const long s = (long)blockIdx.x * groups + gid;
  T *cur = buf, *nxt = buf + G, *sl = buf + 2 * G;
  const bool active = s < S;
  if (Lin) {
    if (active && t < r[k0]) cur[t] = Lin[s * r[k0] + t];
  } else if (t == 0) {
    cur[0] = T(1);
  }
  __syncthreads();
  for (int k = k0; k < k1; ++k) {
    const long rk = r[k], rk1 = r[k + 1], nk = n[k];
    const int ik = active ? idx[s * d + k] : 0;
    if (store && active && t < rk) Lbuf[lofs[k] + s * rk + t] = cur[t];
    if (active) {
      const T *g = core + cofs[k] + rk * ik;  // G_k(a, ik, b) at g[a + rk nk b]
      for (long e = t; e < rk * rk1; e += G) {
        const long a = e % rk, b = e / rk;
        sl[a + slice_ld * b] = g[a + rk * nk * b];
      }
    }
    __syncthreads();
    if (active && t < rk1) {
      const T *col = sl + slice_ld * t;
      T acc = T(0);
      for (long a = 0; a < rk; ++a) acc += cur[a] * col[a];
      nxt[t] = acc;
    }
    __syncthreads();
    T *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  if (k1 < d) {  // uniform over the grid
    if (active && t < r[k1]) Lout[s * r[k1] + t] = cur[t];
    return;
  }
  double l = 0.0;
  if (active && t == 0) {
    const T rs = cur[0] - vals[s];
    res[s] = rs;
    l = 0.5 * (double)rs * (double)rs;
  }
  l = block_sum(l, red);
  if (threadIdx.x == 0) lpart[blockIdx.x] = l;
}

// Right interfaces R_k (S x r_{k+1}), R_{d-1} = 1.
template <class T>
__global__ void k_completion_right(int d, const long *__restrict__ n, const long *__restrict__ r,
                                   const long *__restrict__ cofs, const long *__restrict__ rofs,
                                   const T *__restrict__ core, long S, const int *__restrict__ idx,
                                   int G, T *__restrict__ Rbuf) {
  extern __shared__ __align__(16) char smem[];  // 16-byte aligned: FP64 LDS accesses must not straddle
  T *buf = reinterpret_cast<T *>(smem);
  const int groups = blockDim.x / G;
  const int gid = threadIdx.x / G, t = threadIdx.x % G;
  const long s = (long)blockIdx.x * groups + gid;
  T *cur = buf + gid * 2 * G, *nxt = cur + G;
  const bool active = s < S;
  if (t == 0) cur[0] = T(1);
  __syncthreads();
  for (int k = d - 1; k >= 0; --k) {
    const long rk = r[k], rk1 = r[k + 1], nk = n[k];
    if (active && t < rk1) Rbuf[rofs[k] + s * rk1 + t] = cur[t];
    const int ik = active ? idx[s * d + k] : 0;
    if (active && t < rk) {
      const T *g = core + cofs[k] + t + rk * ik;
      T acc = T(0);
      for (long b = 0; b < rk1; ++b) acc += g[rk * nk * b] * cur[b];
      nxt[t] = acc;
    }
    __syncthreads();
    T *tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
}

// Partial gradients of every core in one launch: block (x, k) takes samples
// [x chunk, (x + 1) chunk) for core k and accumulates
//   acc[i][a][b] += res_s L_k(s)_a R_k(s)_b,  i = i_k(s),
// in LDS, thread t owning the pairs (a, b) = t, t + blockDim, ... (no two
// threads touch one entry, and every thread adds its samples in order, so
// the result is deterministic).  The block's n_k r_k r_{k+1} sums go to
// part[gblk cofs_k + x csize_k + e], e = a + r_k (i + n_k b) (core layout).
// single != 0: one core (k0), L is the interface itself (Lbuf, no offset)
// and the partials go to part + x csize.
template <class T>
__global__ __launch_bounds__(256) void k_completion_grad_all(int d, const long *__restrict__ n,
                                                             const long *__restrict__ r,
                                                             const long *__restrict__ cofs,
                                                             const long *__restrict__ lofs,
                                                             const long *__restrict__ rofs, long S,
                                                             const int *__restrict__ idx,
                                                             const T *__restrict__ Lbuf,
                                                             const T *__restrict__ Rbuf,
                                                             const T *__restrict__ res, long chunk,
                                                             T *__restrict__ part, int k0, int single) {
  extern __shared__ __align__(16) char smem[];
  T *acc = reinterpret_cast<T *>(smem);
  const int k = blockIdx.y + k0;
  const long rk = r[k], rk1 = r[k + 1], nk = n[k];
  const long pairs = rk * rk1, csize = rk * nk * rk1;
  for (long e = threadIdx.x; e < csize; e += blockDim.x) acc[e] = T(0);
  __syncthreads();
  const long s0 = (long)blockIdx.x * chunk, s1 = min(S, s0 + chunk);
  const T *L = single ? Lbuf : Lbuf + lofs[k], *R = Rbuf + rofs[k];
  for (long s = s0; s < s1; ++s) {
    const long i = idx[s * d + k];
    const T w = res[s];
    T *ai = acc + i * pairs;
    for (long p = threadIdx.x; p < pairs; p += blockDim.x) {
      const long a = p % rk, b = p / rk;
      ai[p] += w * L[s * rk + a] * R[s * rk1 + b];
    }
  }
  __syncthreads();
  T *out = part + (single ? 0L : (long)gridDim.x * cofs[k]) + (long)blockIdx.x * csize;
  for (long e = threadIdx.x; e < csize; e += blockDim.x) {
    const long a = e % rk, rest = e / rk, i = rest % nk, b = rest / nk;
    out[e] = acc[i * pairs + a + rk * b];
  }
}

// dgrad[e] = sum over the gblk partial blocks of e's core (fixed order).
// konly >= 0: only core konly, whose partials start at part_base.
template <class T>
__global__ void k_sum_parts_all(int d, const long *__restrict__ cofs, int gblk, T *__restrict__ part_base,
                                T *__restrict__ out, int konly) {
  const long e0 = konly >= 0 ? cofs[konly] : 0, total = konly >= 0 ? cofs[konly + 1] : cofs[d];
  for (long e = e0 + blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    int k = konly >= 0 ? konly : 0;
    while (e >= cofs[k + 1]) ++k;
    const long csize = cofs[k + 1] - cofs[k], off = e - cofs[k];
    const T *p = part_base + (konly >= 0 ? 0L : (long)gblk * cofs[k]) + off;
    double acc = 0.0;
    for (int q = 0; q < gblk; ++q) acc += (double)p[(long)q * csize];
    out[e] = (T)acc;
  }
}

template <class T>
__global__ void k_fill_one(T *p, long count) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) p[i] = T(1);
}

__global__ void k_fill_zero_int(int *p, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0;
}

// flag[0] != 0 if some sample index is outside 0 <= idx[s*d+k] < n[k]
__global__ void k_check_idx(int d, const long *__restrict__ n, long S, const int *__restrict__ idx,
                            int *__restrict__ flag) {
  int bad = 0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < S * d; e += (long)gridDim.x * blockDim.x) {
    const int v = idx[e];
    bad |= (v < 0 || v >= n[e % d]);
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}


template <class T>
ttg_status adjoint_completion(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                              const T *dcore, int64_t S, const int32_t *didx, const T *dvals,
                              double *loss, T *dgrad) {
  TTG_CHECK(d >= 1 && n && r && dcore && didx && dvals && dgrad && S >= 1, TTG_BAD_ARG,
            "bad completion arguments");
  TTG_CHECK(r[0] == 1 && r[d] == 1, TTG_BAD_ARG, "completion needs boundary ranks 1");
  int64_t rmax = 1;
  for (int64_t i = 0; i <= d; ++i) rmax = std::max(rmax, r[i]);
  TTG_CHECK(rmax <= 256, TTG_BAD_ARG, "completion supports ranks up to 256");
  int G = 1;
  while (G < rmax) G <<= 1;
  const int block = std::max(G, 256);
  const int groups = block / G;
  hipStream_t st = ctx->stream;
  std::vector<int64_t> cofs(d + 1, 0), lofs(d + 1, 0), rofs(d + 1, 0);
  for (int64_t k = 0; k < d; ++k) {
    cofs[k + 1] = cofs[k] + r[k] * n[k] * r[k + 1];
    lofs[k + 1] = lofs[k] + S * r[k];
    rofs[k + 1] = rofs[k] + S * r[k + 1];
  }
  const int64_t nblk = (S + groups - 1) / groups;
  // gradient blocks: about 1024 over all cores, each at least 64 samples
  const int64_t gblk = std::max<int64_t>(1, std::min<int64_t>((1024 + d - 1) / d, (S + 63) / 64));
  const int64_t chunk = (S + gblk - 1) / gblk;
  int64_t maxcore = 1;
  for (int64_t k = 0; k < d; ++k) maxcore = std::max(maxcore, r[k] * n[k] * r[k + 1]);
  TTG_CHECK(maxcore * (int64_t)sizeof(T) <= 64 * 1024, TTG_BAD_ARG,
            "completion gradient: a core above 64 KB (r_k n_k r_{k+1}) is not supported");
  // Left interfaces of every core (S sum_k r_k values) are stored when they
  // fit the budget (TTG_ADJOINT_LBUDGET_MB, default 4096 MiB); otherwise they
  // are recomputed core by core during the gradient pass from two S x rmax
  // buffers (same arithmetic, same chunks: bitwise the same result)
  size_t budget = size_t(4096) << 20;
  if (const char *env = std::getenv("TTG_ADJOINT_LBUDGET_MB")) budget = size_t(std::max(0L, std::atol(env))) << 20;
  const bool recompute = (size_t)lofs[d] * sizeof(T) > budget;
  ArenaSizer z;
  z.take<long>(5 * (d + 1));
  if (recompute) {
    z.take<T>(S * rmax);
    z.take<T>(S * rmax);
  } else {
    z.take<T>(lofs[d]);
  }
  z.take<T>(rofs[d]);
  z.take<T>(S);
  z.take<double>(nblk);
  z.take<T>(recompute ? gblk * maxcore : gblk * cofs[d]);
  z.take<int>(64);
  ctx->arena.reserve(z.used + 4096);
  ctx->arena.reset();
  long *tab = ctx->arena.take<long>(5 * (d + 1));
  T *Lbuf = recompute ? nullptr : ctx->arena.take<T>(lofs[d]);
  T *Lc[2] = {recompute ? ctx->arena.take<T>(S * rmax) : nullptr, recompute ? ctx->arena.take<T>(S * rmax) : nullptr};
  T *Rbuf = ctx->arena.take<T>(rofs[d]);
  T *res = ctx->arena.take<T>(S);
  double *lpart = ctx->arena.take<double>(nblk);
  T *part = ctx->arena.take<T>(recompute ? gblk * maxcore : gblk * cofs[d]);
  int *dflag = ctx->arena.take<int>(64);
  std::vector<long> htab(5 * (d + 1));
  for (int64_t k = 0; k < d; ++k) {
    htab[k] = n[k];
    htab[(d + 1) + k] = r[k];
    htab[2 * (d + 1) + k] = cofs[k];
    htab[3 * (d + 1) + k] = lofs[k];
  }
  for (int64_t k = 0; k <= d; ++k) htab[4 * (d + 1) + k] = rofs[k];
  htab[(d + 1) + d] = r[d];
  htab[2 * (d + 1) + d] = cofs[d];
  htab[3 * (d + 1) + d] = lofs[d];
  TTG_HIP(hipMemcpyAsync(tab, htab.data(), htab.size() * sizeof(long), hipMemcpyHostToDevice, st));
  const long *dn = tab, *dr = tab + (d + 1), *dco = tab + 2 * (d + 1), *dlo = tab + 3 * (d + 1);
  // every sample index is checked against the mode sizes before any kernel
  // gathers core entries with it
  {
    hipLaunchKernelGGL(k_fill_zero_int, dim3(1), dim3(64), 0, st, dflag, 1);
    const unsigned cb = (unsigned)std::min<int64_t>((S * d + 255) / 256, 2048);
    hipLaunchKernelGGL(k_check_idx, dim3(cb), dim3(256), 0, st, (int)d, dn, (long)S, didx, dflag);
    TTG_HIP(hipGetLastError());
    int hflag = 0;
    TTG_HIP(hipMemcpyAsync(&hflag, dflag, sizeof(int), hipMemcpyDeviceToHost, st));
    TTG_HIP(hipStreamSynchronize(st));
    TTG_CHECK(hflag == 0, TTG_BAD_ARG, "sample index out of range (needs 0 <= idx[s*d+k] < n[k])");
  }
  const size_t shm = (size_t)groups * 2 * G * sizeof(T);
  const int slice_ld = (int)rmax + 1;
  const size_t shl = (size_t)groups * (2 * G + (size_t)slice_ld * G) * sizeof(T);
  TTG_CHECK(shl <= 128 * 1024, TTG_BAD_ARG, "completion: rank too large for the staged slices");
  TTG_HIP(hipFuncSetAttribute((const void *)k_completion_left<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)shl));
  hipLaunchKernelGGL((k_completion_left<T>), dim3((unsigned)nblk), dim3(block), shl, st, (int)d, dn,
                     dr, dco, dlo, dcore, (long)S, didx, dvals, G, slice_ld, Lbuf, res, lpart, 0, (int)d,
                     recompute ? 0 : 1, (const T *)nullptr, (T *)nullptr);
  TTG_HIP(hipGetLastError());
  const long *dro = tab + 4 * (d + 1);
  hipLaunchKernelGGL((k_completion_right<T>), dim3((unsigned)nblk), dim3(block), shm, st, (int)d,
                     dn, dr, dco, dro, dcore, (long)S, didx, G, Rbuf);
  TTG_HIP(hipGetLastError());
  const size_t sha = (size_t)maxcore * sizeof(T);
  TTG_HIP(hipFuncSetAttribute((const void *)k_completion_grad_all<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)sha));
  if (!recompute) {
    hipLaunchKernelGGL((k_completion_grad_all<T>), dim3((unsigned)gblk, (unsigned)d), dim3(256), sha, st, (int)d,
                       dn, dr, dco, dlo, dro, (long)S, didx, Lbuf, Rbuf, res, (long)chunk, part, 0, 0);
    TTG_HIP(hipGetLastError());
    const int blocks = (int)std::min<int64_t>((cofs[d] + 255) / 256, 2048);
    hipLaunchKernelGGL((k_sum_parts_all<T>), dim3(blocks), dim3(256), 0, st, (int)d, dco, (int)gblk, part, dgrad,
                       -1);
    TTG_HIP(hipGetLastError());
  } else {
    // L_0 = 1; per core: its gradient from L_k and R_k, then L_{k+1} = L_k G_k
    hipLaunchKernelGGL((k_fill_one<T>), dim3((unsigned)std::min<int64_t>((S + 255) / 256, 2048)), dim3(256), 0, st,
                       Lc[0], (long)S);
    for (int64_t k = 0; k < d; ++k) {
      T *Lk = Lc[k & 1], *Ln = Lc[(k + 1) & 1];
      hipLaunchKernelGGL((k_completion_grad_all<T>), dim3((unsigned)gblk, 1u), dim3(256), sha, st, (int)d, dn, dr,
                         dco, dlo, dro, (long)S, didx, (const T *)Lk, Rbuf, res, (long)chunk, part, (int)k, 1);
      TTG_HIP(hipGetLastError());
      const int blocks = (int)std::min<int64_t>((cofs[k + 1] - cofs[k] + 255) / 256, 2048);
      hipLaunchKernelGGL((k_sum_parts_all<T>), dim3(blocks), dim3(256), 0, st, (int)d, dco, (int)gblk, part, dgrad,
                         (int)k);
      TTG_HIP(hipGetLastError());
      if (k + 1 < d) {
        hipLaunchKernelGGL((k_completion_left<T>), dim3((unsigned)nblk), dim3(block), shl, st, (int)d, dn, dr, dco,
                           dlo, dcore, (long)S, didx, dvals, G, slice_ld, (T *)nullptr, res, lpart, (int)k,
                           (int)k + 1, 0, (const T *)Lk, Ln);
        TTG_HIP(hipGetLastError());
      }
    }
  }
  std::vector<double> hl(nblk);
  TTG_HIP(hipMemcpyAsync(hl.data(), lpart, nblk * sizeof(double), hipMemcpyDeviceToHost, st));
  TTG_HIP(hipStreamSynchronize(st));
  if (loss) {
    double s = 0.0;
    for (auto v : hl) s += v;
    *loss = s;
  }
  return TTG_OK;
}
template ttg_status adjoint_completion<double>(ttg_ctx *, int64_t, const int64_t *, const int64_t *, const double *, int64_t, const int32_t *, const double *, double *, double *);
template ttg_status adjoint_completion<float>(ttg_ctx *, int64_t, const int64_t *, const int64_t *, const float *, int64_t, const int32_t *, const float *, double *, float *);

// ---------------------------------------------------------------------------
// bilinear <y, A x>
// ---------------------------------------------------------------------------
// Core layouts: y_k (ry, n, ry'), A_k (ra, n, m, ra'), x_k (rx, m, rx').
// Interfaces: Phi_k (ry_k, ra_k, rx_k) left, Psi_k (ry_k, ra_k, rx_k) right.
struct BiShape {
  long ry0, ry1, ra0, ra1, rx0, rx1, n, m;
};

// T2[al, e, j, b] = sum_{c,a,i} Phi[c,a,al] y[c,i,e] A[a,i,j,b]
template <class T>
__global__ void k_bi_t2(BiShape s, const T *__restrict__ phi, const T *__restrict__ y,
                        const T *__restrict__ A, T *__restrict__ t2) {
  const long total = s.rx0 * s.ry1 * s.m * s.ra1;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < total; o += (long)gridDim.x * blockDim.x) {
    const long al = o % s.rx0;
    long q = o / s.rx0;
    const long e = q % s.ry1;
    q /= s.ry1;
    const long j = q % s.m, b = q / s.m;
    double acc = 0.0;
    for (long i = 0; i < s.n; ++i)
      for (long a = 0; a < s.ra0; ++a) {
        const double av = (double)A[a + s.ra0 * (i + s.n * (j + s.m * b))];
        if (av == 0.0) continue;
        double inner = 0.0;
        for (long c = 0; c < s.ry0; ++c)
          inner += (double)phi[c + s.ry0 * (a + s.ra0 * al)] * (double)y[c + s.ry0 * (i + s.n * e)];
        acc += inner * av;
      }
    t2[o] = (T)acc;
  }
}
// Phi'[e,b,f] = sum_{al,j} T2[al,e,j,b] x[al,j,f]
template <class T>
__global__ void k_bi_phi_next(BiShape s, const T *__restrict__ t2, const T *__restrict__ x,
                              T *__restrict__ phi1) {
  const long total = s.ry1 * s.ra1 * s.rx1;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < total; o += (long)gridDim.x * blockDim.x) {
    const long e = o % s.ry1;
    long q = o / s.ry1;
    const long b = q % s.ra1, f = q / s.ra1;
    double acc = 0.0;
    for (long j = 0; j < s.m; ++j)
      for (long al = 0; al < s.rx0; ++al)
        acc += (double)t2[al + s.rx0 * (e + s.ry1 * (j + s.m * b))] * (double)x[al + s.rx0 * (j + s.m * f)];
    phi1[o] = (T)acc;
  }
}
// gx[al,j,f] = sum_{e,b} T2[al,e,j,b] Psi'[e,b,f]
template <class T>
__global__ void k_bi_gx(BiShape s, const T *__restrict__ t2, const T *__restrict__ psi1,
                        T *__restrict__ gx) {
  const long total = s.rx0 * s.m * s.rx1;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < total; o += (long)gridDim.x * blockDim.x) {
    const long al = o % s.rx0;
    long q = o / s.rx0;
    const long j = q % s.m, f = q / s.m;
    double acc = 0.0;
    for (long b = 0; b < s.ra1; ++b)
      for (long e = 0; e < s.ry1; ++e)
        acc += (double)t2[al + s.rx0 * (e + s.ry1 * (j + s.m * b))] * (double)psi1[e + s.ry1 * (b + s.ra1 * f)];
    gx[o] = (T)acc;
  }
}
// V[al,j,e,b] = sum_f x[al,j,f] Psi'[e,b,f]
template <class T>
__global__ void k_bi_v(BiShape s, const T *__restrict__ x, const T *__restrict__ psi1, T *__restrict__ v) {
  const long total = s.rx0 * s.m * s.ry1 * s.ra1;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < total; o += (long)gridDim.x * blockDim.x) {
    const long al = o % s.rx0;
    long q = o / s.rx0;
    const long j = q % s.m;
    q /= s.m;
    const long e = q % s.ry1, b = q / s.ry1;
    double acc = 0.0;
    for (long f = 0; f < s.rx1; ++f)
      acc += (double)x[al + s.rx0 * (j + s.m * f)] * (double)psi1[e + s.ry1 * (b + s.ra1 * f)];
    v[o] = (T)acc;
  }
}
// W[a,al,i,e] = sum_{j,b} A[a,i,j,b] V[al,j,e,b]
template <class T>
__global__ void k_bi_w(BiShape s, const T *__restrict__ A, const T *__restrict__ v, T *__restrict__ w) {
  const long total = s.ra0 * s.rx0 * s.n * s.ry1;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < total; o += (long)gridDim.x * blockDim.x) {
    const long a = o % s.ra0;
    long q = o / s.ra0;
    const long al = q % s.rx0;
    q /= s.rx0;
    const long i = q % s.n, e = q / s.n;
    double acc = 0.0;
    for (long b = 0; b < s.ra1; ++b)
      for (long j = 0; j < s.m; ++j)
        acc += (double)A[a + s.ra0 * (i + s.n * (j + s.m * b))] *
               (double)v[al + s.rx0 * (j + s.m * (e + s.ry1 * b))];
    w[o] = (T)acc;
  }
}
// Psi[c,a,al] = sum_{i,e} y[c,i,e] W[a,al,i,e];  gy[c,i,e] = sum_{a,al} Phi[c,a,al] W[a,al,i,e]
template <class T>
__global__ void k_bi_psi(BiShape s, const T *__restrict__ y, const T *__restrict__ w, T *__restrict__ psi) {
  const long total = s.ry0 * s.ra0 * s.rx0;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < total; o += (long)gridDim.x * blockDim.x) {
    const long c = o % s.ry0;
    long q = o / s.ry0;
    const long a = q % s.ra0, al = q / s.ra0;
    double acc = 0.0;
    for (long e = 0; e < s.ry1; ++e)
      for (long i = 0; i < s.n; ++i)
        acc += (double)y[c + s.ry0 * (i + s.n * e)] * (double)w[a + s.ra0 * (al + s.rx0 * (i + s.n * e))];
    psi[o] = (T)acc;
  }
}
template <class T>
__global__ void k_bi_gy(BiShape s, const T *__restrict__ phi, const T *__restrict__ w, T *__restrict__ gy) {
  const long total = s.ry0 * s.n * s.ry1;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < total; o += (long)gridDim.x * blockDim.x) {
    const long c = o % s.ry0;
    long q = o / s.ry0;
    const long i = q % s.n, e = q / s.n;
    double acc = 0.0;
    for (long al = 0; al < s.rx0; ++al)
      for (long a = 0; a < s.ra0; ++a)
        acc += (double)phi[c + s.ry0 * (a + s.ra0 * al)] * (double)w[a + s.ra0 * (al + s.rx0 * (i + s.n * e))];
    gy[o] = (T)acc;
  }
}
template <class T> __global__ void k_fill_one(T *p) { p[0] = T(1); }

static dim3 grid_for(long total) { return dim3((unsigned)std::min<long>((total + 255) / 256, 4096)); }

template <class T>
ttg_status adjoint_bilinear(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                            const int64_t *ry, const T *dycore, const int64_t *ra,
                            const T *dacore, const int64_t *rx, const T *dxcore, double *value,
                            T *dgx, T *dgy, void *ext_ws, size_t ext_bytes) {
  // ext_ws != nullptr: the scratch comes from caller-provided device memory of
  // ext_bytes (no allocation); otherwise from the context's workspace
  TTG_CHECK(d >= 1 && n && m && ry && ra && rx && dycore && dacore && dxcore, TTG_BAD_ARG,
            "bad bilinear arguments");
  TTG_CHECK(ry[0] == 1 && ra[0] == 1 && rx[0] == 1 && ry[d] == 1 && ra[d] == 1 && rx[d] == 1,
            TTG_BAD_ARG, "bilinear form needs boundary ranks 1");
  hipStream_t st = ctx->stream;
  std::vector<int64_t> oy(d + 1, 0), oa(d + 1, 0), ox(d + 1, 0), oi(d + 2, 0);
  int64_t tmax = 1;
  for (int64_t k = 0; k < d; ++k) {
    oy[k + 1] = oy[k] + ry[k] * n[k] * ry[k + 1];
    oa[k + 1] = oa[k] + ra[k] * n[k] * m[k] * ra[k + 1];
    ox[k + 1] = ox[k] + rx[k] * m[k] * rx[k + 1];
    tmax = std::max(tmax, rx[k] * ry[k + 1] * m[k] * ra[k + 1]);
    tmax = std::max(tmax, rx[k] * m[k] * ry[k + 1] * ra[k + 1]);
    tmax = std::max(tmax, ra[k] * rx[k] * n[k] * ry[k + 1]);
  }
  for (int64_t k = 0; k <= d; ++k) oi[k + 1] = oi[k] + ry[k] * ra[k] * rx[k];
  ArenaSizer z;
  z.take<T>(oi[d + 1]);
  z.take<T>(oi[d + 1]);
  z.take<T>(tmax);
  z.take<T>(tmax);
  T *phi, *psi, *t2, *w;
  if (ext_ws) {
    TTG_CHECK(z.used <= ext_bytes, TTG_BAD_ARG,
              "ranks too large for the no-allocation scratch (" + std::to_string(z.used) + " bytes needed, " +
                  std::to_string(ext_bytes) + " available)");
    ArenaView v{static_cast<char *>(ext_ws), ext_bytes, 0};
    phi = v.take<T>(oi[d + 1]);
    psi = v.take<T>(oi[d + 1]);
    t2 = v.take<T>(tmax);
    w = v.take<T>(tmax);
  } else {
    ctx->arena.reserve(z.used + 4096);
    ctx->arena.reset();
    phi = ctx->arena.take<T>(oi[d + 1]);
    psi = ctx->arena.take<T>(oi[d + 1]);
    t2 = ctx->arena.take<T>(tmax);
    w = ctx->arena.take<T>(tmax);
  }
  auto shape = [&](int64_t k) {
    BiShape s;
    s.ry0 = ry[k]; s.ry1 = ry[k + 1]; s.ra0 = ra[k]; s.ra1 = ra[k + 1];
    s.rx0 = rx[k]; s.rx1 = rx[k + 1]; s.n = n[k]; s.m = m[k];
    return s;
  };
  hipLaunchKernelGGL((k_fill_one<T>), dim3(1), dim3(1), 0, st, phi);
  hipLaunchKernelGGL((k_fill_one<T>), dim3(1), dim3(1), 0, st, psi + oi[d]);
  for (int64_t k = 0; k < d; ++k) {  // left chain (tt_mvdot.m:16-57)
    BiShape s = shape(k);
    hipLaunchKernelGGL((k_bi_t2<T>), grid_for(s.rx0 * s.ry1 * s.m * s.ra1), dim3(256), 0, st, s,
                       phi + oi[k], dycore + oy[k], dacore + oa[k], t2);
    hipLaunchKernelGGL((k_bi_phi_next<T>), grid_for(s.ry1 * s.ra1 * s.rx1), dim3(256), 0, st, s, t2,
                       dxcore + ox[k], phi + oi[k + 1]);
  }
  for (int64_t k = d - 1; k >= 0; --k) {  // right chain
    BiShape s = shape(k);
    hipLaunchKernelGGL((k_bi_v<T>), grid_for(s.rx0 * s.m * s.ry1 * s.ra1), dim3(256), 0, st, s,
                       dxcore + ox[k], psi + oi[k + 1], t2);
    hipLaunchKernelGGL((k_bi_w<T>), grid_for(s.ra0 * s.rx0 * s.n * s.ry1), dim3(256), 0, st, s,
                       dacore + oa[k], t2, w);
    hipLaunchKernelGGL((k_bi_psi<T>), grid_for(s.ry0 * s.ra0 * s.rx0), dim3(256), 0, st, s,
                       dycore + oy[k], w, psi + oi[k]);
    if (dgy)
      hipLaunchKernelGGL((k_bi_gy<T>), grid_for(s.ry0 * s.n * s.ry1), dim3(256), 0, st, s,
                         phi + oi[k], w, dgy + oy[k]);
  }
  if (dgx) {
    for (int64_t k = 0; k < d; ++k) {
      BiShape s = shape(k);
      hipLaunchKernelGGL((k_bi_t2<T>), grid_for(s.rx0 * s.ry1 * s.m * s.ra1), dim3(256), 0, st, s,
                         phi + oi[k], dycore + oy[k], dacore + oa[k], t2);
      hipLaunchKernelGGL((k_bi_gx<T>), grid_for(s.rx0 * s.m * s.rx1), dim3(256), 0, st, s, t2,
                         psi + oi[k + 1], dgx + ox[k]);
    }
  }
  TTG_HIP(hipGetLastError());
  T hv = T(0);
  TTG_HIP(hipMemcpyAsync(&hv, phi + oi[d], sizeof(T), hipMemcpyDeviceToHost, st));
  TTG_HIP(hipStreamSynchronize(st));
  if (value) *value = (double)hv;
  return TTG_OK;
}
template ttg_status adjoint_bilinear<double>(ttg_ctx *, int64_t, const int64_t *, const int64_t *, const int64_t *, const double *, const int64_t *, const double *, const int64_t *, const double *, double *, double *, double *, void *, size_t);
template ttg_status adjoint_bilinear<float>(ttg_ctx *, int64_t, const int64_t *, const int64_t *, const int64_t *, const float *, const int64_t *, const float *, const int64_t *, const float *, double *, float *, float *, void *, size_t);

// ---------------------------------------------------------------------------
// batched rounding: per-tensor sweep on the shared stream (first version)
// ---------------------------------------------------------------------------
ttg_status round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                     const double *dcore, double eps, const int64_t *rmax, int prec,
                     int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                     ttg_info *info);
ttg_status round_capacity(int64_t d, const int64_t *n, const int64_t *r, int64_t *cap);

}  // namespace ttg
