// Sums and inner products of TT tensors on the device.
//
// * a X + b Y in the rank concatenation of @tt_tensor/plus.m:41-86, with the
//   scalar factors applied to core 1 as @tt_tensor/mtimes.m:10-27 does:
//   every core assembled by one launch (plus_assemble), or, inside the
//   rounding sweep, never assembled at all (plus_round_dev in ttg_round.hip).
// * <X, Y> by the interface chain of @tt_tensor/dot.m:92-111: two GEMMs per
//   core, p_{k+1} = X_k^T (p_k Y_k), on the device.
#include "ttg_internal.h"

#include <cstring>

namespace ttg {
namespace {

// Block (x, y) covers core y with a grid stride over its elements.  Element
// (al, j, be) of core k (al fastest) is X's when al < rx0 and be < rx1, Y's
// when al >= rz0 - ry0 and be >= rz1 - ry1 (the trailing block of
// plus.m:82-83), zero otherwise; a d = 1 sum adds the two cores (plus.m:30-33).
// Core sizes are below 2^31 (checked by the caller), so the index arithmetic
// is 32-bit.
__device__ void plus_core_body(const PlusCore &c) {
  const unsigned rz0 = (unsigned)c.rz0, nn = (unsigned)c.n;
  const unsigned total = (unsigned)(c.rz0 * c.n * c.rz1);
  const unsigned oy0 = (unsigned)(c.rz0 - c.ry0), oy1 = (unsigned)(c.rz1 - c.ry1);
  const unsigned rx0 = (unsigned)c.rx0, rx1 = (unsigned)c.rx1, ry0 = (unsigned)c.ry0;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    double v;
    if (c.sum) {
      v = c.a * c.x[e] + c.b * c.y[e];
    } else {
      const unsigned al = e % rz0, col = e / rz0, j = col % nn, be = col / nn;
      if (al < rx0 && be < rx1)
        v = c.a * c.x[al + rx0 * (j + nn * be)];
      else if (al >= oy0 && be >= oy1)
        v = c.b * c.y[(al - oy0) + ry0 * (j + nn * (be - oy1))];
      else
        v = 0.0;
    }
    c.z[e] = v;
  }
}
__global__ void k_plus_assemble(const PlusCore *tab) { plus_core_body(tab[blockIdx.y]); }
// one core, its description passed by value (no table upload)
__global__ void k_plus_assemble_one(PlusCore c) { plus_core_body(c); }


}  // namespace

void plus_assemble(hipStream_t s, const PlusCore *tab_dev, int ncores, int64_t max_core) {
  if (ncores <= 0 || max_core <= 0) return;
  const int bx = (int)std::min<int64_t>((max_core + 255) / 256, 2048);
  hipLaunchKernelGGL(k_plus_assemble, dim3(bx, ncores), dim3(256), 0, s, tab_dev);
  TTG_HIP(hipGetLastError());
}

void plus_assemble_one(hipStream_t s, const PlusCore &c) {
  const int64_t sz = c.rz0 * c.n * c.rz1;
  TTG_CHECK(sz < (int64_t(1) << 31), TTG_BAD_ARG, "a core of the sum has 2^31 or more entries");
  const int bx = (int)std::min<int64_t>((sz + 255) / 256, 2048);
  hipLaunchKernelGGL(k_plus_assemble_one, dim3(bx), dim3(256), 0, s, c);
  TTG_HIP(hipGetLastError());
}

// Ranks of a X + b Y (plus.m:46-59): boundary ranks must agree, inner ranks add.
std::vector<int64_t> plus_ranks(int64_t d, const int64_t *n, const int64_t *rx, const int64_t *ry) {
  TTG_CHECK(d >= 1 && n && rx && ry, TTG_BAD_ARG, "d, n and both rank vectors are required");
  for (int64_t i = 0; i < d; ++i) TTG_CHECK(n[i] >= 1, TTG_BAD_ARG, "mode sizes must be >= 1");
  for (int64_t i = 0; i <= d; ++i) TTG_CHECK(rx[i] >= 1 && ry[i] >= 1, TTG_BAD_ARG, "ranks must be >= 1");
  TTG_CHECK(rx[0] == ry[0], TTG_BAD_ARG, "Inconsistent sizes in mode 1");
  TTG_CHECK(rx[d] == ry[d], TTG_BAD_ARG, "Inconsistent sizes in the last mode");
  std::vector<int64_t> rz(d + 1);
  for (int64_t i = 0; i <= d; ++i) rz[i] = rx[i] + ry[i];
  rz[0] = rx[0];
  rz[d] = rx[d];
  if (d == 1)
    for (int64_t i = 0; i <= 1; ++i) TTG_CHECK(rx[i] == ry[i], TTG_BAD_ARG, "d = 1 sums need equal ranks");
  return rz;
}

// The table of cores [k0, k1) of a X + b Y, written to pinned memory and
// uploaded to tab_dev on ctx's stream; returns the largest core size.
int64_t plus_table(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx, double a,
                   const int64_t *ry, const double *dy, double b, const std::vector<int64_t> &rz,
                   double *const *zdst, int64_t k0, int64_t k1, PlusCore *tab_dev) {
  std::vector<PlusCore> tab;
  int64_t ox = 0, oy = 0, mx = 0;
  for (int64_t k = 0; k < d; ++k) {
    if (k >= k0 && k < k1) {
      PlusCore c{};
      c.x = dx + ox;
      c.y = dy + oy;
      c.z = zdst[k - k0];
      c.rx0 = rx[k];
      c.rx1 = rx[k + 1];
      c.ry0 = ry[k];
      c.ry1 = ry[k + 1];
      c.rz0 = rz[k];
      c.rz1 = rz[k + 1];
      c.n = n[k];
      c.a = k == 0 ? a : 1.0;  // the scalars sit in core 1 (mtimes.m:10-27)
      c.b = k == 0 ? b : 1.0;
      c.sum = d == 1;
      const int64_t sz = rz[k] * n[k] * rz[k + 1];
      TTG_CHECK(sz < (int64_t(1) << 31), TTG_BAD_ARG, "a core of the sum has 2^31 or more entries");
      mx = std::max(mx, sz);
      tab.push_back(c);
    }
    ox += rx[k] * n[k] * rx[k + 1];
    oy += ry[k] * n[k] * ry[k + 1];
  }
  const size_t bytes = tab.size() * sizeof(PlusCore);
  void *hp = pinned_acquire(ctx, bytes);
  std::memcpy(hp, tab.data(), bytes);
  TTG_HIP(hipMemcpyAsync(tab_dev, hp, bytes, hipMemcpyHostToDevice, ctx->stream));
  pinned_release(ctx, ctx->stream);
  return mx;
}

// Z = a X + b Y (plus.m:41-86, scalars on core 1): one launch over all cores.
ttg_status plus_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx, double a,
                    const int64_t *ry, const double *dy, double b, int64_t *rz_out, double *dz, int64_t cap,
                    int64_t *len_out) {
  std::vector<int64_t> rz = plus_ranks(d, n, rx, ry);
  int64_t need = 0;
  for (int64_t k = 0; k < d; ++k) need += rz[k] * n[k] * rz[k + 1];
  if (len_out) *len_out = need;
  if (cap < need) return TTG_NEED_SPACE;
  std::vector<double *> zdst(d);
  int64_t o = 0;
  for (int64_t k = 0; k < d; ++k) {
    zdst[k] = dz + o;
    o += rz[k] * n[k] * rz[k + 1];
  }
  ctx->arena.reserve(d * sizeof(PlusCore) + 4096);
  ctx->arena.reset();
  PlusCore *tab = ctx->arena.take<PlusCore>(d);
  const int64_t mx = plus_table(ctx, d, n, rx, dx, a, ry, dy, b, rz, zdst.data(), 0, d, tab);
  plus_assemble(ctx->stream, tab, (int)d, mx);
  TTG_HIP(hipStreamSynchronize(ctx->stream));
  for (int64_t k = 0; k <= d; ++k) rz_out[k] = rz[k];
  return TTG_OK;
}

// <X, Y> (dot.m:92-111, boundary ranks 1): p_0 = 1,
// p_{k+1} (rx_{k+1} x ry_{k+1}) = X_k^T (p_k Y_k) with X_k seen as
// (rx_k n_k) x rx_{k+1} and p_k Y_k as (rx_k n_k) x ry_{k+1}.
ttg_status dot_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx,
                   const int64_t *ry, const double *dy, double *value) {
  TTG_CHECK(d >= 1 && n && rx && ry && dx && dy && value, TTG_BAD_ARG, "null argument");
  for (int64_t i = 0; i < d; ++i) TTG_CHECK(n[i] >= 1, TTG_BAD_ARG, "mode sizes must be >= 1");
  for (int64_t i = 0; i <= d; ++i) TTG_CHECK(rx[i] >= 1 && ry[i] >= 1, TTG_BAD_ARG, "ranks must be >= 1");
  TTG_CHECK(rx[0] == 1 && ry[0] == 1 && rx[d] == 1 && ry[d] == 1, TTG_BAD_ARG,
            "dot needs boundary ranks 1");
  hipStream_t st = ctx->stream;
  const int cus = ctx->num_cus;
  int64_t pmax = 1, tmax = 1, wsmax = 1;
  for (int64_t k = 0; k < d; ++k) {
    pmax = std::max(pmax, rx[k + 1] * ry[k + 1]);
    tmax = std::max(tmax, rx[k] * n[k] * ry[k + 1]);
    wsmax = std::max(wsmax, gemm_ws_elems(rx[k], n[k] * ry[k + 1], ry[k], cus));
    wsmax = std::max(wsmax, gemm_ws_elems(rx[k + 1], ry[k + 1], rx[k] * n[k], cus));
  }
  ArenaSizer z;
  z.take<double>(pmax);
  z.take<double>(pmax);
  z.take<double>(tmax);
  z.take<double>(wsmax);
  ctx->arena.reserve(z.used + 4096);
  ctx->arena.reset();
  double *p[2] = {ctx->arena.take<double>(pmax), ctx->arena.take<double>(pmax)};
  double *T = ctx->arena.take<double>(tmax);
  double *ws = ctx->arena.take<double>(wsmax);
  double *hp = static_cast<double *>(pinned_acquire(ctx, sizeof(double)));
  hp[0] = 1.0;
  TTG_HIP(hipMemcpyAsync(p[0], hp, sizeof(double), hipMemcpyHostToDevice, st));
  pinned_release(ctx, st);
  int64_t ox = 0, oy = 0;
  int cur = 0;
  for (int64_t k = 0; k < d; ++k) {
    const int64_t a0 = rx[k], a1 = rx[k + 1], b0 = ry[k], b1 = ry[k + 1], nk = n[k];
    // T (a0 x nk b1) = p (a0 x b0) * Y_k (b0 x nk b1)
    gemm<double>(st, false, false, a0, nk * b1, b0, 1.0, p[cur], a0, dy + oy, b0, 0.0, T, a0, ws, wsmax, cus);
    // p' (a1 x b1) = X_k^T ((a0 nk) x a1)^T * T ((a0 nk) x b1)
    gemm<double>(st, true, false, a1, b1, a0 * nk, 1.0, dx + ox, a0 * nk, T, a0 * nk, 0.0, p[cur ^ 1], a1, ws,
                 wsmax, cus);
    cur ^= 1;
    ox += a0 * nk * a1;
    oy += b0 * nk * b1;
  }
  TTG_HIP(hipMemcpyAsync(hp, p[cur], sizeof(double), hipMemcpyDeviceToHost, st));
  TTG_HIP(hipStreamSynchronize(st));
  *value = hp[0];
  return TTG_OK;
}

}  // namespace ttg
