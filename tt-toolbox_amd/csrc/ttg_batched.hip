// Batched rounding of many independent TT tensors (uneven d, n, r).
//
// Two executions, chosen per context by a tuner:
//   * persistent launch (FP64): one launch of k_round_persist in which every
//     workgroup rounds whole tensors taken longest-first from a device work
//     counter (greedy LPT; the cost estimate is the QR work sum_i m_i k_i^2
//     at the QR-capped ranks).  The launch (counter reset + kernel) is
//     captured in a hipGraph and replayed while its arguments are unchanged.
//     Launch shapes: 256 threads with 128.5 KB of LDS for the Jacobi
//     matrix (one workgroup per CU), 256 threads with 48 KB (two per CU), or
//     512 threads with 128.5 KB;
//   * W concurrent chains of the single-tensor rounder: W worker contexts,
//     each with its own HIP stream and device workspace, driven by W host
//     threads over the same LPT order (any precision, any size).
// The first batch of at least 64 tensors on a context is tuned on
// interleaved slices of the LPT order: each persistent shape rounds an eighth
// of the batch, 16 stream workers a thirty-second, and the rest runs with the
// candidate of least time per unit of estimated work, which is kept for
// later batches (smaller untuned batches use the 512-thread shape).  The
// environment variable TTG_BATCH_MODE (1-3: persistent shapes, 4: stream
// workers) fixes the choice instead.
#include "ttg_internal.h"
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <numeric>
#include <thread>

namespace ttg {

ttg_status round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                     const double *dcore, double eps, const int64_t *rmax, int prec,
                     int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                     ttg_info *info);
ttg_status round_capacity(int64_t d, const int64_t *n, const int64_t *r, int64_t *cap);

static ttg_ctx *make_worker(const ttg_ctx *parent) {
  ttg_ctx *w = new ttg_ctx();
  w->device = parent->device;
  w->num_cus = parent->num_cus;
  w->worker = true;
  if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
    delete w;
    throw Error(TTG_HIP_ERROR, "cannot create a worker stream");
  }
  w->own_stream = true;
  return w;
}

// Releases every HIP object a context created, in dependency order: the work
// on its streams is finished first, then worker contexts (recursively), then
// the captured graph (before the stream it was captured on), the side stream
// and its events, profiler events, pinned host memory and the owned stream.
// The device arenas are freed by their destructors when the context is
// deleted.  Nothing is left for the runtime's exit-time teardown.
void release_ctx(ttg_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  destroy_workers(ctx);
  if (ctx->pgraph) (void)hipGraphExecDestroy(ctx->pgraph);
  ctx->pgraph = nullptr;
  ctx->pgraph_key.clear();
  gram_graph_release(ctx);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  ctx->side = nullptr;
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  ctx->ev_fork = ctx->ev_join = nullptr;
  for (auto e : ctx->prof.pool) (void)hipEventDestroy(e);
  ctx->prof.pool.clear();
  ctx->prof.used = 0;
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  ctx->host_pinned = nullptr;
  ctx->host_pinned_cap = 0;
  if (ctx->pinned_ev) (void)hipEventDestroy(ctx->pinned_ev);
  ctx->pinned_ev = nullptr;
  ctx->pinned_pending = false;
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  ctx->stream = nullptr;
}

void *pinned_acquire(ttg_ctx *ctx, size_t bytes) {
  if (ctx->pinned_pending) {
    TTG_HIP(hipEventSynchronize(ctx->pinned_ev));
    ctx->pinned_pending = false;
  }
  if (!ctx->host_pinned || ctx->host_pinned_cap < bytes) {
    if (ctx->host_pinned) TTG_HIP(hipHostFree(ctx->host_pinned));
    ctx->host_pinned = nullptr;
    const size_t cap = std::max<size_t>(bytes, 4096);
    TTG_HIP(hipHostMalloc(&ctx->host_pinned, cap, hipHostMallocDefault));
    ctx->host_pinned_cap = cap;
  }
  return ctx->host_pinned;
}

void pinned_release(ttg_ctx *ctx, hipStream_t s) {
  if (!ctx->pinned_ev) TTG_HIP(hipEventCreateWithFlags(&ctx->pinned_ev, hipEventDisableTiming));
  TTG_HIP(hipEventRecord(ctx->pinned_ev, s));
  ctx->pinned_pending = true;
}

void destroy_workers(ttg_ctx *ctx) {
  for (ttg_ctx *w : ctx->workers) {
    release_ctx(w);
    delete w;
  }
  ctx->workers.clear();
}

struct BatchArgs {
  const int64_t *d, *nofs, *n, *rofs, *r, *cofs, *cofs_out;
  const double *dcore;
  double *dcore_out;
  int64_t *r_out;
  double eps;
  int prec;
  const int64_t *rmax = nullptr;  // null, or packed like r
};

// Rounds tensors order[begin, end) on `nw` workers; per-tensor status and info.
static void run_batch(ttg_ctx *ctx, const BatchArgs &a, const std::vector<int64_t> &order,
                      int64_t begin, int64_t end, int nw, std::vector<ttg_status> &st,
                      std::vector<ttg_info> &inf, std::vector<std::string> &err) {
  while ((int)ctx->workers.size() < nw) ctx->workers.push_back(make_worker(ctx));
  std::atomic<int64_t> next{begin};
  auto body = [&](int wi) {
    ttg_ctx *w = ctx->workers[wi];
    if (hipSetDevice(w->device) != hipSuccess) return;
    for (;;) {
      const int64_t j = next.fetch_add(1);
      if (j >= end) break;
      const int64_t t = order[j];
      int64_t cap = 0, used = 0;
      try {
        round_capacity(a.d[t], a.n + a.nofs[t], a.r + a.rofs[t], &cap);
        st[t] = round_dev(w, a.d[t], a.n + a.nofs[t], a.r + a.rofs[t], a.dcore + a.cofs[t], a.eps,
                          a.rmax ? a.rmax + a.rofs[t] : nullptr, a.prec, a.r_out + a.rofs[t], a.dcore_out + a.cofs_out[t], cap,
                          &used, &inf[t]);
      } catch (const Error &e) {
        st[t] = e.status;
        err[t] = e.what();
      } catch (...) {
        st[t] = TTG_HIP_ERROR;
        err[t] = "unexpected exception";
      }
    }
  };
  std::vector<std::thread> th;
  th.reserve(nw);
  for (int wi = 0; wi < nw; ++wi) th.emplace_back(body, wi);
  for (auto &x : th) x.join();
}

// Persistent launch over tensors sel[0..m) (indices into the batch, LPT order).
struct PShape {
  int nt, per_cu;
  int64_t ycap;
};
static const PShape kPShapes[3] = {{256, 1, 16448}, {256, 2, 6144}, {512, 1, 16448}};

static void run_persist(ttg_ctx *ctx, const BatchArgs &a, int64_t count, const std::vector<int64_t> &sel,
                        const PShape &sh, std::vector<ttg_status> &st, std::vector<ttg_info> &inf,
                        std::vector<std::string> &err) {
  const int64_t m = (int64_t)sel.size();
  if (m == 0) return;
  hipStream_t s = ctx->stream;
  int64_t ntot = 0, rtot = 0, slot_core = 1, slot_kmax = 1;
  for (int64_t t = 0; t < count; ++t) {
    ntot = std::max(ntot, a.nofs[t] + a.d[t]);
    rtot = std::max(rtot, a.rofs[t] + a.d[t] + 1);
  }
  std::vector<PTensorDesc> desc(count);
  for (int64_t t = 0; t < count; ++t) desc[t] = PTensorDesc{(int32_t)a.d[t], 0, a.nofs[t], a.rofs[t], a.cofs[t], a.cofs_out[t]};
  for (int64_t t : sel) {
    const int64_t dd = a.d[t], *nn = a.n + a.nofs[t], *rr = a.r + a.rofs[t];
    int64_t rq = rr[0];
    for (int64_t i = 0; i < dd; ++i) {
      slot_core = std::max(slot_core, rr[i] * nn[i] * rr[i + 1]);
      slot_kmax = std::max(slot_kmax, rr[i + 1]);
      slot_kmax = std::max(slot_kmax, rr[i]);
      const int64_t rn = (i + 1 < dd) ? std::min(rq * nn[i], rr[i + 1]) : rr[dd];
      slot_kmax = std::max(slot_kmax, std::min(rq, nn[i] * rn));
      rq = rn;
    }
  }
  const int slots = (int)std::min<int64_t>(m, (int64_t)ctx->num_cus * sh.per_cu);
  const int64_t slot_elems = persist_slot_elems(slot_core, slot_kmax);
  ArenaSizer z;
  z.take<PTensorDesc>(count);
  z.take<int>(m);
  z.take<int64_t>(ntot);
  z.take<int64_t>(rtot);
  z.take<int64_t>(rtot);
  z.take<int>(4 * count);
  z.take<int>(64);
  z.take<double>((size_t)slots * slot_elems);
  ctx->parena.reserve(z.used + 4096);
  ctx->parena.reset();
  PTensorDesc *ddesc = ctx->parena.take<PTensorDesc>(count);
  int *dorder = ctx->parena.take<int>(m);
  int64_t *dn = ctx->parena.take<int64_t>(ntot);
  int64_t *dr = ctx->parena.take<int64_t>(rtot);
  int64_t *drout = ctx->parena.take<int64_t>(rtot);
  int *dstat = ctx->parena.take<int>(4 * count);
  int *dctr = ctx->parena.take<int>(64);
  double *dws = ctx->parena.take<double>((size_t)slots * slot_elems);
  std::vector<int> hord(sel.begin(), sel.end());
  TTG_HIP(hipMemcpyAsync(ddesc, desc.data(), count * sizeof(PTensorDesc), hipMemcpyHostToDevice, s));
  TTG_HIP(hipMemcpyAsync(dorder, hord.data(), m * sizeof(int), hipMemcpyHostToDevice, s));
  TTG_HIP(hipMemcpyAsync(dn, a.n, ntot * sizeof(int64_t), hipMemcpyHostToDevice, s));
  TTG_HIP(hipMemcpyAsync(dr, a.r, rtot * sizeof(int64_t), hipMemcpyHostToDevice, s));
  TTG_HIP(hipMemcpyAsync(drout, a.r_out, rtot * sizeof(int64_t), hipMemcpyHostToDevice, s));
  persist_prepare(sh.nt, sh.ycap);
  const int kSweeps = jacobi_sweep_cap();
  // the captured launch is replayed while every argument is unchanged: the
  // key holds every value and pointer persist_launch bakes into the graph
  const std::vector<uint64_t> key = {(uint64_t)sh.nt, (uint64_t)slots, (uint64_t)m, (uint64_t)count,
                                     (uint64_t)(uintptr_t)ddesc, (uint64_t)(uintptr_t)dorder,
                                     (uint64_t)(uintptr_t)dn, (uint64_t)(uintptr_t)dr,
                                     (uint64_t)(uintptr_t)drout, (uint64_t)(uintptr_t)dstat,
                                     (uint64_t)(uintptr_t)dctr,
                                     (uint64_t)(uintptr_t)a.dcore, (uint64_t)(uintptr_t)a.dcore_out,
                                     (uint64_t)(uintptr_t)dws, (uint64_t)slot_core, (uint64_t)slot_kmax,
                                     (uint64_t)sh.ycap, (uint64_t)kSweeps, (uint64_t)std::hash<double>{}(a.eps),
                                     (uint64_t)(uintptr_t)ctx->parena.base};
  if (ctx->worker) {
    // on a stream worker: eager launch (no capture beside the other workers' allocations)
    persist_launch(s, sh.nt, slots, (int)m, ddesc, dorder, dn, dr, a.dcore, a.dcore_out, drout, dws, slot_core,
                   slot_kmax, dctr, dstat, a.eps, kSweeps, sh.ycap);
  } else if (!ctx->pgraph || ctx->pgraph_key != key) {
    if (ctx->pgraph) {
      TTG_HIP(hipGraphExecDestroy(ctx->pgraph));
      ctx->pgraph = nullptr;
    }
    hipGraph_t g = nullptr;
    TTG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      persist_launch(s, sh.nt, slots, (int)m, ddesc, dorder, dn, dr, a.dcore, a.dcore_out, drout, dws, slot_core,
                     slot_kmax, dctr, dstat, a.eps, kSweeps, sh.ycap);
    } catch (...) {
      (void)hipStreamEndCapture(s, &g);  // leave the stream usable
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    TTG_HIP(hipStreamEndCapture(s, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TTG_HIP(ie);
    ctx->pgraph = ge;
    ctx->pgraph_key = key;
  }
  if (!ctx->worker) TTG_HIP(hipGraphLaunch(ctx->pgraph, s));
  std::vector<int> hstat(4 * count);
  TTG_HIP(hipMemcpyAsync(a.r_out, drout, rtot * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  TTG_HIP(hipMemcpyAsync(hstat.data(), dstat, 4 * count * sizeof(int), hipMemcpyDeviceToHost, s));
  TTG_HIP(hipStreamSynchronize(s));
  for (int64_t t : sel) {
    const int code = hstat[4 * t];
    inf[t] = ttg_info{0, hstat[4 * t + 1], hstat[4 * t + 2], 0.0};
    if (code & 2) {
      st[t] = TTG_NONFINITE;
      err[t] = "input core contains NaN or Inf";
    } else if (code & 4) {
      st[t] = TTG_HIP_ERROR;
      err[t] = "persistent workspace too small";
    } else if (code & 1) {
      st[t] = TTG_NO_CONVERGENCE;
    }
  }
}

// A single small FP64 rounding (every QR-capped rank <= 64) as a one-tensor
// persistent launch: one workgroup runs the whole round with no host round
// trip, which beats the multi-CU bond-by-bond path where every bond is a few
// tiny kernels and a host synchronisation.  Returns false if the tensor is
// not eligible (the caller then uses the general path).
bool round_small_persist(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r, const double *dcore,
                         double eps, const int64_t *rmax, int64_t *r_out, double *dcore_out, int64_t cap,
                         int64_t *len_out, ttg_info *info, ttg_status *status) {
  const char *env = std::getenv("TTG_SMALL_PERSIST");
  if (env && env[0] == '0') return false;
  if (!persist_supported(d, n, r)) return false;
  std::vector<int64_t> rc(r, r + d + 1);
  for (int64_t i = 0; i + 1 < d; ++i) rc[i + 1] = std::min(rc[i] * n[i], rc[i + 1]);
  for (int64_t i = 0; i <= d; ++i)
    if (rc[i] > 64) return false;
  if (rmax)
    for (int64_t i = 1; i < d; ++i)
      if (rmax[i] < rc[i]) return false;  // a binding rmax cap: general path
  auto t0 = std::chrono::steady_clock::now();
  int64_t need = 0, o = 0;
  round_capacity(d, n, r, &need);
  if (cap < need) {
    if (len_out) *len_out = need;
    *status = TTG_NEED_SPACE;
    return true;
  }
  const int64_t zero = 0;
  for (int64_t i = 0; i <= d; ++i) r_out[i] = r[i];
  const BatchArgs a{&d, &zero, n, &zero, r, &zero, &zero, dcore, dcore_out, r_out, eps, TTG_PREC_FP64};
  std::vector<ttg_status> st(1, TTG_OK);
  std::vector<ttg_info> inf(1, ttg_info{0, 0, 0, 0.0});
  std::vector<std::string> err(1);
  run_persist(ctx, a, 1, std::vector<int64_t>{0}, kPShapes[2], st, inf, err);
  if (st[0] != TTG_OK && st[0] != TTG_NO_CONVERGENCE) throw Error(st[0], err[0]);
  for (int64_t i = 0; i < d; ++i) o += r_out[i] * n[i] * r_out[i + 1];
  if (len_out) *len_out = o;
  if (info) {
    *info = inf[0];
    info->err_est = 4.0 * 1.2e-16 * (double)d;
    info->orth_est = 1.2e-16 * (double)d;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  *status = st[0];
  return true;
}

ttg_status round_batched_dev(ttg_ctx *ctx, int64_t count, const int64_t *d, const int64_t *nofs,
                             const int64_t *n, const int64_t *rofs, const int64_t *r,
                             const int64_t *cofs, const double *dcore, double eps, const int64_t *rmax,
                             int prec, int64_t *r_out, const int64_t *cofs_out, double *dcore_out,
                             ttg_info *info) {
  TTG_CHECK(count >= 0, TTG_BAD_ARG, "count must be >= 0");
  TTG_CHECK(count == 0 || (d && nofs && n && rofs && r && cofs && dcore && r_out && cofs_out && dcore_out),
            TTG_BAD_ARG, "null pointer argument");
  if (info) *info = ttg_info{0, 0, 0, 0.0};
  if (count == 0) return TTG_OK;
  auto t0 = std::chrono::steady_clock::now();
  // cost estimates (also validates every tensor before any work starts)
  std::vector<double> cost(count);
  for (int64_t t = 0; t < count; ++t) {
    int64_t cap = 0;
    ttg_status s = round_capacity(d[t], n + nofs[t], r + rofs[t], &cap);
    if (s != TTG_OK) return s;
    const int64_t *nn = n + nofs[t], *rr = r + rofs[t];
    if (rmax)
      for (int64_t i = 1; i < d[t]; ++i)
        TTG_CHECK(rmax[rofs[t] + i] >= 1, TTG_BAD_ARG, "rmax entries must be >= 1");
    double c = 0.0, rk = (double)rr[0];
    for (int64_t i = 0; i + 1 < d[t]; ++i) {
      const double m = rk * (double)nn[i], k = std::min(m, (double)rr[i + 1]);
      c += m * k * k;
      rk = k;
    }
    cost[t] = c + 1.0;
  }
  std::vector<int64_t> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return cost[x] > cost[y]; });
  // the inputs come from the caller's stream
  TTG_HIP(hipStreamSynchronize(ctx->stream));
  BatchArgs a{d, nofs, n, rofs, r, cofs, cofs_out, dcore, dcore_out, r_out, eps, prec};
  a.rmax = rmax;
  std::vector<ttg_status> st(count, TTG_OK);
  std::vector<ttg_info> inf(count, ttg_info{0, 0, 0, 0.0});
  std::vector<std::string> err(count);
  // tensors whose rmax binds (an entry below the left-capped rank,
  // round.m:48) go to the stream workers: the one-workgroup kernel has no
  // rank cap.  The rest keep the LPT order.
  std::vector<int64_t> order_w;
  if (rmax) {
    std::vector<int64_t> keep;
    for (int64_t t : order) {
      bool binds = false;
      int64_t rk = r[rofs[t]];
      for (int64_t i = 0; i + 1 < d[t] && !binds; ++i) {
        rk = std::min(rk * n[nofs[t] + i], r[rofs[t] + i + 1]);
        binds = rmax[rofs[t] + i + 1] < rk;
      }
      (binds ? order_w : keep).push_back(t);
    }
    order.swap(keep);
  }
  // persistent launch for FP64 batches that fit one workgroup per tensor
  bool persist_ok = (prec == TTG_PREC_FP64) && !order.empty();
  for (int64_t t : order) {
    if (!persist_ok) break;
    persist_ok = persist_supported(d[t], n + nofs[t], r + rofs[t]);
  }
  if (!persist_ok) {  // everything goes to the stream workers, in LPT order
    order.insert(order.end(), order_w.begin(), order_w.end());
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return cost[x] > cost[y]; });
    order_w.swap(order);
    order.clear();
  }
  const int64_t np = (int64_t)order.size(), nw = (int64_t)order_w.size();
  if (persist_ok) {
    auto work = [&](const std::vector<int64_t> &sel) {
      double c = 0.0;
      for (int64_t t : sel) c += cost[t];
      return c;
    };
    if (ctx->batch_mode == 0) {
      // TTG_BATCH_MODE=1..4 fixes the execution (1-3: persistent launch
      // shapes, 4: stream workers) instead of tuning
      const char *env = std::getenv("TTG_BATCH_MODE");
      if (env && env[0] >= '1' && env[0] <= '4' && env[1] == '\0') ctx->batch_mode = env[0] - '0';
    }
    if (ctx->batch_mode == 0 && np >= 64) {
      // tune on interleaved slices of the LPT order (every tensor is rounded
      // exactly once): the three persistent shapes on an eighth each, the
      // stream workers on a thirty-second, the rest with the fastest
      std::vector<std::vector<int64_t>> sl(5);
      for (int64_t j = 0; j < np; ++j) {
        const int64_t c8 = j % 8;
        sl[c8 < 3 ? c8 : (j % 32 == 3 ? 3 : 4)].push_back(order[j]);
      }
      double best = 1e300;
      int bestm = 3;
      for (int c = 0; c < 4; ++c) {
        if (sl[c].empty()) continue;
        auto s0 = std::chrono::steady_clock::now();
        if (c < 3) run_persist(ctx, a, count, sl[c], kPShapes[c], st, inf, err);
        else run_batch(ctx, a, sl[c], 0, (int64_t)sl[c].size(), (int)std::min<int64_t>(16, sl[c].size()), st, inf, err);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
        const double rate = sec / work(sl[c]);
        if (rate < best) {
          best = rate;
          bestm = c + 1;
        }
      }
      ctx->batch_mode = bestm;
      if (bestm <= 3) run_persist(ctx, a, count, sl[4], kPShapes[bestm - 1], st, inf, err);
      else run_batch(ctx, a, sl[4], 0, (int64_t)sl[4].size(), (int)std::min<int64_t>(16, sl[4].size()), st, inf, err);
    } else {
      const int mode = ctx->batch_mode == 0 ? 3 : ctx->batch_mode;
      if (mode <= 3) {
        run_persist(ctx, a, count, order, kPShapes[mode - 1], st, inf, err);
      } else {
        run_batch(ctx, a, order, 0, np, (int)std::min<int64_t>(16, np), st, inf, err);
      }
    }
  }
  int64_t begin = 0;
  if (nw > 0 && ctx->batch_workers == 0) {
    // tune the worker count on consecutive slices of the (cost-ordered) batch
    const int cand[3] = {8, 16, 32};
    const int64_t slice = nw / 8;
    if (slice >= 4) {
      double best = 1e300;
      int bestw = 16;
      for (int ci = 0; ci < 3 && begin + slice <= nw; ++ci) {
        auto s0 = std::chrono::steady_clock::now();
        run_batch(ctx, a, order_w, begin, begin + slice, cand[ci], st, inf, err);
        double cs = 0.0;
        for (int64_t j = begin; j < begin + slice; ++j) cs += cost[order_w[j]];
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
        const double rate = sec / cs;  // seconds per unit of estimated work
        if (rate < best) {
          best = rate;
          bestw = cand[ci];
        }
        begin += slice;
      }
      ctx->batch_workers = bestw;
    } else {
      ctx->batch_workers = 16;
    }
  }
  if (begin < nw)
    run_batch(ctx, a, order_w, begin, nw, (int)std::min<int64_t>(ctx->batch_workers, nw - begin), st, inf, err);
  ttg_status worst = TTG_OK;
  ttg_info acc{0, 0, 0, 0.0};
  for (int64_t t = 0; t < count; ++t) {
    if (st[t] != TTG_OK && st[t] != TTG_NO_CONVERGENCE) {
      ctx->last_error = "tensor " + std::to_string(t) + ": " + err[t];
      return st[t];
    }
    if (st[t] == TTG_NO_CONVERGENCE) worst = st[t];
    acc.qr_fallbacks += inf[t].qr_fallbacks;
    acc.jacobi_sweeps += inf[t].jacobi_sweeps;
    acc.jacobi_maxsweeps = std::max(acc.jacobi_maxsweeps, inf[t].jacobi_maxsweeps);
    acc.qr_retries += inf[t].qr_retries;
  }
  acc.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (info) *info = acc;
  return worst;
}

}  // namespace ttg
