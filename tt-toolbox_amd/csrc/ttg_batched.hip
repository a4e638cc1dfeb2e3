// Batched rounding of many independent TT tensors (uneven d, n, r).
//
// Each rounding is a chain of latency-bound bond steps that occupies only a
// few CUs at a time, so a batch is run as W concurrent chains: W worker
// contexts, each with its own HIP stream and device workspace, driven by W
// host threads that take tensors longest-first from a shared counter
// (a greedy LPT schedule; the cost estimate is the QR work sum_i m_i k_i^2 at
// the QR-capped ranks).  W is tuned once per context: the first batch large
// enough is timed on a prefix with W in {8, 16, 32} and the fastest is kept.
#include "ttg_internal.h"
#include <algorithm>
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
  if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
    delete w;
    throw Error(TTG_HIP_ERROR, "cannot create a worker stream");
  }
  w->own_stream = true;
  return w;
}

void destroy_workers(ttg_ctx *ctx) {
  for (ttg_ctx *w : ctx->workers) {
    (void)hipStreamSynchronize(w->stream);
    (void)hipStreamDestroy(w->stream);
    for (auto e : w->prof.pool) (void)hipEventDestroy(e);
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
                          nullptr, a.prec, a.r_out + a.rofs[t], a.dcore_out + a.cofs_out[t], cap,
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

ttg_status round_batched_dev(ttg_ctx *ctx, int64_t count, const int64_t *d, const int64_t *nofs,
                             const int64_t *n, const int64_t *rofs, const int64_t *r,
                             const int64_t *cofs, const double *dcore, double eps, int prec,
                             int64_t *r_out, const int64_t *cofs_out, double *dcore_out,
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
  const BatchArgs a{d, nofs, n, rofs, r, cofs, cofs_out, dcore, dcore_out, r_out, eps, prec};
  std::vector<ttg_status> st(count, TTG_OK);
  std::vector<ttg_info> inf(count, ttg_info{0, 0, 0, 0.0});
  std::vector<std::string> err(count);
  int64_t begin = 0;
  if (ctx->batch_workers == 0) {
    // tune the worker count on consecutive slices of the (cost-ordered) batch
    const int cand[3] = {8, 16, 32};
    const int64_t slice = count / 8;
    if (slice >= 4) {
      double best = 1e300;
      int bestw = 16;
      for (int ci = 0; ci < 3 && begin + slice <= count; ++ci) {
        auto s0 = std::chrono::steady_clock::now();
        run_batch(ctx, a, order, begin, begin + slice, cand[ci], st, inf, err);
        double cs = 0.0;
        for (int64_t j = begin; j < begin + slice; ++j) cs += cost[order[j]];
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
  if (begin < count)
    run_batch(ctx, a, order, begin, count, (int)std::min<int64_t>(ctx->batch_workers, count - begin), st,
              inf, err);
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
  }
  acc.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (info) *info = acc;
  return worst;
}

}  // namespace ttg
