// Host orchestration of TT rounding on one GPU (sweeps of @tt_tensor/round.m).
//
// Sweep 1 (round.m:39-53), bond i = 0..d-2, left unfolding A = core_i as
// (r_i n_i) x r_{i+1}:
//   tall  -> CholeskyQR2 (plain: FP64 Grams; mixed: FP32 MFMA Grams of the
//            FP64 data, escalated to FP64 Grams when rejected; FP64 Cholesky
//            and solves in both); rejected factorizations
//            (Cholesky breakdown, diag(R) spread > 1e7, second-pass R far from
//            I) are redone by one-workgroup Householder QR;
//   wide  -> Householder QR (Q square, rank capped at r_i n_i, round.m:48);
//   nrm[i+1] = ||R||_F, R /= nrm, core_{i+1} <- R core_{i+1} (MFMA GEMM).
// Sweep 2 (round.m:59-89), bond i = d-1..1, W = core_i as r_i x n_i r_{i+1}:
//   QR of the tall orientation, one-sided block Jacobi SVD of the square
//   factor stacked over the identity, descending sort + my_chop2 on the
//   device, then core_i <- V(:,1:r1)^T and core_{i-1} <- core_{i-1} U S.
// Finalise (round.m:94-107): normalise core_1, balance all cores by
// exp(mean(log|nrm|)).
//
// The first sweep runs speculatively (one host synchronisation at its end;
// a rejected bond and everything after it is recomputed).  The second sweep
// synchronises once per bond to learn r1, which sizes the next launches.
#include "ttg_internal.h"
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <cmath>
#include <functional>
#include <algorithm>

namespace ttg {


static constexpr int kMaxRank = 512;
static constexpr int kMaxSweeps = 60;

// Jacobi sweep cap: kMaxSweeps, or TTG_JACOBI_MAX_SWEEPS (1..kMaxSweeps) so
// that the NO_CONVERGENCE status can be forced from the host in tests (the
// iteration then ends after that many sweeps; every kernel exits normally)
int jacobi_sweep_cap() {
  const char *env = std::getenv("TTG_JACOBI_MAX_SWEEPS");
  if (!env) return kMaxSweeps;
  const int v = std::atoi(env);
  return (v >= 1 && v <= kMaxSweeps) ? v : kMaxSweeps;
}

// capped ranks of the first sweep (round.m:48)
static std::vector<int64_t> capped_ranks(int64_t d, const int64_t *n, const int64_t *r) {
  std::vector<int64_t> rr(r, r + d + 1);
  for (int64_t i = 0; i + 1 < d; ++i) rr[i + 1] = std::min(rr[i] * n[i], rr[i + 1]);
  return rr;
}

static void validate_tt(int64_t d, const int64_t *n, const int64_t *r) {
  TTG_CHECK(d >= 1, TTG_BAD_ARG, "d must be >= 1");
  TTG_CHECK(n != nullptr && r != nullptr, TTG_BAD_ARG, "n and r are required");
  for (int64_t i = 0; i < d; ++i) TTG_CHECK(n[i] >= 1, TTG_BAD_ARG, "mode sizes must be >= 1");
  for (int64_t i = 0; i <= d; ++i) TTG_CHECK(r[i] >= 1, TTG_BAD_ARG, "ranks must be >= 1");
}

template <class T> struct Rounder {
  ttg_ctx *ctx;
  hipStream_t st;
  int cus;
  int64_t d;
  std::vector<int64_t> n, r, ofs;
  // buf[0]: input cores (sweep 1 never writes it); buf[1]: outputs of the
  // R*core GEMMs; buf[2]: Q factors.  Every sweep-1 bond therefore keeps its
  // inputs and can be recomputed after a rejected Cholesky-QR.
  T *buf[3];
  std::vector<T *> cp;
  // scratch
  double *Winv1, *Winv2;  // inverses of the diagonal blocks of R1 / R2
  double *G, *R1, *R2, *R4, *Rinv, *Rprod, *Rtmp, *Y, *sv, *nrm, *stats, *bstats, *Yws;
  T *RT, *Q1, *Q2, *Tt, *Qs, *US, *gws;
  int64_t gws_elems;
  double *hA, *hQ, *htau, *hws;
  int *perm, *iflag, *scal_i, *jerr, *jsweeps;
  unsigned *bar;
  double *scal_d;
  int64_t max_core, kmax, rsq;
  bool mixed;
  // optional: dst (kk x n_i r_{i+1}) = Rm (kk x r_i) * core_i for generated
  // cores (fused A*x: the product core is built on chip, never stored)
  const std::function<void(int64_t, const T *, int64_t, T *)> *fmul = nullptr;
  int64_t fallbacks = 0, sweeps_total = 0, sweeps_max = 0, retries = 0, svd_skipped = 0;
  bool allow_notrunc = true;
  int next_level = 2;  // QR level after the last rejected svd_bond
  bool noconv = false;

  // a buffer region for core k that is not the current one
  T *other(int64_t k) const {
    for (int b = 0; b < 3; ++b)
      if (cp[k] != buf[b] + ofs[k]) return buf[b] + ofs[k];
    return buf[0] + ofs[k];
  }

  static size_t jac_elems(int64_t kmax) { return (size_t)(2 * kmax) * (size_t)(kmax + 64); }
  // V (m x kq <= one core) + T + two HB x n blocks + split-K partials
  static size_t hh_ws(int64_t max_core, int64_t kmax) {
    return (size_t)max_core + (size_t)(3 * 32 + 32 * 32) * (size_t)(kmax + 64) + 4096;
  }

  template <class A> static void layout(A &a, int64_t d, int64_t total, int64_t max_core, int64_t kmax, int64_t rsq) {
    for (int b = 0; b < 3; ++b) a.template take<T>(total);
    for (int q = 0; q < 7; ++q) a.template take<double>(rsq);
    for (int q = 0; q < 2; ++q) a.template take<double>((kmax / 32 + 1) * 33 * 32);  // Winv1, Winv2
    a.template take<double>(jac_elems(kmax));
    a.template take<double>(jac_elems(kmax) + (kmax + 64) * (kmax + 64));  // polar correction
    a.template take<double>(kmax + 64);      // sv
    a.template take<double>(d + 8);          // nrm (+ balance scales)
    a.template take<double>(64);             // stats
    a.template take<double>(8 * (d + 8));    // per-bond stats
    a.template take<T>(rsq);                 // RT
    for (int q = 0; q < 5; ++q) a.template take<T>(max_core);  // Q1 Q2 Tt Qs US
    a.template take<T>(std::max<int64_t>(32 * kmax * kmax, 1 << 16));  // gemm split-K
    a.template take<double>(max_core);       // hA
    a.template take<double>(max_core);       // hQ
    a.template take<double>(kmax + 64);      // htau
    a.template take<double>(hh_ws(max_core, kmax));  // blocked Householder workspace
    a.template take<int>(kmax + 64);         // perm
    a.template take<int>(4 * (d + 8));       // iflag (per bond)
    a.template take<int>(16);                // scal_i
    a.template take<int>(16);                // jerr
    a.template take<int>(16);                // jsweeps
    a.template take<unsigned>(jacobi_bar_words(kMaxSweeps));
    a.template take<double>(16);             // scal_d
  }

  void carve(Arena &a, int64_t total) {
    for (int b = 0; b < 3; ++b) buf[b] = a.take<T>(total);
    G = a.take<double>(rsq);
    R1 = a.take<double>(rsq);
    R2 = a.take<double>(rsq);
    Rinv = a.take<double>(rsq);
    Rprod = a.take<double>(rsq);
    Rtmp = a.take<double>(rsq);
    R4 = a.take<double>(rsq);
    Winv1 = a.take<double>((kmax / 32 + 1) * 33 * 32);
    Winv2 = a.take<double>((kmax / 32 + 1) * 33 * 32);
    Y = a.take<double>(jac_elems(kmax));
    Yws = a.take<double>(jac_elems(kmax) + (kmax + 64) * (kmax + 64));
    sv = a.take<double>(kmax + 64);
    nrm = a.take<double>(d + 8);
    stats = a.take<double>(64);
    bstats = a.take<double>(8 * (d + 8));
    RT = a.take<T>(rsq);
    Q1 = a.take<T>(max_core);
    Q2 = a.take<T>(max_core);
    Tt = a.take<T>(max_core);
    Qs = a.take<T>(max_core);
    US = a.take<T>(max_core);
    gws_elems = std::max<int64_t>(32 * kmax * kmax, 1 << 16);
    gws = a.take<T>(gws_elems);
    hA = a.take<double>(max_core);
    hQ = a.take<double>(max_core);
    htau = a.take<double>(kmax + 64);
    hws = a.take<double>(hh_ws(max_core, kmax));
    perm = a.take<int>(kmax + 64);
    iflag = a.take<int>(4 * (d + 8));
    scal_i = a.take<int>(16);
    jerr = a.take<int>(16);
    jsweeps = a.take<int>(16);
    bar = a.take<unsigned>(jacobi_bar_words(kMaxSweeps));
    scal_d = a.take<double>(16);
  }

  // ---- small helpers ------------------------------------------------------
  // 4 KB of pinned host memory for the per-bond read-backs (a plain DMA; the
  // synchronisation then waits only for the GPU)
  double *pinned(size_t bytes = 4096) {
    if (!ctx->host_pinned || ctx->host_pinned_cap < bytes) {
      // (every earlier use of the buffer has been waited for: each read-back
      // synchronises, and the finalisation reuses it only after the sweeps)
      if (ctx->host_pinned) TTG_HIP(hipHostFree(ctx->host_pinned));
      ctx->host_pinned = nullptr;
      const size_t cap = std::max<size_t>(bytes, 4096);
      TTG_HIP(hipHostMalloc(&ctx->host_pinned, cap, hipHostMallocDefault));
      ctx->host_pinned_cap = cap;
    }
    return static_cast<double *>(ctx->host_pinned);
  }
  void g(bool ta, bool tb, int64_t M, int64_t N, int64_t K, const T *A, int64_t lda, const T *B,
         int64_t ldb, T *C, int64_t ldc) {
    gemm<T>(st, ta, tb, M, N, K, T(1), A, lda, B, ldb, T(0), C, ldc, gws, gws_elems, cus);
  }
  void gram(const T *A, int64_t m, int64_t k, double *Gout, bool g32) {
    if constexpr (std::is_same<T, double>::value) {
      if (g32)  // FP32 MFMA Gram of FP64 data (its error only perturbs Q's orthogonality)
        gemm_f64_via_f32(st, true, false, k, k, m, A, m, A, m, Gout, k, reinterpret_cast<float *>(gws),
                         2 * gws_elems, cus);
      else
        gemm<double>(st, true, false, k, k, m, 1.0, A, m, A, m, 0.0, Gout, k, gws, gws_elems, cus);
    } else {
      gemm_f32_f64out(st, true, false, k, k, m, A, m, A, m, Gout, k, gws, gws_elems, cus);
    }
  }
  // dst (T) = src (double)
  void to_T(T *dst, const double *src, int64_t count) {
    if constexpr (std::is_same<T, double>::value) {
      TTG_HIP(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
      convert<float, double>(st, dst, src, count);
    }
  }

  // One Cholesky-QR pass: R = chol(A^T A (+ shift)), Qout = A R^{-1} by triangular solve.
  // W (optional): receives the inverses of R's diagonal blocks, which the
  // solve then uses (returns whether it did)
  bool cholqr_pass(const T *A, int64_t m, int64_t k, double shift_rel, T *Qout, double *Rout,
                   int *flag, double *st2, bool g32, double *W, double *fro = nullptr) {
    gram(A, m, k, G, g32);
    const bool w = chol_factor(st, k, G, k, shift_rel, Rout, flag, st2, nullptr, W, fro);
    trsm_right_upper<T>(st, m, k, A, m, Rout, Qout, m, nullptr, w ? W : nullptr);
    return w;
  }

  // Tall QR (m >= k): Qout (m x k, T), Rout (k x k, FP64), by CholeskyQR2.
  // flags[0..1] and stt[0..3] feed the acceptance test (qr_accept), checked
  // by the caller.  QR level 0: the second pass by the near-identity series
  // (gram_nearid): R2 = I + Phi(E), Q = Q1 (I - Phi(E)).  FP64 Grams:
  // accepted if max|E| < 1e-8 (both truncations O(1e-16)).  The mixed form's
  // sweep-1 bonds (g32): Gram products in FP32 MFMA (their error only
  // perturbs the orthogonality of Q; every Q = A R^{-1} step is FP64),
  // accepted if max|E| < 1e-5 (Q orthogonal and Q R2 = Q1 to 1e-10, inside
  // the form's 1e-6 budget).  Sweep-2
  // unfoldings carry the bond's singular-value spread, beyond that bound with
  // FP32 Grams, so they use FP64 Grams in both forms.  Level 1: FP64 Grams,
  // Cholesky + substitution second pass (accepted if its R is within 0.1 of
  // I).  Level 2 (caller): Householder.
  bool nearid_level(int level) const { return level == 0; }
  // bound (level 0 only): also start, on the side stream, ||R1||_F and
  // ||R1^{-1}||_F for the no-truncation test (R = R2 R1 with ||R2 - I|| of
  // order k max|E| <= 5e-6, so the bounds carry over with the test's factor
  // 2 margin); joined by the caller through ctx->ev_join.
  void qr_tall_fast(const T *A, int64_t m, int64_t k, T *Qout, double *Rout, int *flags,
                    double *stt, int level, bool allow_g32, bool bound = false) {
    const bool g32 = mixed && level == 0 && allow_g32;
    // (with the bound: ||R1||_F from the factorisation itself when it can)
    const bool w1 = cholqr_pass(A, m, k, 0.0, Q1, R1, flags + 0, stt + 0, g32, Winv1, bound ? scal_d + 0 : nullptr);
    if (bound && nearid_level(level)) {
      TTG_HIP(hipEventRecord(ctx->ev_fork, st));
      TTG_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
      if (k <= 256) {
        trsm_right_upper<double>(ctx->side, k, k, nullptr, k, R1, Rtmp, k, nullptr,
                                 w1 ? Winv1 : nullptr);  // Rtmp = I R1^{-1}
      } else {
        eye_plus(ctx->side, k, R4, nullptr, 0.0);
        trsm_right_upper<double>(ctx->side, k, k, R4, k, R1, Rtmp, k);
      }
      fro_norm<double>(ctx->side, Rtmp, k * k, scal_d + 1);
      TTG_HIP(hipEventRecord(ctx->ev_join, ctx->side));
      if (!w1) fro_norm<double>(st, R1, k * k, scal_d + 0);
    }
    if (nearid_level(level)) {
      gram_nearid(st, m, k, Q1, g32, G, Rinv, R2, stt + 3, gws, gws_elems, cus);
      g(false, false, m, k, k, Q1, m, Rinv, k, Qout, m);  // Q = Q1 R2^{-1}
    } else {
      cholqr_pass(Q1, m, k, 0.0, Qout, R2, flags + 1, stt + 2, g32, Winv2);
    }
    gemm<double>(st, false, false, k, k, k, 1.0, R2, k, R1, k, 0.0, Rout, k, nullptr, 0, cus);
  }
  // level after a rejection: a failed FP64 first pass (Cholesky breakdown or
  // diagonal spread) fails the same way at level 1, so it goes to Householder
  int qr_next(const int *hflags, const double *hstats, int level, bool allow_g32) const {
    const bool g32 = mixed && level == 0 && allow_g32;
    if (level == 0 && !g32 && (hflags[0] || !(hstats[0] < 1e7))) return 2;
    return level + 1;
  }
  bool qr_accept(const int *hflags, const double *hstats, int level, bool allow_g32) const {
    // flags[1] is the second pass's Cholesky, which only levels >= 1 run
    // (level 0's second pass is the near-identity series): at level 0 it is
    // not written, and not read
    if (hflags[0] || (!nearid_level(level) && hflags[1])) return false;
    if (!(hstats[0] < 1e7)) return false;   // diag(R1) spread: kappa estimate
    // Q1 nearly orthonormal: max|Q1^T Q1 - I| (level 0) or max|R2 - I| (level 1)
    const bool g32 = mixed && level == 0 && allow_g32;
    if (!(hstats[3] < (nearid_level(level) ? (g32 ? 1e-5 : 1e-8) : 0.1))) return false;
    return true;
  }
  // Householder fallback (FP64) for the tall/wide QR of A (m x k).
  void qr_householder(const T *A, int64_t m, int64_t k, T *Qout, double *Rout) {
    const int64_t kq = std::min(m, k);
    if constexpr (std::is_same<T, double>::value) {
      TTG_HIP(hipMemcpyAsync(hA, A, m * k * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
      convert<double, float>(st, hA, A, m * k);
    }
    TTG_CHECK(householder_ws_elems(m, k, cus) <= hh_ws(max_core, kmax), TTG_HIP_ERROR,
              "Householder workspace too small");
    householder_qr(st, m, k, hA, hQ, Rout, htau, hws, cus);
    to_T(Qout, hQ, m * kq);
    ++fallbacks;
  }

  // ---- sweep 1 -----------------------------------------------------------
  // Bond i reads core i from (i == 0 ? buf0 : buf1) and core i+1 from buf0;
  // writes Q to buf2 and R*core_{i+1} to buf1.  Returns true when the
  // Cholesky-QR acceptance check is pending.
  // level: 0 = the form's CholeskyQR2, 1 = FP64 CholeskyQR2, 2 = Householder
  bool sweep1_bond(int64_t i, int level,
                   const std::function<void(int64_t, T *)> &gen) {
    const int64_t m = r[i] * n[i], k = r[i + 1];
    const int64_t kk = std::min(m, k);
    const T *A = (i == 0 ? buf[0] : buf[1]) + ofs[i];
    T *Qdst = buf[2] + ofs[i];
    double *R = Rprod;
    bool pending = false;
    if (m >= k && level < 2) {
      qr_tall_fast(A, m, k, Qdst, R, iflag + 4 * i, bstats + 8 * i, level, true);
      pending = true;
    } else {
      qr_householder(A, m, k, Qdst, R);
    }
    cp[i] = Qdst;
    // nrm(i+1) = ||R||_F ; R /= nrm   (round.m:41-44)
    fro_norm<double>(st, R, kk * k, nrm + i + 1);
    scale_by_inv<double>(st, R, kk * k, nrm + i + 1);
    // core_{i+1} <- R core_{i+1}   (round.m:45-47)
    const int64_t N1 = n[i + 1] * r[i + 2];
    // R in the cores' type (the FP64 form reads R itself, no copy)
    const T *Rt;
    if constexpr (std::is_same<T, double>::value) {
      Rt = R;
    } else {
      to_T(RT, R, kk * k);
      Rt = RT;
    }
    T *dst = buf[1] + ofs[i + 1];
    if (fmul && *fmul) {
      (*fmul)(i + 1, Rt, kk, dst);  // core_{i+1} generated on chip inside the product
    } else {
      T *src = buf[0] + ofs[i + 1];
      if (gen) gen(i + 1, src);
      g(false, false, kk, N1, k, Rt, kk, src, k, dst, kk);
    }
    cp[i + 1] = dst;
    r[i + 1] = kk;  // round.m:48
    return pending;
  }

  // ---- sweep 2 -----------------------------------------------------------
  void svd_bond(int64_t i, double ep, int64_t rmax_i, int level, int64_t &r1,
                bool &qr_ok, int64_t &kj, bool &wide, bool &skipped) {
    const int64_t kw = r[i], N = n[i] * r[i + 1];
    wide = (N >= kw);
    const T *W = cp[i];
    // stats[0..5]: QR statistics; the QR flags and the shortcut's decision sit
    // right behind them, so that one copy brings all three to the host
    int *fl = reinterpret_cast<int *>(stats + 8);
    int *skip = reinterpret_cast<int *>(stats + 10);
    double *R = Rprod;
    kj = wide ? kw : N;
    const int64_t mt = wide ? N : kw;  // rows of the tall orientation
    const T *tallA = W;
    if (wide) {
      transpose<T>(st, W, kw, N, Tt);
      tallA = Tt;
    }
    // (fl[0] is written by every first-pass Cholesky, fl[1] by the second
    // pass's at levels >= 1, the only ones that read it: no clearing needed)
    const bool bound = wide && allow_notrunc && level == 0;
    if (level < 2) {
      qr_tall_fast(tallA, mt, kj, Qs, R, fl, stats, level, false, bound);
    } else {
      qr_householder(tallA, mt, kj, Qs, R);
    }
    JacobiPlan pl = jacobi_plan(kj);
    int h_fl[4] = {0, 0, 0, 0};
    double h_st[6] = {0, 0, 0, 0, 0, 0};
    skipped = false;
    // No-truncation shortcut (wide bonds): when 1/||R^{-1}||_F, a lower bound
    // of the smallest singular value, clears my_chop2's threshold, every
    // singular value is kept and the SVD would only change the gauge of the
    // bond, so it is completed from its QR factors, W = R^T Q^T:
    // core_i = Q^T (orthonormal rows, as V' would have), core_{i-1} <-
    // core_{i-1} R^T.  Same tensor, same ranks.  The decision and the QR
    // acceptance are read in one synchronisation; a rejected QR is redone
    // before any SVD work.
    if (bound) {
      TTG_HIP(hipStreamWaitEvent(st, ctx->ev_join, 0));
      notrunc_decide(st, scal_d + 0, scal_d + 1, ep, kj, rmax_i, skip);
      double *h_io = pinned();
      TTG_HIP(hipMemcpyAsync(h_io, stats, 11 * sizeof(double), hipMemcpyDeviceToHost, st));
      TTG_HIP(hipStreamSynchronize(st));
      std::memcpy(h_st, h_io, sizeof(h_st));
      std::memcpy(h_fl, h_io + 8, sizeof(h_fl));
      int h_skip = 0;
      std::memcpy(&h_skip, h_io + 10, sizeof(int));
      qr_ok = level == 2 || qr_accept(h_fl, h_st, level, false);
      if (!qr_ok) {
        next_level = qr_next(h_fl, h_st, level, false);
        return;
      }
      if (h_skip) {
        skipped = true;
        ++svd_skipped;
        r1 = kj;
        return;
      }
    }
    // stopping cosine^2: plain form 1e-16 (the last sweep's rotations leave
    // O(1e-16) behind); mixed form 1e-8 (O(1e-8) left, far below its 1e-6
    // error budget, one sweep fewer)
    jacobi_svd(st, pl, R, /*transS=*/wide, Y, jsweeps, bar, jerr, jacobi_sweep_cap(), Yws, nullptr,
               mixed ? 1e-8 : 1e-16);
    sort_chop(st, Y, pl.rows, kj, pl.kpad, ep, rmax_i, sv, perm, scal_i, scal_d + 8);
    // one synchronisation: r1, Jacobi status, QR acceptance data
    int h_i[4] = {0, 0, 0, 0};
    {
      double *hp = pinned();  // [0..10] stats/flags (as above), then three ints
      int *hpi = reinterpret_cast<int *>(hp + 12);
      TTG_HIP(hipMemcpyAsync(hp, stats, 11 * sizeof(double), hipMemcpyDeviceToHost, st));
      TTG_HIP(hipMemcpyAsync(hpi, scal_i, sizeof(int), hipMemcpyDeviceToHost, st));
      TTG_HIP(hipMemcpyAsync(hpi + 1, jerr, sizeof(int), hipMemcpyDeviceToHost, st));
      TTG_HIP(hipMemcpyAsync(hpi + 2, jsweeps, sizeof(int), hipMemcpyDeviceToHost, st));
      TTG_HIP(hipStreamSynchronize(st));
      std::memcpy(h_st, hp, sizeof(h_st));
      std::memcpy(h_fl, hp + 8, sizeof(h_fl));
      std::memcpy(h_i, hpi, 3 * sizeof(int));
    }
    qr_ok = level == 2 || qr_accept(h_fl, h_st, level, false);
    if (!qr_ok) next_level = qr_next(h_fl, h_st, level, false);
    r1 = h_i[0];
    if (qr_ok) {
      if (h_i[1] == 2) throw Error(TTG_HIP_ERROR, "Jacobi grid barrier timed out");
      if (h_i[1] == 3) throw Error(TTG_HIP_ERROR, "Jacobi round barrier did not complete");
      if (h_i[1] == 1) noconv = true;
      sweeps_total += h_i[2];
      sweeps_max = std::max<int64_t>(sweeps_max, h_i[2]);
    }
  }

  void sweep2_bond(int64_t i, double ep, int64_t rmax_i) {
    int64_t r1, kj;
    bool ok, wide, skipped;
    // escalate on rejection: FP64 CholeskyQR2 with a Cholesky second pass,
    // then Householder
    int level = 0;
    svd_bond(i, ep, rmax_i, level, r1, ok, kj, wide, skipped);
    while (!ok) {
      level = next_level;
      ++retries;
      svd_bond(i, ep, rmax_i, level, r1, ok, kj, wide, skipped);
    }
    const int64_t kw = r[i], N = n[i] * r[i + 1];
    JacobiPlan pl = jacobi_plan(kj);
    T *core_i = other(i);
    if (wide && skipped) {
      // no truncation: U S = R^T, core_i = Qs^T
      transpose<T>(st, Rprod, kw, kw, US);
      transpose<T>(st, Qs, N, kw, core_i);
    } else if (wide) {
      // W^T = Qs R, R^T J = Y:  U S = Y[0:kw, perm],  V = Qs J[:, perm]
      gather_cols<T>(st, Y, pl.rows, 0, kw, perm, r1, US, kw);
      gather_cols<T>(st, Y, pl.rows, kw, kw, perm, r1, Q1, kw);  // J(:, perm(1:r1))
      // core_i (r1 x N) = J_r^T Qs^T
      g(true, true, r1, N, kw, Q1, kw, Qs, N, core_i, r1);
    } else {
      // W = Qs R, R J = Y:  U S = Qs Y[0:N, perm],  V = J[:, perm]
      gather_cols<T>(st, Y, pl.rows, 0, N, perm, r1, Q1, N);
      g(false, false, kw, r1, N, Qs, kw, Q1, N, US, kw);
      gather_cols_t<T>(st, Y, pl.rows, N, N, perm, r1, core_i);
    }
    cp[i] = core_i;
    // core_{i-1} <- core_{i-1} * (U S)   (round.m:81-83)
    const int64_t mp = r[i - 1] * n[i - 1];
    T *dst = other(i - 1);
    g(false, false, mp, r1, kw, cp[i - 1], mp, US, kw, dst, mp);
    cp[i - 1] = dst;
    r[i] = r1;
  }

  // ---- driver --------------------------------------------------------------
  void run(double eps, const int64_t *rmax, const std::function<void(int64_t, T *)> &gen) {
    // sweep 1, speculative: one synchronisation at its end.  After a rejected
    // bond the sweep restarts there one QR level up (mixed -> FP64
    // CholeskyQR2 -> Householder) and continues bond by bond with an
    // immediate check, so rank-deficient inputs, where many bonds fail,
    // cost one retry per failing bond instead of a restart each.
    std::vector<int> force(d, 0);
    const std::vector<int64_t> r_in = r;
    int64_t start = 0;
    bool careful = false;
    std::vector<int> hfl(4 * (d + 8));
    std::vector<double> hst(8 * (d + 8));
    while (d > 1 && !careful) {
      std::vector<int64_t> pending;
      TTG_HIP(hipMemsetAsync(iflag + 4 * start, 0, 4 * (d - 1 - start) * sizeof(int), st));
      for (int64_t i = start; i + 1 < d; ++i)
        if (sweep1_bond(i, force[i], gen)) pending.push_back(i);
      if (pending.empty()) break;
      TTG_HIP(hipMemcpyAsync(hfl.data(), iflag, hfl.size() * sizeof(int), hipMemcpyDeviceToHost, st));
      TTG_HIP(hipMemcpyAsync(hst.data(), bstats, hst.size() * sizeof(double), hipMemcpyDeviceToHost, st));
      TTG_HIP(hipStreamSynchronize(st));
      int64_t bad = -1;
      for (int64_t i : pending)
        if (!qr_accept(&hfl[4 * i], &hst[8 * i], force[i], true)) {
          bad = i;
          break;
        }
      if (bad < 0) break;
      force[bad] = qr_next(&hfl[4 * bad], &hst[8 * bad], force[bad], true);
      ++retries;
      for (int64_t j = bad + 1; j <= d; ++j) r[j] = r_in[j];
      start = bad;
      careful = true;
    }
    if (careful) {
      for (int64_t i = start; i + 1 < d; ++i) {
        for (;;) {
          TTG_HIP(hipMemsetAsync(iflag + 4 * i, 0, 4 * sizeof(int), st));
          if (!sweep1_bond(i, force[i], gen)) break;
          TTG_HIP(hipMemcpyAsync(hfl.data() + 4 * i, iflag + 4 * i, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
          TTG_HIP(hipMemcpyAsync(hst.data() + 8 * i, bstats + 8 * i, 8 * sizeof(double), hipMemcpyDeviceToHost, st));
          TTG_HIP(hipStreamSynchronize(st));
          if (qr_accept(&hfl[4 * i], &hst[8 * i], force[i], true)) break;
          force[i] = qr_next(&hfl[4 * i], &hst[8 * i], force[i], true);
          ++retries;
          r[i + 1] = r_in[i + 1];
        }
      }
    }
    const double ep = d > 1 ? eps / std::sqrt((double)(d - 1)) : INFINITY;
    for (int64_t i = d - 1; i >= 1; --i) {
      const int64_t rm = rmax ? rmax[i] : INT64_MAX;
      sweep2_bond(i, ep, rm);
    }
  }
};

// dst/ofs of the packed output with final ranks
static int64_t packed_size(int64_t d, const int64_t *n, const int64_t *r) {
  int64_t s = 0;
  for (int64_t i = 0; i < d; ++i) s += r[i] * n[i] * r[i + 1];
  return s;
}

template <class T>
static ttg_status round_impl(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                             const double *dcore, double eps, const int64_t *rmax, int64_t *r_out,
                             double *dcore_out, int64_t cap, int64_t *len_out, ttg_info *info,
                             const std::function<void(int64_t, T *)> &gen,
                             const std::vector<int64_t> *gen_ranks, bool mixed,
                             const std::function<void(int64_t, const T *, int64_t, T *)> *fmul = nullptr) {
  auto t0 = std::chrono::steady_clock::now();
  validate_tt(d, n, r);
  if (rmax)
    for (int64_t i = 1; i < d; ++i) TTG_CHECK(rmax[i] >= 1, TTG_BAD_ARG, "rmax entries must be >= 1");
  TTG_CHECK(eps >= 0.0 && std::isfinite(eps), TTG_BAD_ARG, "eps must be finite and >= 0");
  hipStream_t st = ctx->stream;
  Rounder<T> R;
  R.ctx = ctx;
  R.st = st;
  R.cus = ctx->num_cus;
  R.d = d;
  R.n.assign(n, n + d);
  R.r.assign(r, r + d + 1);
  R.mixed = mixed;
  R.fmul = fmul;
  if (!ctx->side) {
    TTG_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    TTG_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    TTG_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
  }
  {
    const char *env = std::getenv("TTG_SVD_ALWAYS");
    R.allow_notrunc = !(env && env[0] == '1');
  }
  std::vector<int64_t> sizes(d);
  R.ofs.assign(d + 1, 0);
  int64_t max_core = 1, kmax = 1;
  for (int64_t i = 0; i < d; ++i) {
    sizes[i] = r[i] * n[i] * r[i + 1];
    R.ofs[i + 1] = R.ofs[i] + sizes[i];
    max_core = std::max(max_core, sizes[i]);
  }
  std::vector<int64_t> rc = capped_ranks(d, n, r);
  for (int64_t i = 0; i <= d; ++i) kmax = std::max(kmax, rc[i]);
  for (int64_t i = 0; i < d; ++i) kmax = std::max(kmax, std::min(rc[i + 1], n[i] * rc[i + 1]));
  TTG_CHECK(kmax <= kMaxRank, TTG_BAD_ARG, "TT ranks above 512 are not supported");
  const int64_t cap_needed = packed_size(d, n, rc.data());
  if (cap < cap_needed) {
    if (len_out) *len_out = cap_needed;
    return TTG_NEED_SPACE;
  }
  int64_t rsq = kmax * kmax;
  for (int64_t i = 0; i + 1 < d; ++i) rsq = std::max(rsq, std::min(rc[i] * n[i], r[i + 1]) * r[i + 1]);
  R.max_core = max_core;
  R.kmax = kmax;
  R.rsq = rsq;
  ArenaSizer z;
  Rounder<T>::layout(z, d, R.ofs[d], max_core, kmax, rsq);
  z.take<int>(16);                       // nonfinite flag
  z.take<int64_t>(3 * (d + 1));          // pack tables
  z.take<void *>(d + 1);
  ctx->arena.reserve(z.used + 4096);
  ctx->arena.reset();
  R.carve(ctx->arena, R.ofs[d]);
  int *nf = ctx->arena.take<int>(16);
  int64_t *dtab = ctx->arena.take<int64_t>(3 * (d + 1));
  void **dsrc = ctx->arena.take<void *>(d + 1);
  // input -> working buffer 0 (converted for the mixed form)
  if (!gen) {
    TTG_HIP(hipMemsetAsync(nf, 0, sizeof(int), st));
    if constexpr (std::is_same<T, double>::value) {
      copy_check_finite(st, R.buf[0], dcore, R.ofs[d], nf);
    } else {
      check_finite(st, dcore, R.ofs[d], nf);
      convert<float, double>(st, R.buf[0], dcore, R.ofs[d]);
    }
    int *hnf = reinterpret_cast<int *>(R.pinned());
    TTG_HIP(hipMemcpyAsync(hnf, nf, sizeof(int), hipMemcpyDeviceToHost, st));
    TTG_HIP(hipStreamSynchronize(st));
    if (*hnf) throw Error(TTG_NONFINITE, "input core contains NaN or Inf");
  } else {
    (void)gen_ranks;
  }
  R.cp.resize(d);
  for (int64_t i = 0; i < d; ++i) R.cp[i] = R.buf[0] + R.ofs[i];
  if (gen) gen(0, R.cp[0]);
  R.run(eps, rmax, gen);
  // finalise: nrm(1) = ||core_1||, balance, pack   (round.m:90-107)
  const int64_t s0 = R.r[0] * R.n[0] * R.r[1];
  fro_norm<T>(st, R.cp[0], s0, R.nrm + 0);
  std::vector<int64_t> htab(3 * (d + 1));
  int64_t o = 0;
  for (int64_t i = 0; i < d; ++i) {
    htab[i] = R.r[i] * R.n[i] * R.r[i + 1];
    htab[(d + 1) + i] = o;
    o += htab[i];
  }
  std::vector<void *> hsrc(d);
  for (int64_t i = 0; i < d; ++i) hsrc[i] = R.cp[i];
  {
    // the tables go through pinned host memory (no pageable staging)
    const size_t tb = htab.size() * sizeof(int64_t), sb = d * sizeof(void *);
    char *hp = reinterpret_cast<char *>(R.pinned(tb + sb));
    std::memcpy(hp, htab.data(), tb);
    std::memcpy(hp + tb, hsrc.data(), sb);
    TTG_HIP(hipMemcpyAsync(dtab, hp, tb, hipMemcpyHostToDevice, st));
    TTG_HIP(hipMemcpyAsync(dsrc, hp + tb, sb, hipMemcpyHostToDevice, st));
  }
  finalize_balance<T>(st, d, reinterpret_cast<T *const *>(dsrc), dtab, dtab + (d + 1), s0, R.nrm,
                      dcore_out, o);
  TTG_HIP(hipStreamSynchronize(st));
  for (int64_t i = 0; i <= d; ++i) r_out[i] = R.r[i];
  if (len_out) *len_out = o;
  if (info) {
    info->qr_fallbacks = R.fallbacks;
    info->jacobi_sweeps = R.sweeps_total;
    info->jacobi_maxsweeps = R.sweeps_max;
    info->qr_retries = R.retries;
    info->svd_skipped = R.svd_skipped;
    info->path = 0;
    // backward-stable sweeps: O(d u); the mixed form's FP32-Gram CholeskyQR2
    // bonds are accepted with Q orthogonal to 1e-10
    info->err_est = mixed ? 1e-10 : 4.0 * 1.2e-16 * (double)d;
    info->orth_est = mixed ? 1e-10 : 1.2e-16 * (double)d;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return R.noconv ? TTG_NO_CONVERGENCE : TTG_OK;
}

ttg_status round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                     const double *dcore, double eps, const int64_t *rmax, int prec,
                     int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                     ttg_info *info) {
  TTG_CHECK(prec == TTG_PREC_FP64 || prec == TTG_PREC_FP32, TTG_BAD_ARG, "unknown precision mode");
  validate_tt(d, n, r);
  TTG_CHECK(eps >= 0.0 && std::isfinite(eps), TTG_BAD_ARG, "eps must be finite and >= 0");
  if (rmax)
    for (int64_t i = 1; i < d; ++i) TTG_CHECK(rmax[i] >= 1, TTG_BAD_ARG, "rmax entries must be >= 1");
  if (prec == TTG_PREC_FP64) {
    ttg_status st = TTG_OK;
    if (round_small_persist(ctx, d, n, r, dcore, eps, rmax, r_out, dcore_out, cap, len_out, info, &st))
      return st;
    // plain form: when every bond is proven to keep all its singular values,
    // one exact right-to-left CholeskyQR2 sweep preconditioned by the Gram
    // factors (no left sweep, no SVD); otherwise the sweeps
    if (round_gram(ctx, d, n, r, dcore, eps, rmax, true, 0.0, r_out, dcore_out, cap, len_out, info)) return TTG_OK;
  } else {
    // mixed form: the Gram-chain round when every bond is proven to keep all
    // its singular values (rows orthogonal to 1e-6 at worst); otherwise the sweeps
    if (round_gram(ctx, d, n, r, dcore, eps, rmax, false, 1e-6, r_out, dcore_out, cap, len_out, info)) return TTG_OK;
  }
  return round_impl<double>(ctx, d, n, r, dcore, eps, rmax, r_out, dcore_out, cap, len_out, info,
                            nullptr, nullptr, prec == TTG_PREC_FP32);
}

ttg_status round_capacity(int64_t d, const int64_t *n, const int64_t *r, int64_t *cap) {
  validate_tt(d, n, r);
  std::vector<int64_t> rc = capped_ranks(d, n, r);
  *cap = packed_size(d, n, rc.data());
  return TTG_OK;
}

// ---- mtimes (unfused) and fused mtimes + round -----------------------------
void validate_mv(int64_t d, const int64_t *n, const int64_t *m, const int64_t *ra,
                 const int64_t *rx) {
  TTG_CHECK(d >= 1, TTG_BAD_ARG, "d must be >= 1");
  TTG_CHECK(n && m && ra && rx, TTG_BAD_ARG, "n, m and the ranks are required");
  for (int64_t i = 0; i < d; ++i) TTG_CHECK(n[i] >= 1 && m[i] >= 1, TTG_BAD_ARG, "mode sizes must be >= 1");
  for (int64_t i = 0; i <= d; ++i) TTG_CHECK(ra[i] >= 1 && rx[i] >= 1, TTG_BAD_ARG, "ranks must be >= 1");
}

ttg_status mtimes_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                      const int64_t *ra, const double *dacore, const int64_t *rx,
                      const double *dxcore, int64_t *ry, double *dycore, int64_t cap,
                      int64_t *len_out) {
  validate_mv(d, n, m, ra, rx);
  std::vector<int64_t> rp(d + 1);
  for (int64_t i = 0; i <= d; ++i) rp[i] = ra[i] * rx[i];
  const int64_t need = packed_size(d, n, rp.data());
  if (len_out) *len_out = need;
  if (cap < need) return TTG_NEED_SPACE;
  int64_t oa = 0, ox = 0, oy = 0;
  for (int64_t i = 0; i < d; ++i) {
    mtimes_core<double>(ctx->stream, ra[i], n[i], m[i], ra[i + 1], dacore + oa, rx[i], rx[i + 1],
                        dxcore + ox, dycore + oy);
    oa += ra[i] * n[i] * m[i] * ra[i + 1];
    ox += rx[i] * m[i] * rx[i + 1];
    oy += rp[i] * n[i] * rp[i + 1];
  }
  TTG_HIP(hipStreamSynchronize(ctx->stream));
  for (int64_t i = 0; i <= d; ++i) ry[i] = rp[i];
  return TTG_OK;
}

ttg_status mtimes_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                            const int64_t *ra, const double *dacore, const int64_t *rx,
                            const double *dxcore, double eps, const int64_t *rmax, int prec,
                            int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                            ttg_info *info) {
  validate_mv(d, n, m, ra, rx);
  std::vector<int64_t> rp(d + 1), oa(d + 1, 0), ox(d + 1, 0);
  for (int64_t i = 0; i <= d; ++i) rp[i] = ra[i] * rx[i];
  for (int64_t i = 0; i < d; ++i) {
    oa[i + 1] = oa[i] + ra[i] * n[i] * m[i] * ra[i + 1];
    ox[i + 1] = ox[i] + rx[i] * m[i] * rx[i + 1];
  }
  hipStream_t st = ctx->stream;
  std::vector<int64_t> nn(n, n + d), mm(m, m + d), raa(ra, ra + d + 1), rxx(rx, rx + d + 1);
  TTG_CHECK(prec == TTG_PREC_FP64 || prec == TTG_PREC_FP32, TTG_BAD_ARG, "unknown precision mode");
  // core 0 is materialised (it is the first QR's input); every later product
  // core only ever feeds sweep 1's R * core product and is built inside it
  std::function<void(int64_t, double *)> gen = [&](int64_t i, double *dst) {
    mtimes_core<double>(st, raa[i], nn[i], mm[i], raa[i + 1], dacore + oa[i], rxx[i], rxx[i + 1],
                        dxcore + ox[i], dst);
  };
  std::function<void(int64_t, const double *, int64_t, double *)> fmul =
      [&](int64_t i, const double *Rm, int64_t kk, double *dst) {
        gemm_mtimes(st, kk, Rm, kk, raa[i], nn[i], mm[i], raa[i + 1], dacore + oa[i], rxx[i], rxx[i + 1],
                    dxcore + ox[i], dst, kk);
      };
  return round_impl<double>(ctx, d, n, rp.data(), nullptr, eps, rmax, r_out, dcore_out, cap,
                            len_out, info, gen, &rp, prec == TTG_PREC_FP32, &fmul);
}

}  // namespace ttg
