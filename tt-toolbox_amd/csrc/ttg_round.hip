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
  bool sweep1_only = false;  // norm: the left QR sweep alone (norm.m:25-41)
  double *tsW = nullptr, *tsA = nullptr, *tsB = nullptr;  // TSQR scratch (TT-SVD)

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
  double *pinned(size_t bytes = 4096) { return static_cast<double *>(pinned_acquire(ctx, bytes)); }
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
  // bound (level 0 only): also start, on the side stream, ||R1||_F and two
  // upper bounds of ||R1^{-1}||_2 for the no-truncation test: ||R1^{-1}||_F
  // (loose by up to sqrt(k)) and ||(M^T M)^8||_F^(1/16), M = R1^{-1} (loose
  // by at most k^(1/16): (M^T M)^8 has the eigenvalues ||M||_2^16 >= ... and
  // its Frobenius norm lies between the largest and sqrt(k) times it), formed
  // by three squarings, each normalised by its Frobenius norm (scal_d[2..5]).
  // R = R2 R1 with ||R2 - I|| of order k max|E| <= 5e-6, so the bounds carry
  // over to R inside the test's margin.  Joined by the caller through
  // ctx->ev_join.
  // ||(M^T M)^8||_F^(1/16) for M = Rtmp (k x k): B = M^T M and three
  // squarings, each normalised by its Frobenius norm (scal_d[2..5]); R4 and
  // Rtmp are overwritten
  void inv_power_bound(hipStream_t s, int64_t k) {
    gemm<double>(s, true, false, k, k, k, 1.0, Rtmp, k, Rtmp, k, 0.0, R4, k, nullptr, 0, cus);  // M^T M
    double *Bp[2] = {R4, Rtmp};
    for (int j = 0; j < 3; ++j) {
      fro_norm<double>(s, Bp[j & 1], k * k, scal_d + 2 + j);
      scale_by_inv<double>(s, Bp[j & 1], k * k, scal_d + 2 + j);
      gemm<double>(s, false, false, k, k, k, 1.0, Bp[j & 1], k, Bp[j & 1], k, 0.0, Bp[(j + 1) & 1], k, nullptr, 0,
                   cus);
    }
    fro_norm<double>(s, Bp[1], k * k, scal_d + 5);
  }
  // the bounds of the no-truncation test from a finished R (k x k, FP64) on
  // the main stream: ||R||_F, ||R^{-1}||_F, the powered bound
  void inv_bounds(const double *Rf, int64_t k) {
    fro_norm<double>(st, Rf, k * k, scal_d + 0);
    eye_plus(st, k, R4, nullptr, 0.0);
    trsm_right_upper<double>(st, k, k, R4, k, Rf, Rtmp, k);  // Rtmp = R^{-1}
    fro_norm<double>(st, Rtmp, k * k, scal_d + 1);
    inv_power_bound(st, k);
  }
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
      inv_power_bound(ctx->side, k);
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
    // the test's bounds: at QR level 0 overlapped with the QR's second pass
    // (from R1, on the side stream), at the fallback levels from the final R.
    // Wide and tall bonds alike: the singular values of W are those of R.
    const bool bound = allow_notrunc;
    if (level < 2) {
      qr_tall_fast(tallA, mt, kj, Qs, R, fl, stats, level, false, bound && level == 0);
    } else {
      qr_householder(tallA, mt, kj, Qs, R);
    }
    if (bound && level > 0) inv_bounds(R, kj);
    JacobiPlan pl = jacobi_plan(kj);
    int h_fl[4] = {0, 0, 0, 0};
    double h_st[6] = {0, 0, 0, 0, 0, 0};
    skipped = false;
    // No-truncation shortcut: when 1/||R^{-1}||_2's upper bound
    // (qr_tall_fast, inv_bounds), a lower bound of the smallest singular
    // value, clears my_chop2's threshold, every singular value is kept and
    // the SVD would only change the gauge of the bond, so it is completed
    // without one (sweep2_bond): wide W = R^T Q^T gives core_i = Q^T
    // (orthonormal rows, as V' would have) and core_{i-1} <- core_{i-1} R^T;
    // tall W (rank N = n_i r_{i+1}, the right cap) gives core_i = I_N and
    // core_{i-1} <- core_{i-1} W.  Same tensor, same ranks.  The decision and
    // the QR acceptance are read in one synchronisation; a rejected QR is
    // redone before any SVD work.
    if (bound) {
      if (level == 0) TTG_HIP(hipStreamWaitEvent(st, ctx->ev_join, 0));
      notrunc_decide(st, scal_d + 0, scal_d + 1, scal_d + 2, ep, kj, rmax_i, skip);
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

  // ---- TT-SVD of a full tensor (@tt_tensor/tt_tensor.m:216-233) -----------
  // Step i: the unfolding c (m x M, m = r_i n_i), then the one-sided Jacobi
  // SVD of a k x k R factor, k = min(m, M), and my_chop2 at ep * ||s||
  // (sort_chop):
  //   wide (M >= m): c^T = Q R by TSQR (R only, Householder tree: stable on
  //     the numerically low-rank unfoldings TT-SVD is for), R J = Y:
  //     U = J(:, perm), S V^T = U^T c (one GEMM over c; Q is never formed)
  //   tall (M < m, small): c = Qs R (CholeskyQR2 under the acceptance tests,
  //     Householder where they fail), R^T J = Y: U = Qs J(:, perm),
  //     S V^T = Y(:, perm)^T
  // core_i = U (m x r_{i+1}); the next unfolding is S V^T seen as
  // (r_{i+1} n_{i+1}) x (M / n_{i+1}).  Cores go to `out`; returns its length.
  int64_t ttsvd(const double *full, int64_t N, const int64_t *nn, double eps, const int64_t *rmax,
                int64_t *rr, double *out) {
    // step 0 reads the caller's tensor; later steps alternate buf[0] / buf[1]
    const double *c = full;
    const double ep = eps / std::sqrt((double)(d - 1));
    rr[0] = 1;
    int64_t o = 0, len = N;  // len: entries of the current unfolding
    for (int64_t i = 0; i + 1 < d; ++i) {
      const int64_t m = rr[i] * nn[i], M = len / m;
      const int64_t kj = std::min(m, M);
      TTG_CHECK(kj <= kMaxRank, TTG_BAD_ARG, "TT-SVD unfoldings need min(rows, columns) <= 512");
      const bool wide = M >= m;
      const int64_t mt = m;
      const double *At = c;
      int level = wide ? 3 : 0;
      if (wide) tsqr_rt(st, c, m, M, tsW, tsA, tsB, Rprod);  // Rprod = R^T
      for (; level < 3;) {
        if (level >= 2) {
          qr_householder(At, mt, kj, Qs, Rprod);
          break;
        }
        qr_tall_fast(At, mt, kj, Qs, Rprod, iflag, bstats, level, false, false);
        double *hp = pinned();
        int *hf = reinterpret_cast<int *>(hp + 8);
        TTG_HIP(hipMemcpyAsync(hp, bstats, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        TTG_HIP(hipMemcpyAsync(hf, iflag, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        TTG_HIP(hipStreamSynchronize(st));
        double h_st[4];
        int h_fl[2];
        std::memcpy(h_st, hp, sizeof(h_st));
        std::memcpy(h_fl, hf, sizeof(h_fl));
        if (qr_accept(h_fl, h_st, level, false)) break;
        level = qr_next(h_fl, h_st, level, false);
        ++retries;
      }
      JacobiPlan pl = jacobi_plan(kj);
      // wide: S = R^T transposed (R J = Y); tall: S = R transposed (R^T J = Y)
      jacobi_svd(st, pl, Rprod, /*transS=*/true, Y, jsweeps, bar, jerr, jacobi_sweep_cap(), Yws, nullptr, 1e-16);
      sort_chop(st, Y, pl.rows, kj, pl.kpad, ep, rmax ? rmax[i + 1] : INT64_MAX, sv, perm, scal_i, scal_d + 8);
      int h_i[3];
      {
        int *hp = reinterpret_cast<int *>(pinned());
        TTG_HIP(hipMemcpyAsync(hp, scal_i, sizeof(int), hipMemcpyDeviceToHost, st));
        TTG_HIP(hipMemcpyAsync(hp + 1, jerr, sizeof(int), hipMemcpyDeviceToHost, st));
        TTG_HIP(hipMemcpyAsync(hp + 2, jsweeps, sizeof(int), hipMemcpyDeviceToHost, st));
        TTG_HIP(hipStreamSynchronize(st));
        std::memcpy(h_i, hp, sizeof(h_i));
      }
      if (h_i[1] == 2) throw Error(TTG_HIP_ERROR, "Jacobi grid barrier timed out");
      if (h_i[1] == 3) throw Error(TTG_HIP_ERROR, "Jacobi round barrier did not complete");
      if (h_i[1] == 1) noconv = true;
      sweeps_total += h_i[2];
      sweeps_max = std::max<int64_t>(sweeps_max, h_i[2]);
      const int64_t r1 = std::max(1, h_i[0]);
      double *cn = buf[i & 1];
      if (wide) {
        gather_cols<double>(st, Y, pl.rows, kj, kj, perm, r1, out + o, m);  // U = J(:, perm)
        g(true, false, r1, M, m, out + o, m, c, m, cn, r1);                 // S V^T = U^T c
      } else {
        gather_cols<double>(st, Y, pl.rows, kj, kj, perm, r1, Q1, kj);      // J(:, perm)
        g(false, false, m, r1, M, Qs, m, Q1, kj, out + o, m);                // U = Qs J_p
        gather_cols_t<double>(st, Y, pl.rows, 0, kj, perm, r1, cn);         // S V^T = Y_p^T
      }
      rr[i + 1] = r1;
      o += m * r1;
      len = r1 * M;
      c = cn;
    }
    // the last core is what remains (tt_tensor.m:234)
    TTG_HIP(hipMemcpyAsync(out + o, c, len * sizeof(double), hipMemcpyDeviceToDevice, st));
    rr[d] = 1;
    return o + len;
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
    } else if (skipped) {
      // tall, no truncation (rank N, the right cap): U S = W, core_i = I_N
      TTG_HIP(hipMemcpyAsync(US, cp[i], kw * N * sizeof(T), hipMemcpyDeviceToDevice, st));
      eye_plus(st, N, core_i, nullptr, 0.0);
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
    if (sweep1_only) return;
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
                             const std::function<void(int64_t, const T *, int64_t, T *)> *fmul = nullptr,
                             double *norm_out = nullptr) {
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
  const int64_t cap_needed = norm_out ? 0 : packed_size(d, n, rc.data());
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
  R.sweep1_only = norm_out != nullptr;
  R.run(eps, rmax, gen);
  if (norm_out) {
    // ||X|| = prod_i nrm(i) * ||last core||  (norm.m:38-40): the factors are
    // multiplied with their binary exponents split off, so the product
    // neither overflows nor loses digits on the way
    fro_norm<T>(st, R.cp[d - 1], R.r[d - 1] * R.n[d - 1] * R.r[d], R.nrm + 0);
    double *hn = R.pinned((d + 1) * sizeof(double));
    TTG_HIP(hipMemcpyAsync(hn, R.nrm, d * sizeof(double), hipMemcpyDeviceToHost, st));
    TTG_HIP(hipStreamSynchronize(st));
    double mant = 1.0;
    long ex = 0;
    for (int64_t i = 0; i < d; ++i) {
      int e = 0;
      mant = std::frexp(mant * hn[i], &e);
      ex += e;
    }
    *norm_out = mant == 0.0 ? 0.0 : std::ldexp(mant, (int)std::max<long>(std::min<long>(ex, 4096), -4096));
    if (info) {
      *info = ttg_info{};
      info->qr_fallbacks = R.fallbacks;
      info->qr_retries = R.retries;
      info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return TTG_OK;
  }
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
  // one launch for every core (the table goes through pinned memory)
  std::vector<MvCore> tab(d);
  int64_t oa = 0, ox = 0, oy = 0, max_units = 1;
  for (int64_t i = 0; i < d; ++i) {
    tab[i] = MvCore{dacore + oa, dxcore + ox, dycore + oy, ra[i], n[i], m[i], ra[i + 1], rx[i], rx[i + 1]};
    TTG_CHECK(rp[i] * n[i] * rp[i + 1] < (int64_t(1) << 31) && ra[i] * n[i] * m[i] * ra[i + 1] < (int64_t(1) << 31),
              TTG_BAD_ARG, "a core of A*x has 2^31 or more entries");
    max_units = std::max(max_units, mtimes_units(rp[i], rx[i + 1]));
    oa += ra[i] * n[i] * m[i] * ra[i + 1];
    ox += rx[i] * m[i] * rx[i + 1];
    oy += rp[i] * n[i] * rp[i + 1];
  }
  ctx->arena.reserve(d * sizeof(MvCore) + 4096);
  ctx->arena.reset();
  MvCore *dtab = ctx->arena.take<MvCore>(d);
  void *hp = pinned_acquire(ctx, d * sizeof(MvCore));
  std::memcpy(hp, tab.data(), d * sizeof(MvCore));
  TTG_HIP(hipMemcpyAsync(dtab, hp, d * sizeof(MvCore), hipMemcpyHostToDevice, ctx->stream));
  pinned_release(ctx, ctx->stream);
  mtimes_all(ctx->stream, dtab, (int)d, max_units, 8.0 * (double)(oa + ox + oy));
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
  // When the product can take the Gram-chain form (capped ranks up to 480,
  // no binding rmax), it is formed once by one k_mtimes_all launch and
  // rounded as a tensor: that form reads every core several times, which the
  // sweeps' on-chip generation of the product cores would have to repeat.
  // Otherwise the product cores are built inside the first sweep's products.
  if (gram_shape_ok(d, n, rp.data(), rmax)) {
    int64_t need = 0;
    round_capacity(d, n, rp.data(), &need);
    if (cap < need) {
      if (len_out) *len_out = need;
      return TTG_NEED_SPACE;
    }
    const int64_t plen = packed_size(d, n, rp.data());
    ctx->sarena.reserve(plen * sizeof(double) + 4096);
    ctx->sarena.reset();
    double *prod = ctx->sarena.take<double>(plen);
    std::vector<int64_t> ry(d + 1);
    int64_t used = 0;
    const ttg_status ms = mtimes_dev(ctx, d, n, m, ra, dacore, rx, dxcore, ry.data(), prod, plen, &used);
    TTG_CHECK(ms == TTG_OK, ms, "A*x staging for the round failed");
    return round_dev(ctx, d, n, rp.data(), prod, eps, rmax, prec, r_out, dcore_out, cap, len_out, info);
  }
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

// ---- sums: a X + b Y fused with the round; norms of tensors and differences
namespace {
// Rounding-sweep hooks for Z = a X + b Y (@tt_tensor/plus.m:41-86, scalars on
// core 1 as @tt_tensor/mtimes.m:10-27): core 1 of Z is assembled (it is the
// first QR's input); every later core only feeds sweep 1's R * core product,
// and R * Z_k splits by the block structure of Z_k into
//   [R(:, 0:rx_k) X_k, R(:, rx_k:rz_k) Y_k]    (column ranges of the result)
// for inner cores and R(:, 0:rx) X_k + R(:, rx:rz) Y_k for the last core, so
// the zero blocks are neither stored nor multiplied.
struct PlusHooks {
  ttg_ctx *ctx;
  int64_t d;
  const int64_t *n, *rx, *ry;
  const double *dx, *dy;
  double a, b;
  std::vector<int64_t> rz, ox, oy;
  std::function<void(int64_t, double *)> gen;
  std::function<void(int64_t, const double *, int64_t, double *)> fmul;

  PlusHooks(ttg_ctx *c, int64_t d_, const int64_t *n_, const int64_t *rx_, const double *dx_, double a_,
            const int64_t *ry_, const double *dy_, double b_)
      : ctx(c), d(d_), n(n_), rx(rx_), ry(ry_), dx(dx_), dy(dy_), a(a_), b(b_) {
    rz = plus_ranks(d, n, rx, ry);
    ox.assign(d + 1, 0);
    oy.assign(d + 1, 0);
    for (int64_t k = 0; k < d; ++k) {
      ox[k + 1] = ox[k] + rx[k] * n[k] * rx[k + 1];
      oy[k + 1] = oy[k] + ry[k] * n[k] * ry[k + 1];
    }
    gen = [this](int64_t k, double *dst) {
      PlusCore c{};
      c.x = dx + ox[k];
      c.y = dy + oy[k];
      c.z = dst;
      c.rx0 = rx[k];
      c.rx1 = rx[k + 1];
      c.ry0 = ry[k];
      c.ry1 = ry[k + 1];
      c.rz0 = rz[k];
      c.rz1 = rz[k + 1];
      c.n = n[k];
      c.a = k == 0 ? a : 1.0;
      c.b = k == 0 ? b : 1.0;
      c.sum = d == 1;
      plus_assemble_one(ctx->stream, c);
    };
    fmul = [this](int64_t k, const double *Rm, int64_t kk, double *dst) {
      hipStream_t st = ctx->stream;
      const int cus = ctx->num_cus;
      const int64_t nk = n[k];
      if (k + 1 < d) {
        gemm<double>(st, false, false, kk, nk * rx[k + 1], rx[k], 1.0, Rm, kk, dx + ox[k], rx[k], 0.0, dst, kk,
                     nullptr, 0, cus);
        gemm<double>(st, false, false, kk, nk * ry[k + 1], ry[k], 1.0, Rm + kk * rx[k], kk, dy + oy[k], ry[k],
                     0.0, dst + kk * nk * rx[k + 1], kk, nullptr, 0, cus);
      } else {
        gemm<double>(st, false, false, kk, nk * rz[d], rx[k], 1.0, Rm, kk, dx + ox[k], rx[k], 0.0, dst, kk,
                     nullptr, 0, cus);
        gemm<double>(st, false, false, kk, nk * rz[d], ry[k], 1.0, Rm + kk * rx[k], kk, dy + oy[k], ry[k], 1.0,
                     dst, kk, nullptr, 0, cus);
      }
    };
  }

  // NaN/Inf in either summand (the fused sweep never copies the input)
  void check_inputs() {
    hipStream_t st = ctx->stream;
    ctx->arena.reserve(4096);
    ctx->arena.reset();
    int *nf = ctx->arena.take<int>(16);
    TTG_HIP(hipMemsetAsync(nf, 0, sizeof(int), st));
    check_finite(st, dx, ox[d], nf);
    check_finite(st, dy, oy[d], nf);
    int hnf = 0;
    TTG_HIP(hipMemcpyAsync(&hnf, nf, sizeof(int), hipMemcpyDeviceToHost, st));
    TTG_HIP(hipStreamSynchronize(st));
    if (hnf) throw Error(TTG_NONFINITE, "input core contains NaN or Inf");
  }
};
}  // namespace

// ||X|| by the left QR sweep of @tt_tensor/norm.m:25-41 (FP64 CholeskyQR2 with
// the sweep's acceptance tests, Householder where they fail).
ttg_status norm_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r, const double *dcore,
                    double *nrm) {
  validate_tt(d, n, r);
  TTG_CHECK(dcore && nrm, TTG_BAD_ARG, "null pointer argument");
  std::vector<int64_t> ro(d + 1);
  return round_impl<double>(ctx, d, n, r, dcore, 0.0, nullptr, ro.data(), nullptr, INT64_MAX, nullptr, nullptr,
                            nullptr, nullptr, false, nullptr, nrm);
}

// ||a X + b Y|| (minus.m:18 with a = 1, b = -1, then norm.m:25-41): the QR
// sweep runs over the sum with the sum's cores generated inside sweep 1's
// products, so the difference of two nearly equal tensors keeps its digits
// (no cancellation of norms or inner products).
ttg_status plus_norm_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx,
                         double a, const int64_t *ry, const double *dy, double b, double *nrm) {
  TTG_CHECK(dx && dy && nrm, TTG_BAD_ARG, "null pointer argument");
  PlusHooks h(ctx, d, n, rx, dx, a, ry, dy, b);
  std::vector<int64_t> ro(d + 1);
  h.check_inputs();
  return round_impl<double>(ctx, d, n, h.rz.data(), nullptr, 0.0, nullptr, ro.data(), nullptr, INT64_MAX,
                            nullptr, nullptr, h.gen, &h.rz, false, &h.fmul, nrm);
}

// round(a X + b Y, eps, rmax) (plus.m:41-86 then round.m).  When the sum can
// take the small-tensor or the Gram-chain form (both read every input core
// more than once), its cores are assembled once into ctx->sarena (one launch,
// one HBM pass) and rounded like any tensor; otherwise the QR/SVD sweeps run
// with the sum generated inside sweep 1's products (never assembled).
ttg_status plus_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx,
                          double a, const int64_t *ry, const double *dy, double b, double eps,
                          const int64_t *rmax, int prec, int64_t *r_out, double *dcore_out, int64_t cap,
                          int64_t *len_out, ttg_info *info) {
  TTG_CHECK(dx && dy && r_out && dcore_out, TTG_BAD_ARG, "null pointer argument");
  TTG_CHECK(prec == TTG_PREC_FP64 || prec == TTG_PREC_FP32, TTG_BAD_ARG, "unknown precision mode");
  TTG_CHECK(eps >= 0.0 && std::isfinite(eps), TTG_BAD_ARG, "eps must be finite and >= 0");
  PlusHooks h(ctx, d, n, rx, dx, a, ry, dy, b);
  const int64_t *rz = h.rz.data();
  if (rmax)
    for (int64_t i = 1; i < d; ++i) TTG_CHECK(rmax[i] >= 1, TTG_BAD_ARG, "rmax entries must be >= 1");
  int64_t need = 0;
  round_capacity(d, n, rz, &need);
  if (cap < need) {
    if (len_out) *len_out = need;
    return TTG_NEED_SPACE;
  }
  bool small = false;
  if (prec == TTG_PREC_FP64) {
    const char *env = std::getenv("TTG_SMALL_PERSIST");
    std::vector<int64_t> rc = capped_ranks(d, n, rz);
    small = !(env && env[0] == '0') && persist_supported(d, n, rz);
    for (int64_t i = 0; i <= d && small; ++i) small = rc[i] <= 64;
    for (int64_t i = 1; i < d && small && rmax; ++i) small = rmax[i] >= rc[i];
  }
  if (small || gram_shape_ok(d, n, rz, rmax)) {
    int64_t zlen = 0;
    for (int64_t k = 0; k < d; ++k) zlen += rz[k] * n[k] * rz[k + 1];
    ctx->sarena.reserve(zlen * sizeof(double) + d * sizeof(PlusCore) + 4096);
    ctx->sarena.reset();
    double *z = ctx->sarena.take<double>(zlen);
    PlusCore *tab = ctx->sarena.take<PlusCore>(d);
    std::vector<double *> zdst(d);
    for (int64_t k = 0, o = 0; k < d; ++k) {
      zdst[k] = z + o;
      o += rz[k] * n[k] * rz[k + 1];
    }
    const int64_t mx = plus_table(ctx, d, n, rx, dx, a, ry, dy, b, h.rz, zdst.data(), 0, d, tab);
    plus_assemble(ctx->stream, tab, (int)d, mx);
    // the table's upload read the pinned buffer, which the round reuses
    TTG_HIP(hipStreamSynchronize(ctx->stream));
    return round_dev(ctx, d, n, rz, z, eps, rmax, prec, r_out, dcore_out, cap, len_out, info);
  }
  h.check_inputs();
  return round_impl<double>(ctx, d, n, rz, nullptr, eps, rmax, r_out, dcore_out, cap, len_out, info, h.gen,
                            &h.rz, prec == TTG_PREC_FP32, &h.fmul);
}

// TT-SVD: tt_tensor(full, eps) of @tt_tensor/tt_tensor.m:175-244, left to
// right, ep = eps / sqrt(d - 1), r_{i+1} = my_chop2(s, ep * ||s||), capped by
// rmax (the reference's constructor has no cap: rmax = NULL matches it).
ttg_status full_to_tt_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const double *dfull, double eps,
                          const int64_t *rmax, int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                          ttg_info *info) {
  auto t0 = std::chrono::steady_clock::now();
  TTG_CHECK(d >= 1 && n && dfull && r_out && dcore_out, TTG_BAD_ARG, "null or empty argument");
  TTG_CHECK(eps >= 0.0 && std::isfinite(eps), TTG_BAD_ARG, "eps must be finite and >= 0");
  int64_t N = 1;
  for (int64_t i = 0; i < d; ++i) {
    TTG_CHECK(n[i] >= 1, TTG_BAD_ARG, "mode sizes must be >= 1");
    TTG_CHECK(N <= (int64_t(1) << 40) / n[i], TTG_BAD_ARG, "full tensor above 2^40 entries");
    N *= n[i];
  }
  if (rmax)
    for (int64_t i = 1; i < d; ++i) TTG_CHECK(rmax[i] >= 1, TTG_BAD_ARG, "rmax entries must be >= 1");
  hipStream_t st = ctx->stream;
  // rank bounds: r_i <= min(prod n[<i], prod n[>=i]) (and rmax): output size bound
  std::vector<int64_t> rb(d + 1, 1);
  {
    int64_t left = 1;
    for (int64_t i = 1; i < d; ++i) {
      left = std::min<int64_t>(left * n[i - 1], int64_t(1) << 40);
      int64_t right = 1;
      for (int64_t j = i; j < d && right < left; ++j) right = std::min<int64_t>(right * n[j], int64_t(1) << 40);
      rb[i] = std::min(left, right);
      if (rmax) rb[i] = std::min(rb[i], rmax[i]);
    }
  }
  int64_t outb = 0, kmax = 1;
  for (int64_t i = 0; i < d; ++i) {
    outb += rb[i] * n[i] * rb[i + 1];
    if (i + 1 < d) kmax = std::max(kmax, std::min<int64_t>(rb[i] * n[i], N / std::max<int64_t>(rb[i] * n[i], 1)));
  }
  kmax = std::max<int64_t>(1, std::min<int64_t>(kmax, kMaxRank));
  const int64_t total = std::max(N, outb);
  if (d == 1) {  // one core: the tensor itself (tt_tensor.m:199-211)
    if (len_out) *len_out = N;
    if (cap < N) return TTG_NEED_SPACE;
    ctx->arena.reserve(4096);
    ctx->arena.reset();
    int *nf = ctx->arena.take<int>(16);
    TTG_HIP(hipMemsetAsync(nf, 0, sizeof(int), st));
    copy_check_finite(st, dcore_out, dfull, N, nf);
    int hnf = 0;
    TTG_HIP(hipMemcpyAsync(&hnf, nf, sizeof(int), hipMemcpyDeviceToHost, st));
    TTG_HIP(hipStreamSynchronize(st));
    if (hnf) throw Error(TTG_NONFINITE, "input contains NaN or Inf");
    r_out[0] = r_out[1] = 1;
    if (info) *info = ttg_info{};
    return TTG_OK;
  }
  Rounder<double> R;
  R.ctx = ctx;
  R.st = st;
  R.cus = ctx->num_cus;
  R.d = d;
  R.mixed = false;
  R.max_core = N;
  R.kmax = kmax;
  R.rsq = kmax * kmax;
  if (!ctx->side) {
    TTG_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    TTG_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    TTG_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
  }
  // TSQR scratch for the largest wide unfolding (m <= kMaxRank rows)
  // (blocks of B = max(256, 2m) <= 1024 columns: W <= (M + B) m <= N + 1024 m,
  // one level's R stack <= (M / B + 1) m^2 <= N / 2 + m^2)
  const int64_t tsw = N + 1024 * kMaxRank, tsl = N / 2 + kMaxRank * kMaxRank;
  ArenaSizer z;
  Rounder<double>::layout(z, d, total, N, kmax, R.rsq);
  z.take<int>(16);
  z.take<double>(tsw);
  z.take<double>(tsl);
  z.take<double>(tsl);
  ctx->arena.reserve(z.used + 4096);
  ctx->arena.reset();
  R.carve(ctx->arena, total);
  int *nf = ctx->arena.take<int>(16);
  R.tsW = ctx->arena.take<double>(tsw);
  R.tsA = ctx->arena.take<double>(tsl);
  R.tsB = ctx->arena.take<double>(tsl);
  TTG_HIP(hipMemsetAsync(nf, 0, sizeof(int), st));
  check_finite(st, dfull, N, nf);
  {
    int *hnf = reinterpret_cast<int *>(R.pinned());
    TTG_HIP(hipMemcpyAsync(hnf, nf, sizeof(int), hipMemcpyDeviceToHost, st));
    TTG_HIP(hipStreamSynchronize(st));
    if (*hnf) throw Error(TTG_NONFINITE, "input contains NaN or Inf");
  }
  std::vector<int64_t> rr(d + 1, 1);
  const int64_t len = R.ttsvd(dfull, N, n, eps, rmax, rr.data(), R.buf[2]);
  if (len_out) *len_out = len;
  if (cap < len) return TTG_NEED_SPACE;
  TTG_HIP(hipMemcpyAsync(dcore_out, R.buf[2], len * sizeof(double), hipMemcpyDeviceToDevice, st));
  TTG_HIP(hipStreamSynchronize(st));
  for (int64_t i = 0; i <= d; ++i) r_out[i] = rr[i];
  if (info) {
    *info = ttg_info{};
    info->qr_fallbacks = R.fallbacks;
    info->qr_retries = R.retries;
    info->jacobi_sweeps = R.sweeps_total;
    info->jacobi_maxsweeps = R.sweeps_max;
    info->err_est = 4.0 * 1.2e-16 * (double)d;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return R.noconv ? TTG_NO_CONVERGENCE : TTG_OK;
}

}  // namespace ttg
