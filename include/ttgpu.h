/*
 * ttgpu — C ABI of the MI355X (gfx950) TT rounding / TT-matvec library.
 *
 * Layout contract (identical to TT-Toolbox's tt_tensor, @tt_tensor/tt_tensor.m:33-41):
 *   a TT tensor is d, mode sizes n[d], ranks r[d+1] and ONE flat column-major
 *   core vector; core k is r[k] x n[k] x r[k+1] (r fastest) and starts at
 *   ps[k] = sum_{j<k} r[j]*n[j]*r[j+1].  A tt_matrix (@tt_matrix/tt_matrix.m:33-39)
 *   is the same with core k shaped r[k] x n[k] x m[k] x r[k+1].
 *
 * All sizes are int64.  Inputs are borrowed read-only, outputs are written
 * into caller memory; nothing allocated by the library crosses the ABI.
 * Functions without the _dev suffix take HOST pointers for core data; the
 * _dev variants take DEVICE pointers (the ctx's device) and never copy the
 * cores through the host.  Size/rank arrays are always host memory.
 *
 * Threading: one context per host thread; calls on one context are
 * serialised on its stream.  Every function returns a ttg_status; on error
 * ttg_last_error(ctx) holds a message.
 */
#ifndef TTGPU_H
#define TTGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  TTG_OK = 0,
  TTG_BAD_ARG = 1,        /* inconsistent sizes/ranks (cf. @tt_tensor/plus.m:50-59 checks) */
  TTG_NEED_SPACE = 2,     /* output capacity too small; *len_out holds the required length */
  TTG_NONFINITE = 3,      /* NaN/Inf met in the input or produced by a factorisation */
  TTG_NO_CONVERGENCE = 4, /* Jacobi sweep cap reached (analogue of the svd->svds warning, round.m:64-74) */
  TTG_HIP_ERROR = 5,      /* a HIP runtime call failed */
  TTG_NO_DEVICE = 6,      /* no usable gfx950 device */
  TTG_IO_ERROR = 7        /* a file cannot be opened, read or written (fmex/tt_read.c:29-34) */
} ttg_status;

typedef enum {
  TTG_PREC_FP64 = 0,      /* plain form: FP64 throughout */
  TTG_PREC_FP32 = 1       /* mixed form (error budget 1e-6 relative): the Gram-chain form when every
                             bond is proven not to truncate (FP64 Gram chains, batched r x r
                             Cholesky, one CholeskyQR pass, rows orthogonal to <= 1e-6); otherwise
                             the sweeps with FP32 MFMA Gram products in sweep 1's CholeskyQR2 and
                             FP64 storage, solves, contractions, r x r Cholesky and Jacobi SVD */
} ttg_precision;

typedef struct ttg_ctx ttg_ctx;

/* Per-call diagnostics (optional out-parameter, may be NULL). */
typedef struct {
  int64_t qr_fallbacks;     /* bonds where CholeskyQR was rejected and Householder QR ran */
  int64_t jacobi_sweeps;    /* total one-sided Jacobi sweeps over all bonds */
  int64_t jacobi_maxsweeps; /* largest sweep count of one SVD */
  double  seconds;          /* wall time of the call on the host clock */
  int64_t qr_retries;       /* bond QRs redone at a higher level (FP64 / Cholesky pass, Householder) */
  int64_t svd_skipped;      /* sweep-2 bonds proven not to truncate: completed from the QR, no SVD */
  int64_t path;             /* 0: QR/SVD sweeps; 1: Gram-chain form (every bond proven to keep all
                               its singular values: result from batched Gram factorisations) */
  double  err_est;          /* estimated relative Frobenius difference from the FP64 result of the
                               sweeps (Gram form: u * sum_i kappa_F(R_i); sweeps: u * d) */
  double  orth_est;         /* estimated max |C_k C_k^T / nrm0^2 - I| of the output cores */
} ttg_info;

/* Context: owns a HIP stream (or borrows `stream` when non-NULL), device workspace.
 * No reference counterpart: MATLAB holds this state implicitly (single interpreter thread). */
ttg_status ttg_ctx_create(int device, void *stream, ttg_ctx **out);
ttg_status ttg_ctx_destroy(ttg_ctx *ctx);
const char *ttg_last_error(const ttg_ctx *ctx);
void *ttg_ctx_stream(const ttg_ctx *ctx);
const char *ttg_version(void);

/* Required output capacity (doubles) for round: the QR-capped input size.
 * Replaces the storage bookkeeping of @tt_tensor/round.m:48,54-57. */
ttg_status ttg_round_capacity(int64_t d, const int64_t *n, const int64_t *r, int64_t *cap);

/* Y = round(X, eps, rmax)  —  replaces @tt_tensor/round.m:1-111 (with core/my_chop2.m:18-34).
 * rmax: NULL (no cap, round.m:22) or d+1 entries (rmax[i] caps r[i], round.m:78).
 * Output: r_out[d+1] and core_out (capacity `cap` doubles); *len_out = used length.
 * prec selects the plain FP64 form or the mixed FP32/FP64 form. */
ttg_status ttg_round(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                     const double *core, double eps, const int64_t *rmax, int prec,
                     int64_t *r_out, double *core_out, int64_t cap, int64_t *len_out,
                     ttg_info *info);
ttg_status ttg_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                         const double *dcore, double eps, const int64_t *rmax, int prec,
                         int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                         ttg_info *info);

/* round on a tt_matrix — replaces @tt_matrix/round.m:23-27 (rounds tt with modes n.*m). */
ttg_status ttg_round_matrix(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                            const int64_t *r, const double *core, double eps,
                            const int64_t *rmax, int prec, int64_t *r_out,
                            double *core_out, int64_t cap, int64_t *len_out, ttg_info *info);

/* y = A*x (tt_matrix times tt_tensor, no rounding) — replaces @tt_matrix/mtimes.m:38-65.
 * Output ranks ry = rA.*rx, core layout (rA rx) x n x (rA' rx') with rank index a + rA*alpha.
 * cap >= sum_k ry[k]*n[k]*ry[k+1]. */
ttg_status ttg_mtimes(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                      const int64_t *ra, const double *acore, const int64_t *rx,
                      const double *xcore, int64_t *ry, double *ycore, int64_t cap,
                      int64_t *len_out);
ttg_status ttg_mtimes_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                          const int64_t *ra, const double *dacore, const int64_t *rx,
                          const double *dxcore, int64_t *ry, double *dycore, int64_t cap,
                          int64_t *len_out);

/* y = round(A*x, eps, rmax) as one call — replaces the composition
 * @tt_matrix/mtimes.m:38-65 -> @tt_tensor/round.m.  When the product's ranks
 * fit the Gram-chain forms (d >= 2, ranks <= 480 after the caps, no binding
 * rmax) the product is formed once in the context's workspace and rounded by
 * them; otherwise the product cores are generated on the device inside the
 * rounding sweep's first products and never stored as a TT. */
ttg_status ttg_mtimes_round(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                            const int64_t *ra, const double *acore, const int64_t *rx,
                            const double *xcore, double eps, const int64_t *rmax, int prec,
                            int64_t *r_out, double *core_out, int64_t cap, int64_t *len_out,
                            ttg_info *info);
ttg_status ttg_mtimes_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                                const int64_t *ra, const double *dacore, const int64_t *rx,
                                const double *dxcore, double eps, const int64_t *rmax, int prec,
                                int64_t *r_out, double *dcore_out, int64_t cap,
                                int64_t *len_out, ttg_info *info);

/* Launch interface of round for repeated rounds of one shape (no reference
 * counterpart: MATLAB has no asynchronous calls).  It applies to tensors the
 * Gram-chain form handles (d >= 2, ranks <= 480 after the caps, no rmax), in
 * the mixed form (prec must be TTG_PREC_FP32; TTG_PREC_FP64 returns TTG_BAD_ARG).
 * ttg_round_gram_prepare rounds dcore once into dcore_out (it allocates,
 * uploads its tables, captures its launch sequence for these shapes,
 * pointers, eps and prec, and synchronises); *proven = 1 when every bond
 * kept its singular values (the result is the round), 0 otherwise (call
 * ttg_round_dev).  ttg_round_gram_launch then re-rounds whatever dcore holds
 * now into dcore_out on the context's stream with no allocation, no host
 * synchronisation and no host copies, and writes the device int *dstatus
 * (1: proven, the output is valid; 0: call ttg_round_dev).  Ranks are those
 * prepare returned.  ttg_round_gram_graph returns the captured graph (a
 * hipGraph_t, owned by the context, valid until the next prepare, a
 * Gram-form round of another shape on the context, or ttg_ctx_destroy) for a caller's
 * hipGraphAddChildGraphNode.  A Gram-form round of another shape on the same
 * context replaces the prepared one (launch then returns TTG_BAD_ARG). */
ttg_status ttg_round_gram_prepare(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                                  const double *dcore, double eps, int prec, int64_t *r_out,
                                  double *dcore_out, int64_t cap, int64_t *len_out, int *proven);
ttg_status ttg_round_gram_launch(ttg_ctx *ctx, int *dstatus);
ttg_status ttg_round_gram_graph(ttg_ctx *ctx, void **graph);

/* Batched rounding of `count` independent TT tensors (uneven d, n, r).
 * Tensor t: d[t], n/r at offsets nofs[t]/rofs[t] of the packed n/r arrays,
 * cores at device offset cofs[t] of dcore.  rmax: NULL (no cap) or packed like
 * r (rmax[rofs[t] + i] caps r_i of tensor t, round.m:78).  Outputs: r_out
 * packed like r, cores at device offset cofs_out[t] of dcore_out (capacity
 * from ttg_round_capacity per tensor).  Replaces a MATLAB loop of
 * round(X{t}, eps, rmax{t}) calls. */
ttg_status ttg_round_batched_dev(ttg_ctx *ctx, int64_t count, const int64_t *d,
                                 const int64_t *nofs, const int64_t *n, const int64_t *rofs,
                                 const int64_t *r, const int64_t *cofs, const double *dcore,
                                 double eps, const int64_t *rmax, int prec, int64_t *r_out,
                                 const int64_t *cofs_out, double *dcore_out, ttg_info *info);

/* Z = a*X + b*Y — replaces @tt_tensor/plus.m:41-86 (rank concatenation: X's cores in the
 * leading blocks, Y's in the trailing blocks) with the scalar products of
 * @tt_tensor/mtimes.m:10-27 (a and b scale core 1; minus.m:18 is a = 1, b = -1).
 * Boundary ranks must agree (plus.m:50-59, else TTG_BAD_ARG); rz = rx + ry inside.
 * cap >= sum_k rz[k]*n[k]*rz[k+1].  All cores are assembled by one launch.  Device pointers. */
ttg_status ttg_plus_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx,
                        const double *dx, double a, const int64_t *ry, const double *dy,
                        double b, int64_t *rz, double *dz, int64_t cap, int64_t *len_out);

/* W = round(a*X + b*Y, eps, rmax) fused — replaces the composition plus.m:41-86 ->
 * @tt_tensor/round.m.  When the sum can take the small-tensor or the Gram-chain form
 * its cores are assembled once on the device; otherwise the QR/SVD sweeps run with
 * the sum generated inside the first sweep's products (zero blocks never stored or
 * multiplied).  Output and capacity as ttg_round_dev for ranks rx + ry. */
ttg_status ttg_plus_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx,
                              const double *dx, double a, const int64_t *ry, const double *dy,
                              double b, double eps, const int64_t *rmax, int prec,
                              int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                              ttg_info *info);

/* ||X||_F — replaces @tt_tensor/norm.m:25-41 (left-to-right QR sweep, product of the
 * R-factor norms and the last core's norm).  Device pointer. */
ttg_status ttg_norm_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                        const double *dcore, double *nrm);

/* ||a*X + b*Y||_F — replaces norm(X - Y) (@tt_tensor/minus.m:18 then norm.m:25-41) without
 * storing the sum: the QR sweep runs over it with its cores generated inside the sweep's
 * products, so differences of nearly equal tensors keep their digits.  Device pointers. */
ttg_status ttg_plus_norm_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx,
                             const double *dx, double a, const int64_t *ry, const double *dy,
                             double b, double *nrm);

/* <X, Y> — replaces @tt_tensor/dot.m:92-111 (interface chain, boundary ranks 1;
 * other boundary ranks return TTG_BAD_ARG).  Device pointers. */
ttg_status ttg_dot_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx,
                       const double *dx, const int64_t *ry, const double *dy, double *value);

/* X = tt_tensor(full, eps) (TT-SVD) — replaces @tt_tensor/tt_tensor.m:175-244: left to
 * right, the unfolding of r_i n_i rows is factored by a QR (CholeskyQR2 under acceptance
 * tests, Householder otherwise) and the one-sided Jacobi SVD of its R factor; the rank is
 * my_chop2(s, eps/sqrt(d-1) * ||s||) (core/my_chop2.m), capped by rmax (NULL: no cap, as
 * the reference).  dfull: prod(n) doubles, column-major (first index fastest), device
 * memory.  Needs min(r_i n_i, prod(n)/(r_i n_i)) <= 512 at every step (else TTG_BAD_ARG).
 * d = 1 returns the vector as one core.  Output as ttg_round_dev (*len_out = used length;
 * TTG_NEED_SPACE when cap is short). */
ttg_status ttg_full_to_tt_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const double *dfull,
                              double eps, const int64_t *rmax, int64_t *r_out, double *dcore_out,
                              int64_t cap, int64_t *len_out, ttg_info *info);

/* SDV TT files — replace fmex/tt_read.c:13-103 (tt_tensor(filename),
 * @tt_tensor/tt_tensor.m:149-165) and fmex/tt_write.c:21-79 (@tt_tensor/tt_save.m:22).
 * Format: 128-byte header {"TT      ", ver {1,0}, inf {128, dtype, 0, 0}, comment[64],
 * i {1, 0, d, 0..}}, int32 l, m (d = m - l + 1), int32 n[d], int32 r[d+1], then the
 * flat core as double (dtype 0) or float (dtype 2; widened on reading).  Complex
 * files (dtype 1, 3) return TTG_BAD_ARG.  ttg_sdv_info reports d, the core length and
 * dtype; the readers need n[dcap >= d], r[dcap + 1] and cap >= core length (else
 * TTG_NEED_SPACE, with *d and *len_out set).  Writers: precision 0 double, 1 float.
 * The _dev forms stream the core between the file and device memory through two
 * pinned 16 MiB buffers (disk and PCIe overlap). */
ttg_status ttg_sdv_info(const char *path, int64_t *d, int64_t *core_len, int *dtype);
ttg_status ttg_sdv_read(const char *path, int64_t dcap, int64_t *d, int64_t *n, int64_t *r,
                        double *core, int64_t cap, int64_t *len_out);
ttg_status ttg_sdv_read_dev(ttg_ctx *ctx, const char *path, int64_t dcap, int64_t *d, int64_t *n,
                            int64_t *r, double *dcore, int64_t cap, int64_t *len_out);
ttg_status ttg_sdv_write(const char *path, int64_t d, const int64_t *n, const int64_t *r,
                         const double *core, int precision);
ttg_status ttg_sdv_write_dev(ttg_ctx *ctx, const char *path, int64_t d, const int64_t *n,
                             const int64_t *r, const double *dcore, int precision);

/* Adjoint of the sampled-entry completion loss 0.5*||P_Omega(X) - y||^2.
 * No reference counterpart (TT-Toolbox has no gradients); entry evaluation
 * follows @tt_tensor/subsref.m:116-122 and the interface products of
 * tt_regression/reg_als.m:22-39,66-92.  idx: S x d int32 (row-major, 0-based).
 * grad: laid out like core.  Device pointers; f32 and f64 forms. */
ttg_status ttg_adjoint_completion_f64(ttg_ctx *ctx, int64_t d, const int64_t *n,
                                      const int64_t *r, const double *dcore, int64_t S,
                                      const int32_t *didx, const double *dvals,
                                      double *loss, double *dgrad);
ttg_status ttg_adjoint_completion_f32(ttg_ctx *ctx, int64_t d, const int64_t *n,
                                      const int64_t *r, const float *dcore, int64_t S,
                                      const int32_t *didx, const float *dvals,
                                      double *loss, float *dgrad);

/* Adjoint of the bilinear form <y, A x>: value and gradients w.r.t. the cores
 * of x and y (either gradient pointer may be NULL).  Interface chain of
 * core/tt_mvdot.m:16-57.  Device pointers. */
ttg_status ttg_adjoint_bilinear_f64(ttg_ctx *ctx, int64_t d, const int64_t *n,
                                    const int64_t *m, const int64_t *ry, const double *dycore,
                                    const int64_t *ra, const double *dacore,
                                    const int64_t *rx, const double *dxcore, double *value,
                                    double *dgx, double *dgy);
ttg_status ttg_adjoint_bilinear_f32(ttg_ctx *ctx, int64_t d, const int64_t *n,
                                    const int64_t *m, const int64_t *ry, const float *dycore,
                                    const int64_t *ra, const float *dacore,
                                    const int64_t *rx, const float *dxcore, double *value,
                                    float *dgx, float *dgy);

/* ttg_adjoint_bilinear_f64 without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (the interface chains, 2 sum_k ry_k ra_k rx_k + 2 max_k(core
 * product) doubles) comes from a static device array of 8 MiB per device,
 * set up when the first context on the device is created; larger ranks
 * return TTG_BAD_ARG.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (on different streams or contexts) race.  Same results
 * as ttg_adjoint_bilinear_f64. */
ttg_status ttg_adjoint_bilinear_f64_noalloc(ttg_ctx *ctx, int64_t d, const int64_t *n,
                                            const int64_t *m, const int64_t *ry,
                                            const double *dycore, const int64_t *ra,
                                            const double *dacore, const int64_t *rx,
                                            const double *dxcore, double *value, double *dgx,
                                            double *dgy);

/* Kernel-class timing (diagnostic; no reference counterpart).  Between
 * begin and end, every launch of the chosen class on this context is
 * bracketed by HIP events on the context's stream; end synchronises and
 * returns the launch count, summed kernel time and the algorithmic flops and
 * bytes of those launches.  Classes: 1 GEMM, 2 Jacobi SVD, 3 Cholesky,
 * 4 triangular inverse, 5 Householder QR, 6 A*x product cores. */
ttg_status ttg_profile_begin(ttg_ctx *ctx, int kernel_class);
ttg_status ttg_profile_end(ttg_ctx *ctx, int64_t *launches, double *ms_total,
                           double *flops_total, double *bytes_total);

#ifdef __cplusplus
}
#endif
#endif /* TTGPU_H */
