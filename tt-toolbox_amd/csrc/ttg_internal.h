// Internal declarations shared by the ttgpu HIP sources (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include <stdexcept>
#include "../../include/ttgpu.h"

namespace ttg {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  ttg_status status;
  Error(ttg_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

#define TTG_HIP(call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw ::ttg::Error(TTG_HIP_ERROR, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TTG_CHECK(cond, st, msg)                                                    \
  do {                                                                              \
    if (!(cond)) throw ::ttg::Error(st, msg);                                       \
  } while (0)

// ---------------------------------------------------------------------------
// device workspace: one growable arena, carved per call
// ---------------------------------------------------------------------------
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  void reserve(size_t bytes) {
    if (bytes <= cap) return;
    if (base) TTG_HIP(hipFree(base));
    base = nullptr;
    cap = 0;
    TTG_HIP(hipMalloc(&base, bytes));
    cap = bytes;
  }
  void reset() { used = 0; }
  template <class T> T *take(size_t count) {
    size_t bytes = ((count * sizeof(T)) + 255) & ~size_t(255);
    if (used + bytes > cap) throw Error(TTG_HIP_ERROR, "workspace arena exhausted");
    T *p = reinterpret_cast<T *>(base + used);
    used += bytes;
    return p;
  }
  ~Arena() {
    if (base) (void)hipFree(base);
  }
};

// Bump allocation inside memory the caller owns (no allocation, no free).
struct ArenaView {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  template <class T> T *take(size_t count) {
    size_t bytes = ((count * sizeof(T)) + 255) & ~size_t(255);
    if (used + bytes > cap) throw Error(TTG_BAD_ARG, "workspace too small");
    T *p = reinterpret_cast<T *>(base + used);
    used += bytes;
    return p;
  }
};

// Size-only twin of Arena used to compute the workspace requirement.
struct ArenaSizer {
  size_t used = 0;
  template <class T> void take(size_t count) { used += ((count * sizeof(T)) + 255) & ~size_t(255); }
};

// Kernel-class timing of one class.  Two ways: an event pair recorded on the
// stream around a scope (begin/end: the time includes the dispatch of the
// scope's kernels), or, for scopes built with `own`, a pair per kernel
// passed to hipExtLaunchKernel, which stamps the kernel's own start and end
// (the execution time a kernel trace reports).
enum KernelClass { KC_NONE = 0, KC_GEMM = 1, KC_JACOBI = 2, KC_CHOL = 3, KC_TRTRI = 4, KC_HQR = 5, KC_MTIMES = 6 };
struct Profiler {
  int kind = KC_NONE;
  std::vector<hipEvent_t> pool;  // pairs
  size_t used = 0;               // pairs in use
  std::vector<double> flops, bytes;
  size_t take_pair();  // a fresh pair (flops and bytes 0); returns its index
  void begin(hipStream_t s);
  void end(hipStream_t s, double fl, double by);
};
// the profiler of the context whose call is running on this thread (or null)
extern thread_local Profiler *g_prof;
struct ProfScope {
  Profiler *p;
  hipStream_t s;
  double fl, by;
  bool own;
  size_t first = ~size_t(0);
  ProfScope(int kind, hipStream_t st, double f, double b, bool own_events = false)
      : p(nullptr), s(st), fl(f), by(b), own(own_events) {
    if (g_prof && g_prof->kind == kind) {
      p = g_prof;
      if (!own) p->begin(s);
    }
  }
  // events for the scope's next kernel (own mode); the scope's flops and
  // bytes go to its first kernel
  void events(hipEvent_t *a, hipEvent_t *b) {
    const size_t i = p->take_pair();
    if (first == ~size_t(0)) first = i;
    *a = p->pool[2 * i];
    *b = p->pool[2 * i + 1];
  }
  ~ProfScope() {
    if (!p) return;
    if (!own) {
      p->end(s, fl, by);
    } else if (first != ~size_t(0)) {
      p->flops[first] += fl;
      p->bytes[first] += by;
    }
  }
};
// Launch inside an own-mode ProfScope: with the kernel's own timestamps when
// the scope is profiling, a plain launch otherwise.
#define TTG_TIMED_LAUNCH(ps, K, grid, block, shm, st, ...)                      \
  do {                                                                          \
    if ((ps).p) {                                                               \
      hipEvent_t e0_, e1_;                                                      \
      (ps).events(&e0_, &e1_);                                                  \
      hipExtLaunchKernelGGL(K, grid, block, shm, st, e0_, e1_, 0, __VA_ARGS__); \
    } else {                                                                    \
      hipLaunchKernelGGL(K, grid, block, shm, st, __VA_ARGS__);                 \
    }                                                                           \
  } while (0)

}  // namespace ttg

constexpr int kGramParts = 7;  // captured parts of the Gram-chain round

struct ttg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int num_cus = 256;
  ttg::Arena arena;       // per-call scratch
  ttg::Arena persist;     // small persistent device scalars (barrier counters, flags)
  void *host_pinned = nullptr;  // staging buffer: take it with ttg::pinned_acquire
  size_t host_pinned_cap = 0;
  hipEvent_t pinned_ev = nullptr;  // recorded after the last copy enqueued out of it
  bool pinned_pending = false;
  std::string last_error;
  // diagnostics of the current call
  int64_t qr_fallbacks = 0, jacobi_sweeps = 0, jacobi_maxsweeps = 0;
  ttg::Profiler prof;
  // batched rounding: worker contexts (own stream and workspace each),
  // created on first use and owned by this context
  std::vector<ttg_ctx *> workers;
  int batch_workers = 0;  // 0: chosen by the tuner on first use
  // persistent batched rounding: descriptors + slot workspace, the captured
  // launch (reset + persistent kernel) with the arguments it was captured for,
  // and the tuned launch shape (0: not tuned yet)
  ttg::Arena parena;
  // side stream + events (default priority) for work that overlaps the main chain
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraphExec_t pgraph = nullptr;
  std::vector<uint64_t> pgraph_key;
  int batch_mode = 0;
  // Gram-chain round: its workspace (job tables included) and its captured
  // launch sequence with the shapes and pointers it was captured for, in
  // seven parts (set-up; right chain | left chain; factorisations; products
  // or exact sweep | inverses and statistics; verdict and output) whose
  // branches replay on two streams at once: a graph's own branches replay
  // one after the other on one queue
  ttg::Arena garena;
  hipGraphExec_t ggraph[kGramParts] = {};
  std::vector<uint64_t> gkey;
  hipGraph_t ggraph_src = nullptr;  // the whole sequence as one graph (child-node form)
  int *gok = nullptr;               // its proof flag (device)
  uint64_t gruns = 0;               // replays so far
  std::vector<uint64_t> gprep;      // key of the prepared launch (ttg_round_gram_prepare)
  // cores of an assembled sum a X + b Y (plus fused with round), kept apart
  // from the rounding workspaces that read them
  ttg::Arena sarena;
  // a batched call's stream worker: runs on its own host thread beside other
  // workers, whose workspace growth (hipMalloc / hipFree / hipHostMalloc /
  // hipHostFree, which synchronise the device) must not meet a stream
  // capture in progress, so a worker launches eagerly and never captures
  bool worker = false;
};

namespace ttg {

// releases the batched-rounding worker contexts of ctx
void destroy_workers(ttg_ctx *ctx);
// The context's pinned host staging buffer, at least `bytes`: first waits for
// the asynchronous copies last enqueued out of (or into) it, so the caller may
// overwrite it; grows it when short.
void *pinned_acquire(ttg_ctx *ctx, size_t bytes);
// marks the copies just enqueued on s as the buffer's users
void pinned_release(ttg_ctx *ctx, hipStream_t s);
// releases every HIP object of ctx (workers, graph, streams, events, pinned
// memory) in dependency order; the arenas go with `delete ctx`
void release_ctx(ttg_ctx *ctx);
// Jacobi sweep cap (60; TTG_JACOBI_MAX_SWEEPS=1..60 lowers it to force NO_CONVERGENCE)
int jacobi_sweep_cap();

// one tensor of a persistent batched rounding (device copy)
struct PTensorDesc {
  int32_t d, pad;
  int64_t nofs, rofs;     // into the flat n and r arrays
  int64_t in_ofs, out_ofs;  // first input core, first output/working core
};
// round_dev's one-workgroup path for small FP64 tensors (false: not eligible)
bool round_small_persist(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r, const double *dcore,
                         double eps, const int64_t *rmax, int64_t *r_out, double *dcore_out, int64_t cap,
                         int64_t *len_out, ttg_info *info, ttg_status *status);
// Gram-chain round for tensors whose bonds provably keep every singular value
// (false: not applicable or not proven; the caller runs the sweeps)
bool round_gram(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r, const double *dcore, double eps,
                const int64_t *rmax, bool exact, double orth_tol, int64_t *r_out, double *dcore_out, int64_t cap,
                int64_t *len_out, ttg_info *info);
// the Gram-chain round's captured parts: launch on ctx's stream (chains
// forked to the side stream and joined back), release
void gram_graph_launch(ttg_ctx *ctx);
void gram_graph_release(ttg_ctx *ctx);
// the shape tests of round_gram (d >= 2, capped ranks <= 256, no binding
// rmax, the form not switched off): true when the form may be tried
bool gram_shape_ok(int64_t d, const int64_t *n, const int64_t *r, const int64_t *rmax);
// launch interface of the Gram-chain round (prepare allocates and synchronises;
// launch only enqueues the captured graph and the proof-flag copy)
ttg_status round_gram_prepare(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r, const double *dcore,
                              double eps, int prec, int64_t *r_out, double *dcore_out, int64_t cap,
                              int64_t *len_out, int *proven);
ttg_status round_gram_launch(ttg_ctx *ctx, int *dstatus);
ttg_status round_gram_graph(ttg_ctx *ctx, void **graph);
// true if one workgroup can round the tensor (d <= 64, ranks <= 512)
bool persist_supported(int64_t d, const int64_t *n, const int64_t *r);
// doubles of per-slot workspace for the largest core / bond rank of a batch
int64_t persist_slot_elems(int64_t slot_core, int64_t slot_kmax);
// ycap: doubles of LDS for the Jacobi matrix [S; J] (larger ones stay in the slot)
void persist_prepare(int nt, int64_t ycap);
// reset of the work counter + the persistent kernel on `slots` workgroups of nt threads
void persist_launch(hipStream_t s, int nt, int slots, int count, const PTensorDesc *tens,
                    const int *order, const int64_t *nall, const int64_t *rall, const double *in,
                    double *out, int64_t *r_out, double *ws, int64_t slot_core, int64_t slot_kmax, int *ctr,
                    int *tstat, double eps, int max_sweeps, int64_t ycap);

// ---------------------------------------------------------------------------
// kernel launchers (implemented in ttg_kernels.hip)
// ---------------------------------------------------------------------------
// Column-major GEMM C = alpha*op(A)*op(B) + beta*C.  ws (may be null) is used
// for split-K partials; ws_elems is its capacity in T elements.
template <class T>
void gemm(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K, T alpha,
          const T *A, int64_t lda, const T *B, int64_t ldb, T beta, T *C, int64_t ldc,
          T *ws, int64_t ws_elems, int num_cus);
// bytes of split-K workspace gemm() may use for the given shape
int64_t gemm_ws_elems(int64_t M, int64_t N, int64_t K, int num_cus);
// FP64 C (M x M) = op(A) op(B) for a product known to be symmetric (a Gram
// matrix): with split-K only the blocks on or above the diagonal are formed
// and the reduction mirrors them (C exactly symmetric); otherwise as gemm
void gemm_sym(hipStream_t s, bool ta, bool tb, int64_t M, int64_t K, const double *A, int64_t lda, const double *B,
              int64_t ldb, double *C, int64_t ldc, double *ws, int64_t wse, int cus);
// FP64 C = (*alpha_dev) op(A) op(B) (alpha_dev: device scalar)
void gemm_scaled(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K, const double *alpha_dev,
                 const double *A, int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc, double *ws,
                 int64_t wse, int cus);
// One problem of a batched FP64 GEMM C = alpha (*alpha_dev) op(A) op(B) (alpha_dev may be null).
struct BGemmJob {
  const void *A, *B;
  void *C;
  const double *alpha_dev;
  double alpha;
  int M, N, K, lda, ldb, ldc;
  int tile0;  // first grid block of the job (set by bgemm_f64)
  // triangular operands whose vanishing part is exactly zero in memory: the
  // K tiles it covers are skipped (the sum is bitwise the full product's).
  // 1: op(B) lower (block column n0 needs K >= n0); 2: op(A) lower (block row
  // m0 needs K < m0 + BM); 4: op(B) upper (K < n0 + BN); 8: op(A) upper (K >= m0)
  int tri;
};
// All jobs in one launch (same op(A), op(B)).  bgemm_plan fills the jobs'
// first blocks on the host and returns the grid size; the table is then
// copied to device memory (jobs_dev) by the caller and bgemm_run launches it.
// shape: 0 = 32 x 64 blocks, 1 = 32 x 32, 2 = 64 x 64 (force >= 0 imposes one)
int bgemm_plan(bool tb, BGemmJob *jobs, int njobs, int *shape, double *flops, int force = -1);
void bgemm_run(hipStream_t st, bool ta, bool tb, const BGemmJob *jobs_dev, int njobs, int total, int shape,
               double flops);

// y[i] = (T)x[i]
template <class TO, class TI> void convert(hipStream_t s, TO *y, const TI *x, int64_t count);
// B (cols x rows) = A^T where A is rows x cols, both column-major
template <class T> void transpose(hipStream_t s, const T *A, int64_t rows, int64_t cols, T *B);
// Frobenius norm of `count` contiguous values -> *out (device), fixed-order reduction
template <class T> void fro_norm(hipStream_t s, const T *x, int64_t count, double *out);
// x[i] *= 1/ *nrm if *nrm != 0 (nrm device scalar)
template <class T> void scale_by_inv(hipStream_t s, T *x, int64_t count, const double *nrm);

// Householder QR (LAPACK dgeqrf+dorgqr semantics): A (m x n, lda=m) -> Q
// (m x kq, explicit) and R (kq x n), kq = min(m,n).  A is overwritten.  With a
// workspace of householder_ws_elems doubles it runs blocked (32-column panels
// on one workgroup, trailing and Q updates as multi-CU GEMMs); without one,
// or for kq <= 32, unblocked on one workgroup.
size_t householder_ws_elems(int64_t m, int64_t n, int num_cus);
void householder_qr(hipStream_t s, int64_t m, int64_t n, double *A, double *Q, double *R,
                    double *tau, double *ws = nullptr, int num_cus = 256);

// One-workgroup Cholesky G + shift = R^T R (upper R), FP64, k x k (k <= 512).
// G is read from its upper triangle; shift = shift_rel * trace(G) / k.
// flag[0] = 0 ok, j+1 if pivot j failed; stats[0] = diag spread, stats[1] = max|R - I|.
// Wout (optional, (k/32 + 1) * 33 * 32 doubles): the inverses R_JJ^{-T} of the
// diagonal blocks for trsm_right_upper; fro (optional): ||R||_F.  Returns
// whether the factorisation wrote them (k <= 256); otherwise they are untouched.
bool chol_factor(hipStream_t s, int64_t k, const double *G, int64_t ldg, double shift_rel,
                 double *R, int *flag, double *stats, unsigned long long *dbg = nullptr,
                 double *Wout = nullptr, double *fro = nullptr);
// Batched one-workgroup Cholesky factorisations (k <= 256, no shift): job b
// factorises G (ld ldg) into R (ld k), flags, statistics, the inverses of
// R's diagonal blocks (Wout, for trsm3_batched) and ||R||_F (fro, optional).
struct CholJob {
  const double *G;
  double *R;
  int *flag;
  double *stats;
  double *Wout;
  double *fro;
  int k, ldg;
};
void chol_batched(hipStream_t s, const CholJob *jobs_dev, int njobs, int64_t kmax);
// the same for 256 < k <= 480 (no block inverses, no ||R||_F)
void chol4_batched(hipStream_t s, const CholJob *jobs_dev, int njobs, int64_t kmax);
// Batched X = A R^{-1} (A == nullptr: A = I) with the block inverses a
// batched Cholesky left (k <= 256 for every job).
struct TrsmJob {
  const double *A;
  const double *R;
  const double *Winv;
  double *X;
  int m, k, lda, ldx;
};
void trsm3_batched(hipStream_t s, const TrsmJob *jobs_dev, int njobs, int64_t mmax, int64_t kmax);
// Batched X = A R^{-1} through the generic blocked substitution (every job's A
// given, Winv unused; for k above 256)
void trsm_batched(hipStream_t s, const TrsmJob *jobs_dev, int njobs, int64_t mmax, int64_t kmax);
// X (m x k) = A R^{-1}, R upper triangular (FP64), blocked substitution.
// Gram of A (m x k, FP32 MFMA when g32) fused with the near-identity second
// CholeskyQR pass: X = I - Phi(E), R2 = I + Phi(E), *maxabs = max|E| (G is
// scratch for the unsplit case; ws holds split-K partials)
void gram_nearid(hipStream_t st, int64_t m, int64_t k, const double *A, bool g32, double *G, double *X,
                 double *R2, double *maxabs, double *ws, int64_t ws_elems, int cus);
// FP32 operands / FP64 operands rounded to FP32 while staged; FP32 MFMA; FP64 result
void gemm_f32_f64out(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K,
                     const float *A, int64_t lda, const float *B, int64_t ldb, double *C,
                     int64_t ldc, float *ws, int64_t wse, int cus);
void gemm_f64_via_f32(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K,
                      const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
                      int64_t ldc, float *ws, int64_t wse, int cus);
template <class T>
void trsm_right_upper(hipStream_t s, int64_t m, int64_t k, const T *A, int64_t lda, const double *R,
                      T *X, int64_t ldx, unsigned long long *dbg = nullptr,
                      const double *Winv = nullptr);

// One-sided block Jacobi SVD on the stacked matrix Y = [S; I] (FP64).
// Setup from S (k x k, column-major, optionally transposed) then rotate
// until converged.  Returns sweeps via *sweeps (device int).
struct JacobiPlan {
  int64_t k = 0;        // matrix order
  int64_t kpad = 0;     // padded column count (2*b*p)
  int b = 0, p = 0;     // block size and number of workgroups
  int64_t rows = 0;     // stacked rows = 2*k
};
JacobiPlan jacobi_plan(int64_t k);
size_t jacobi_ws_bytes(const JacobiPlan &pl);
// ws: jacobi_polar_ws_elems(pl) doubles of scratch for the final polar correction.
// dbg (optional): 4 timestamps (100 MHz clock) per outer step of the first sweep.
// stop_cos2: the iteration ends after a sweep none of whose rotations had a
// pair cosine^2 above it; the columns are then orthogonal to O(stop_cos2).
size_t jacobi_polar_ws_elems(const JacobiPlan &pl);
// unsigned words of the barrier/statistics buffer jacobi_svd needs
int jacobi_bar_words(int max_sweeps);
void jacobi_svd(hipStream_t s, const JacobiPlan &pl, const double *S, bool transS, double *Y,
                int *sweeps_dev, unsigned *barrier_dev, int *err_dev, int max_sweeps, double *ws,
                unsigned long long *dbg = nullptr, double stop_cos2 = 1e-16, const int *skip = nullptr);

// Sort the column norms of Y's top k rows (descending) and apply my_chop2.
// sv[k], perm[k] written; out_i[0] = r1 (after rmax cap), out_d[0] = ||s||.
void sort_chop(hipStream_t s, const double *Y, int64_t ldy, int64_t k, int64_t kpad,
               double ep, int64_t rmax, double *sv, int *perm, int *out_i, double *out_d,
               const int *skip = nullptr);

// skip = 1 iff 1/x >= 1.05 ep nrmR, k <= rmax and nrmR > 0 (no truncation at
// the bond), x the smaller upper bound of ||R^{-1}||_2: nrmX = ||R^{-1}||_F or
// the one from the normalised squarings' norms pw[0..3] (pw may be null)
void notrunc_decide(hipStream_t s, const double *nrmR, const double *nrmX, const double *pw, double ep,
                    int64_t k, int64_t rmax, int *skip);
// X = I + sign * U (k x k; U may be null)
void eye_plus(hipStream_t s, int64_t k, double *X, const double *U, double sign);

// Gather columns perm[0..r1) of Y (rows x *) with row offset into dst (rows_out x r1).
template <class T>
void gather_cols(hipStream_t s, const double *Y, int64_t ldy, int64_t row0, int64_t rows,
                 const int *perm, int64_t r1, T *dst, int64_t ldd);
// Like gather_cols but writes the transpose: dst (r1 x rows), ld = r1.
template <class T>
void gather_cols_t(hipStream_t s, const double *Y, int64_t ldy, int64_t row0, int64_t rows,
                   const int *perm, int64_t r1, T *dst);

// balance: nrm[0] = ||core0||, core0 /= nrm[0]; nrm0 = exp(mean log|nrm|); all *= nrm0.
// Cores are gathered from src pointers (device array of pointers) into a packed output.
template <class TI>
void finalize_balance(hipStream_t s, int64_t d, TI *const *src_dev, const int64_t *sizes_dev,
                      const int64_t *ofs_dev, int64_t size0, double *nrm_dev, double *out,
                      int64_t total);

// mtimes: one product core (rA*rx) x n x (rA'*rx') (mtimes.m:38-65 layout)
template <class T>
void mtimes_core(hipStream_t s, int64_t ra0, int64_t n, int64_t m, int64_t ra1, const double *A,
                 int64_t rx0, int64_t rx1, const double *X, T *Y);
// every product core of A*x in one launch (table of cores in device memory)
struct MvCore {
  const double *A, *X;
  double *Y;
  int64_t ra0, n, m, ra1, rx0, rx1;
};
// max_units: the largest mtimes_units(ra0 rx0, rx1) over the cores
int64_t mtimes_units(int64_t r0, int64_t rx1);
// bytes: algorithmic HBM traffic (A, X read, Y written) for kernel-class timing
void mtimes_all(hipStream_t s, const MvCore *tab_dev, int ncores, int64_t max_units, double bytes);
// C (M x n ra1 rx1) = R (M x ra0 rx0, ld ldr) * (product core of A and X), the
// product core generated on chip (never written to memory).
void gemm_mtimes(hipStream_t s, int64_t M, const double *R, int64_t ldr, int64_t ra0, int64_t n,
                 int64_t m, int64_t ra1, const double *A, int64_t rx0, int64_t rx1, const double *X,
                 double *C, int64_t ldc);

// One core of Z = a X + b Y in the rank concatenation of
// @tt_tensor/plus.m:41-86: X's core in the leading block, Y's in the
// trailing block, zeros elsewhere (sum != 0: d = 1, the cores add).
struct PlusCore {
  const double *x, *y;
  double *z;
  int64_t rx0, rx1, ry0, ry1, rz0, rz1, n;
  double a, b;
  int sum, pad;
};
// assembles ncores cores (table in device memory) in one launch
void plus_assemble(hipStream_t s, const PlusCore *tab_dev, int ncores, int64_t max_core);
void plus_assemble_one(hipStream_t s, const PlusCore &c);
// ranks of a X + b Y (boundary ranks must agree; throws TTG_BAD_ARG)
std::vector<int64_t> plus_ranks(int64_t d, const int64_t *n, const int64_t *rx, const int64_t *ry);
// uploads the table of cores [k0, k1) of a X + b Y (destinations zdst[k - k0])
// to tab_dev through pinned memory on ctx's stream; returns the largest core
int64_t plus_table(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx, double a,
                   const int64_t *ry, const double *dy, double b, const std::vector<int64_t> &rz,
                   double *const *zdst, int64_t k0, int64_t k1, PlusCore *tab_dev);

// TSQR: Rt (m x m) = R^T of the QR of A^T, A (m x M, ld m).  W, Sa, Sb:
// scratch of tsqr_ws_elems(m, M, &level) doubles (W) and `level` doubles each.
int64_t tsqr_ws_elems(int64_t m, int64_t M, int64_t *level_elems);
void tsqr_rt(hipStream_t s, const double *A, int64_t m, int64_t M, double *W, double *Sa, double *Sb, double *Rt);

// check for non-finite values; sets *flag nonzero if any
void check_finite(hipStream_t s, const double *x, int64_t count, int *flag);
void copy_check_finite(hipStream_t s, double *y, const double *x, int64_t count, int *flag);

}  // namespace ttg
