// ttgpu device kernels for gfx950 (CDNA4, wave64).
//
// Dense building blocks of the TT rounding sweep:
//   * MFMA GEMM (FP64 16x16x4, FP32 16x16x4), column-major, split-K with a
//     fixed-order reduction (results are bitwise reproducible run to run);
//   * one-workgroup Householder QR (LAPACK dgeqr2 + dorg2r semantics);
//   * one-workgroup Cholesky of a k x k Gram matrix, wave-per-column
//     triangular inverse;
//   * cooperative block one-sided Jacobi SVD (Hestenes rotations, de Rijk /
//     Rutishauser formulas) with column blocks resident in LDS;
//   * descending sort of singular values + the my_chop2 rank rule;
//   * gathers, transposes, norms and the final balance of round.m:94-107.
#include "ttg_internal.h"
#include "ttg_device.h"
#include <cmath>
#include <cfloat>
#include <algorithm>
#include <type_traits>

namespace ttg {

thread_local Profiler *g_prof = nullptr;

size_t Profiler::take_pair() {
  if (pool.size() < 2 * (used + 1)) {
    hipEvent_t a, b;
    TTG_HIP(hipEventCreate(&a));
    TTG_HIP(hipEventCreate(&b));
    pool.push_back(a);
    pool.push_back(b);
  }
  flops.push_back(0.0);
  bytes.push_back(0.0);
  return used++;
}
void Profiler::begin(hipStream_t s) {
  const size_t i = take_pair();
  TTG_HIP(hipEventRecord(pool[2 * i], s));
}
void Profiler::end(hipStream_t s, double fl, double by) {
  const size_t i = used - 1;
  TTG_HIP(hipEventRecord(pool[2 * i + 1], s));
  flops[i] = fl;
  bytes[i] = by;
}

// ---------------------------------------------------------------------------
// MFMA GEMM
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <class T> struct MF;
// Accumulator layouts (measured on gfx950 with an in-bounds one-wave probe):
//   f64 16x16x4: lane l, element v holds D[l/16 + 4v][l%16]
//   f32 16x16x4: lane l, element v holds D[4(l/16) + v][l%16]
template <> struct MF<double> {
  typedef v4d V;
  static __device__ __forceinline__ V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct MF<float> {
  typedef v4f V;
  static __device__ __forceinline__ V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

constexpr int GBM = 64, GBN = 64, GBK = 64;
// The tile loads address an operand by 32-bit byte offsets within 64 of its
// columns: leading dimensions must stay below 2^22 elements.
constexpr int64_t kGemmMaxLd = int64_t(1) << 22;
static inline void gemm_check_ld(int64_t lda, int64_t ldb) {
  TTG_CHECK(lda < kGemmMaxLd && ldb < kGemmMaxLd, TTG_BAD_ARG,
            "GEMM operand with a leading dimension of 2^22 elements or more");
}

// C(MxN) = alpha*op(A)*op(B) + beta*C.  Block BM x BN (64 x 64; 32 x 64 when
// 64 x 64 blocks would not fill the CUs; 32 x 32 when even 32 x 64 blocks
// would leave one workgroup per CU, so that two share each CU and cover each
// other's load latency), 4 waves (2x2), each wave a (BM/2) x (BN/2)
// sub-tile made of (BM/32) x (BN/32) MFMA 16x16x4 tiles.  Operand layout of
// v_mfma_*_16x16x4: lane l supplies A[l%16][l/16] and B[l/16][l%16]; the
// accumulator element v of lane l is D[MF<T>::row(l, v)][l%16].  The waves
// form the transposed tiles (op(B)^T op(A)^T), so that lane l holds rows
// l%16 of C and 16 lanes store 16 consecutive rows (coalesced).  K tiles
// go through two LDS buffers (gemm_acc): while one is multiplied, the next
// tile is written to the other and the one after is loaded into registers.
// Split-K: blockIdx.z selects K range [z*kchunk, ...); with gridDim.z > 1 the
// raw partial product goes to P + z*M*N (ld M) and k_splitk_reduce finishes.
// TI is the storage type of A and B, T the MFMA type: TI = double, T = float
// rounds the operands to FP32 while staging them (FP32 Gram products of FP64
// data).
// One element through a buffer descriptor: raw (stride 0) load at byte offset
// voff; an offset at or past the descriptor's range reads 0.
template <class TI> __device__ __forceinline__ TI buffer_load(__amdgpu_buffer_rsrc_t r, unsigned voff);
template <> __device__ __forceinline__ double buffer_load<double>(__amdgpu_buffer_rsrc_t r, unsigned voff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}
template <> __device__ __forceinline__ float buffer_load<float>(__amdgpu_buffer_rsrc_t r, unsigned voff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
}

// K tile of a BM x BN block: 32 for FP64 (two LDS buffers of a 32 x 32 tile
// take 33 KB, so four workgroups fit a CU's 160 KB and the two Gram chains'
// products, launched on two streams, run side by side; at K tile 64 a 32 x 32
// tile takes 68 KB and one chain's product fills every CU), 64 for FP32
template <class T, int BM, int BN> struct GemmKT {
  static constexpr int value = sizeof(T) == 8 ? 32 : 64;
};

// One BM x BN block of op(A) op(B) over K range [kbeg, kend) with its
// top-left corner at (m0, n0), accumulated in acc (layout below), staged
// through the caller's two LDS buffers As[2], Bs[2] (K tile KT).  Tile t is
// multiplied from buffer t % 2 while tile t+1 (in registers since tile t-2)
// is written to the other buffer and tile t+3 is loaded into the register
// stage it frees (two register stages: a load has two tiles' time to
// arrive): one workgroup barrier per K tile.  Rows of the LDS tiles are padded by one
// element (odd stride), so the transposed writes of a K-contiguous operand
// (64 lanes, 64 rows, one column) fall in distinct banks.
template <class TI, class T, bool TA, bool TB, int BM, int BN, int KT>
__device__ __forceinline__ void gemm_acc(int M, int N, int kbeg, int kend, int m0, int n0,
                                         const TI *__restrict__ A, long lda, const TI *__restrict__ B, long ldb,
                                         T (*As)[KT][BM + 1], T (*Bs)[KT][BN + 1],
                                         typename MF<T>::V (&acc)[BM / 32][BN / 32]) {
  static_assert(!TB || BN == 64, "op(B) = B^T stages 64 columns per K row");
  static_assert(KT == 32 || KT == 64, "K tile 32 or 64");
  constexpr int MI = BM / 32;            // 16-row MFMA tiles per wave
  constexpr int NJ = BN / 32;            // 16-column MFMA tiles per wave
  constexpr int LA = BM * KT / 256;      // A elements each thread stages per K tile
  constexpr int LB = BN * KT / 256;      // B elements each thread stages per K tile
  constexpr int KQ = 256 / KT;           // threads along M or N when the operand is K-contiguous
  typedef typename MF<T>::V V;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[i][j] = V{0, 0, 0, 0};
  // consecutive threads read consecutive elements of the operand's contiguous
  // dimension (M or K for A, K or N for B); prefetched in the storage type (a
  // conversion here would stall on the load).  Loads are buffer loads from a
  // descriptor based at the tile's first element: each thread's element
  // offsets are fixed (computed once), and elements outside the M x K (K x N)
  // window read 0 through the descriptor's range check (for the slow index of
  // the operand: the range ends at the window's edge) or through an offset
  // past any range (for the fast index: that thread's elements, chosen once
  // for M and N and per tile for K).  No branch and no 64-bit address
  // arithmetic per element.
  constexpr int NS = 2;  // register stages: K tiles in flight while one is multiplied
  TI ra[NS][LA], rb[NS][LB];
  constexpr int AQ = 256 / BM;  // A, M contiguous: BM consecutive threads read BM consecutive rows
  const int kc = tid % KT, kq = tid / KT;
  constexpr unsigned OOB = 0x80000000u;  // past every range (ranges are below 2^31 bytes)
  unsigned va[LA], vb[LB];
#pragma unroll
  for (int i = 0; i < LA; ++i) {
    if (!TA) {
      const bool ok = m0 + tid % BM < M;
      va[i] = ok ? (unsigned)(tid % BM + (long)(tid / BM + AQ * i) * lda) * (unsigned)sizeof(TI) : OOB;
    } else {
      va[i] = (unsigned)(kc + (long)(kq + KQ * i) * lda) * (unsigned)sizeof(TI);
    }
  }
#pragma unroll
  for (int i = 0; i < LB; ++i) {
    if (!TB) {
      vb[i] = (unsigned)(kc + (long)(kq + KQ * i) * ldb) * (unsigned)sizeof(TI);
    } else {
      const bool ok = n0 + lane < N;
      vb[i] = ok ? (unsigned)(lane + (long)(wave + 4 * i) * ldb) * (unsigned)sizeof(TI) : OOB;
    }
  }
  auto bytes_ok = [](long b) { return (int)(b <= 0 ? 0 : (b > 0x7fffffffL ? 0x7fffffffL : b)); };
  auto fetch = [&](int st, int k0) {
    // A: tile at (m0, k0) of op(A); the range ends at column kend (!TA: K is
    // the slow index) or at row M of op(A) (TA: M is the slow index)
    const TI *abase = TA ? A + k0 + (long)m0 * lda : A + m0 + (long)k0 * lda;
    const long abytes = TA ? (long)(M - m0) * lda * (long)sizeof(TI) : (long)(kend - k0) * lda * (long)sizeof(TI);
    const __amdgpu_buffer_rsrc_t ar =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<TI *>(abase), (short)0, bytes_ok(abytes), 0x00020000);
    const TI *bbase = TB ? B + n0 + (long)k0 * ldb : B + k0 + (long)n0 * ldb;
    const long bbytes = TB ? (long)(kend - k0) * ldb * (long)sizeof(TI) : (long)(N - n0) * ldb * (long)sizeof(TI);
    const __amdgpu_buffer_rsrc_t br =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<TI *>(bbase), (short)0, bytes_ok(bbytes), 0x00020000);
    // this thread's K index past kend (K-contiguous operands): every offset
    // gets the range-breaking top bit (offsets are below 2^31)
    const unsigned kout = (k0 + kc < kend) ? 0u : OOB;
#pragma unroll
    for (int i = 0; i < LA; ++i) ra[st][i] = buffer_load<TI>(ar, TA ? (va[i] | kout) : va[i]);
#pragma unroll
    for (int i = 0; i < LB; ++i) rb[st][i] = buffer_load<TI>(br, TB ? vb[i] : (vb[i] | kout));
  };
  auto stash = [&](int buf, int st) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      if (!TA) As[buf][tid / BM + AQ * i][tid % BM] = T(ra[st][i]);
      else As[buf][kc][kq + KQ * i] = T(ra[st][i]);
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      if (!TB) Bs[buf][kc][kq + KQ * i] = T(rb[st][i]);
      else Bs[buf][wave + 4 * i][lane] = T(rb[st][i]);
    }
  };
  // software-pipelined over the tile's K steps: the operands of step s+1
  // are read from LDS before step s's MFMAs issue (two register sets), and
  // the scheduling groups keep that order (reads of s+1, then MFMAs of s), so
  // an MFMA waits for reads issued one step earlier instead of just before it
  auto multiply = [&](int buf) {
    constexpr int KS = KT / 4;
    T a[2][MI], b[2][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i) a[0][i] = As[buf][(lane >> 4)][wm * (BM / 2) + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < NJ; ++j) b[0][j] = Bs[buf][(lane >> 4)][wn * (BN / 2) + j * 16 + (lane & 15)];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int c = s & 1, nx = c ^ 1;
      if (s + 1 < KS) {
#pragma unroll
        for (int i = 0; i < MI; ++i)
          a[nx][i] = As[buf][4 * (s + 1) + (lane >> 4)][wm * (BM / 2) + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          b[nx][j] = Bs[buf][4 * (s + 1) + (lane >> 4)][wn * (BN / 2) + j * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = MF<T>::op(b[c][j], a[c][i], acc[i][j]);
      if (s + 1 < KS) {
        __builtin_amdgcn_sched_group_barrier(0x100, MI + NJ, 0);  // LDS reads of step s+1
        __builtin_amdgcn_sched_group_barrier(0x008, MI * NJ, 0);  // MFMAs of step s
      }
    }
  };
  // tile t sits in register stage t % NS until it is written to LDS buffer
  // t % 2; tiles t+1 .. t+NS are in flight while tile t is multiplied
  const int nt = kbeg < kend ? (kend - kbeg + KT - 1) / KT : 0;
  if (nt == 0) return;
  if (nt == 8) {
    // the chain products' K range (256 at K tile 32), fully unrolled: with no
    // loop-carried or conditional loads the compiler's waits before each
    // stash are exact (the newer stage stays in flight) instead of draining
    // every outstanding load
#pragma unroll
    for (int u = 0; u < NS; ++u) fetch(u, kbeg + u * KT);
    __syncthreads();  // the caller's earlier use of the buffers is over
    stash(0, 0);
    fetch(0, kbeg + NS * KT);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      multiply(t & 1);
      if (t + 1 < 8) {
        stash((t + 1) & 1, (t + 1) % NS);
        if (t + 1 + NS < 8) fetch((t + 1) % NS, kbeg + (t + 1 + NS) * KT);
      }
      __syncthreads();
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < NS; ++u)
    if (u < nt) fetch(u, kbeg + u * KT);
  __syncthreads();  // the caller's earlier use of the buffers is over
  stash(0, 0);
  if (NS < nt) fetch(0, kbeg + NS * KT);
  __syncthreads();
  for (int t0 = 0; t0 < nt; t0 += NS) {
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int t = t0 + u;
      if (t >= nt) break;
      multiply(t & 1);
      if (t + 1 < nt) {
        const int st = (u + 1) % NS;
        stash((t + 1) & 1, st);  // buffer last read by tile t-1 (before the previous barrier)
        if (t + 1 + NS < nt) fetch(st, kbeg + (t + 1 + NS) * KT);
      }
      __syncthreads();
    }
  }
}

// One BM x BN block of C over K range [kbeg, kend) with its top-left corner at
// (m0, n0).  P != nullptr: the raw partial product goes to P (ld M) instead of
// C (split-K).  Shared by k_gemm (grid tiles) and k_bgemm (batched jobs).
template <class TI, class T, class TO, bool TA, bool TB, int BM, int BN>
__device__ __forceinline__ void gemm_block(int M, int N, int kbeg, int kend, int m0, int n0, T alpha,
                                           const TI *__restrict__ A, long lda, const TI *__restrict__ B,
                                           long ldb, T beta, TO *__restrict__ C, long ldc,
                                           T *__restrict__ P) {
  constexpr int KT = GemmKT<T, BM, BN>::value;
  __shared__ T As[2][KT][BM + 1];
  __shared__ T Bs[2][KT][BN + 1];
  constexpr int MI = BM / 32, NJ = BN / 32;
  typedef typename MF<T>::V V;
  V acc[MI][NJ];
  gemm_acc<TI, T, TA, TB, BM, BN, KT>(M, N, kbeg, kend, m0, n0, A, lda, B, ldb, As, Bs, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  // acc[i][j][v] of lane l: C(m0 + wm*BM/2 + i*16 + l%16, n0 + wn*BN/2 + j*16 + row(l, v))
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + wm * (BM / 2) + i * 16 + (lane & 15);
        const int gn = n0 + wn * (BN / 2) + j * 16 + MF<T>::row(lane, v);
        if (gm < M && gn < N) {
          if (P) {
            P[gm + (long)gn * M] = acc[i][j][v];
          } else {
            T val = alpha * acc[i][j][v];
            if (beta != T(0)) val += beta * T(C[gm + (long)gn * ldc]);
            C[gm + (long)gn * ldc] = TO(val);
          }
        }
      }
}

// alpha_dev (optional device scalar) multiplies alpha.
// Workgroups are dispatched to the 8 XCDs round-robin (block w on XCD w % 8),
// and each XCD has its own L2.  This bijection gives XCD x the contiguous
// tile range [x q + min(x, rem), ...) (q = total / 8, rem = total % 8), so the
// M tiles that share one B panel (consecutive in a job's tile order) run on
// one XCD and read the panel from HBM once instead of once per XCD.
__device__ __forceinline__ int xcd_tile(int w, int total) {
  constexpr int X = 8;
  const int q = total / X, rem = total % X;
  const int x = w % X, local = w / X;
  return x * q + min(x, rem) + local;
}

template <class TI, class T, class TO, bool TA, bool TB, int BM = GBM, int BN = GBN>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, int kchunk, T alpha,
                                              const TI *__restrict__ A, long lda,
                                              const TI *__restrict__ B, long ldb, T beta,
                                              TO *__restrict__ C, long ldc, T *__restrict__ P,
                                              const T *__restrict__ alpha_dev, int sym) {
  // XCD-aware placement: the grid (M tiles, N tiles, K chunks) is walked as
  // one tile index, K chunk slowest, and within a chunk M fastest when op(B)
  // is the larger operand (N >= M), N fastest otherwise, so each XCD's
  // contiguous run of tiles shares the larger operand's panels (and a split-K
  // product gives each XCD its own K range)
  const int gx = (int)gridDim.x, gy = (int)gridDim.y, gxy = gx * gy;
  const int t = xcd_tile((int)(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z)), gxy * (int)gridDim.z);
  const int tz = t / gxy, tr = t - tz * gxy;
  const bool mfast = N >= M;
  const int im = mfast ? tr % gx : tr / gy, in = mfast ? tr / gx : tr % gy;
  // sym: the product is symmetric (M == N) and only blocks that reach the
  // diagonal or above are formed; k_splitk_reduce mirrors them
  if (sym && im * BM >= in * BN + BN) return;
  const int kbeg = tz * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const T a = alpha_dev ? alpha * alpha_dev[0] : alpha;
  T *part = gridDim.z > 1 ? P + (long)tz * M * N : nullptr;
  gemm_block<TI, T, TO, TA, TB, BM, BN>(M, N, kbeg, kend, im * BM, in * BN, a, A, lda, B, ldb,
                                        beta, C, ldc, part);
}

// Batched GEMM: job j owns tiles [tile0_j, tile0_{j+1}); every job has its
// own shape, operands and (optional) device scale; full K per block.
template <class T, bool TA, bool TB, int BM, int BN>
__global__ __launch_bounds__(256) void k_bgemm(const BGemmJob *__restrict__ jobs, int njobs) {
  const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  int j = 0;
  while (j + 1 < njobs && jobs[j + 1].tile0 <= tile) ++j;
  const BGemmJob jb = jobs[j];
  const int t = tile - jb.tile0;
  const int tm = (jb.M + BM - 1) / BM, tn = (jb.N + BN - 1) / BN;
  // consecutive tiles (one XCD) share the larger operand's panel: M fastest
  // when op(B) is the larger, N fastest when op(A) is
  const bool mfast = jb.N >= jb.M;
  const int im = mfast ? t % tm : t / tn, in = mfast ? t / tm : t % tn;
  const T a = jb.alpha_dev ? T(jb.alpha * jb.alpha_dev[0]) : T(jb.alpha);
  int kb = 0, ke = jb.K;
  if (jb.tri & 1) kb = max(kb, in * BN);
  if (jb.tri & 8) kb = max(kb, im * BM);
  if (jb.tri & 2) ke = min(ke, im * BM + BM);
  if (jb.tri & 4) ke = min(ke, in * BN + BN);
  gemm_block<T, T, T, TA, TB, BM, BN>(jb.M, jb.N, kb, ke, im * BM, in * BN, a,
                                      static_cast<const T *>(jb.A), jb.lda, static_cast<const T *>(jb.B), jb.ldb,
                                      T(0), static_cast<T *>(jb.C), jb.ldc, nullptr);
}

int bgemm_plan(bool tb, BGemmJob *jobs, int njobs, int *shape, double *flops, int force) {
  // 64 x 64 blocks when they give at least two per CU (256 CUs), else
  // 32 x 64; 32 x 32 when op(B) = B and the batch has few blocks
  int tiles_all = 0, tiles_big = 0;
  for (int j = 0; j < njobs; ++j) {
    tiles_all += ((jobs[j].M + 31) / 32) * ((jobs[j].N + 63) / 64);
    tiles_big += ((jobs[j].M + 63) / 64) * ((jobs[j].N + 63) / 64);
  }
  int sh = tiles_big >= 512 ? 2 : (!tb && tiles_all < 1024 ? 1 : 0);
  if (force >= 0 && force <= 2 && !(force == 1 && tb)) sh = force;
  const int bm = sh == 2 ? 64 : 32, bn = sh == 1 ? 32 : 64;
  int total = 0;
  double fl = 0.0;
  for (int j = 0; j < njobs; ++j) {
    gemm_check_ld(jobs[j].lda, jobs[j].ldb);
    jobs[j].tile0 = total;
    total += ((jobs[j].M + bm - 1) / bm) * ((jobs[j].N + bn - 1) / bn);
    fl += 2.0 * jobs[j].M * (double)jobs[j].N * jobs[j].K;
  }
  if (shape) *shape = sh;
  if (flops) *flops = fl;
  return total;
}

void bgemm_run(hipStream_t st, bool ta, bool tb, const BGemmJob *jobs_dev, int njobs, int total, int shape,
               double flops) {
  if (njobs <= 0 || total <= 0) return;
  ProfScope prof(KC_GEMM, st, flops, 0.0, true);
  const dim3 grid((unsigned)total);
#define TTG_BGEMM(TA_, TB_)                                                                                   \
  do {                                                                                                        \
    if (shape == 1)                                                                                           \
      TTG_TIMED_LAUNCH(prof, (k_bgemm<double, TA_, false, 32, 32>), grid, dim3(256), 0, st, jobs_dev, njobs); \
    else if (shape == 2)                                                                                      \
      TTG_TIMED_LAUNCH(prof, (k_bgemm<double, TA_, TB_, 64, 64>), grid, dim3(256), 0, st, jobs_dev, njobs);   \
    else                                                                                                      \
      TTG_TIMED_LAUNCH(prof, (k_bgemm<double, TA_, TB_, 32, 64>), grid, dim3(256), 0, st, jobs_dev, njobs);   \
  } while (0)
  if (!ta && !tb) TTG_BGEMM(false, false);
  else if (ta && !tb) TTG_BGEMM(true, false);
  else if (!ta && tb) TTG_BGEMM(false, true);
  else TTG_BGEMM(true, true);
#undef TTG_BGEMM
  TTG_HIP(hipGetLastError());
}

// Fixed-order sum of the split-K partials (accumulated in double).
// sym (M == N, only the blocks on or above the diagonal were formed): an
// entry below the diagonal takes its mirror's sums, so C is exactly symmetric.
template <class T, class TO>
__global__ void k_splitk_reduce(int M, int N, int S, const T *__restrict__ P, double alpha,
                                double beta, TO *__restrict__ C, long ldc, const T *__restrict__ alpha_dev,
                                int sym) {
  if (alpha_dev) alpha *= (double)alpha_dev[0];
  const long total = (long)M * N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int gm = (int)(idx % M);
    const long gn = idx / M;
    if (sym && gm > gn) continue;  // written as the mirror of (gn, gm)
    double s = 0.0;
    for (int z = 0; z < S; ++z) s += (double)P[(long)z * total + idx];
    double v = alpha * s;
    if (beta != 0.0) v += beta * (double)C[gm + gn * ldc];
    C[gm + gn * ldc] = (TO)v;
    if (sym && gm < gn) C[gn + gm * ldc] = (TO)v;
  }
}

static int splitk_count(int64_t M, int64_t N, int64_t K, int num_cus) {
  const int64_t tiles = ((M + GBM - 1) / GBM) * ((N + GBN - 1) / GBN);
  if (tiles >= num_cus || K < 512) return 1;
  int64_t s = (num_cus + tiles - 1) / tiles;
  s = std::min<int64_t>(s, K / 256);
  s = std::min<int64_t>(s, 32);
  return (int)std::max<int64_t>(s, 1);
}

static bool gemm_half_tiles(int64_t M, int64_t N, int S, int num_cus) {
  const int64_t tiles = ((M + GBM - 1) / GBM) * ((N + GBN - 1) / GBN);
  return tiles * S < num_cus;
}

// 32 x 32 blocks: 32 x 64 blocks would give at most one workgroup per CU
// (op(B) = B only)
static bool gemm_quarter_tiles(int64_t M, int64_t N, int S, int num_cus, bool tb) {
  const int64_t tiles = ((M + 31) / 32) * ((N + GBN - 1) / GBN);
  return !tb && tiles * S <= num_cus;
}

int64_t gemm_ws_elems(int64_t M, int64_t N, int64_t K, int num_cus) {
  const int s = splitk_count(M, N, K, num_cus);
  return s > 1 ? (int64_t)s * M * N : 0;
}

template <class TI, class T, class TO>
static void gemm_impl(hipStream_t st, bool ta, bool tb, int64_t M, int64_t N, int64_t K, T alpha,
                      const TI *A, int64_t lda, const TI *B, int64_t ldb, T beta, TO *C, int64_t ldc,
                      T *ws, int64_t ws_elems, int num_cus, const T *alpha_dev = nullptr, bool sym = false) {
  if (M <= 0 || N <= 0) return;
  gemm_check_ld(lda, ldb);
  int S = splitk_count(M, N, K, num_cus);
  if (S > 1 && (ws == nullptr || (int64_t)S * M * N > ws_elems)) S = 1;
  // symmetric products skip their lower blocks only when split-K's reduction
  // (which mirrors them) runs
  const int sy = (sym && S > 1 && M == N) ? 1 : 0;
  int kchunk = (int)K;
  if (S > 1) {
    kchunk = (int)(((K + S - 1) / S + GBK - 1) / GBK * GBK);
    S = (int)((K + kchunk - 1) / kchunk);
  }
  if (K <= 0) kchunk = 0;
  // 32-row tiles when 64 x 64 tiles would leave CUs idle (bond-sized products)
  const bool half = gemm_half_tiles(M, N, S, num_cus);
  const bool quarter = half && gemm_quarter_tiles(M, N, S, num_cus, tb);
  const int bm = half ? 32 : GBM, bn = quarter ? 32 : GBN;
  dim3 grid((unsigned)((M + bm - 1) / bm), (unsigned)((N + bn - 1) / bn), (unsigned)S);
  ProfScope prof(KC_GEMM, st, 2.0 * (double)M * (double)N * (double)K,
                 (double)sizeof(TI) * ((double)M * K + (double)K * N) + (double)sizeof(TO) * M * N, true);
#define TTG_GEMM_LAUNCH(TA_, TB_)                                                                                \
  do {                                                                                                           \
    if (quarter)                                                                                                 \
      TTG_TIMED_LAUNCH(prof, (k_gemm<TI, T, TO, TA_, false, 32, 32>), grid, dim3(256), 0, st, (int)M, (int)N,    \
                       (int)K, kchunk, alpha, A, (long)lda, B, (long)ldb, beta, C, (long)ldc, ws, alpha_dev, sy);    \
    else if (half)                                                                                               \
      TTG_TIMED_LAUNCH(prof, (k_gemm<TI, T, TO, TA_, TB_, 32>), grid, dim3(256), 0, st, (int)M, (int)N, (int)K,  \
                       kchunk, alpha, A, (long)lda, B, (long)ldb, beta, C, (long)ldc, ws, alpha_dev, sy);            \
    else                                                                                                         \
      TTG_TIMED_LAUNCH(prof, (k_gemm<TI, T, TO, TA_, TB_, GBM>), grid, dim3(256), 0, st, (int)M, (int)N, (int)K, \
                       kchunk, alpha, A, (long)lda, B, (long)ldb, beta, C, (long)ldc, ws, alpha_dev, sy);            \
  } while (0)
  if (!ta && !tb) TTG_GEMM_LAUNCH(false, false);
  else if (ta && !tb) TTG_GEMM_LAUNCH(true, false);
  else if (!ta && tb) TTG_GEMM_LAUNCH(false, true);
  else TTG_GEMM_LAUNCH(true, true);
#undef TTG_GEMM_LAUNCH
  TTG_HIP(hipGetLastError());
  if (S > 1) {
    const long total = (long)M * N;
    const int blocks = (int)std::min<long>((total + 255) / 256, 4096);
    TTG_TIMED_LAUNCH(prof, (k_splitk_reduce<T, TO>), dim3(blocks), dim3(256), 0, st, (int)M, (int)N, S,
                     (const T *)ws, (double)alpha, (double)beta, C, (long)ldc, alpha_dev, sy);
    TTG_HIP(hipGetLastError());
  }
}

template <>
void gemm<double>(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K, double alpha,
                  const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
                  double *C, int64_t ldc, double *ws, int64_t wse, int cus) {
  gemm_impl<double, double, double>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, ws, wse, cus);
}
template <>
void gemm<float>(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K, float alpha,
                 const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
                 int64_t ldc, float *ws, int64_t wse, int cus) {
  gemm_impl<float, float, float>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, ws, wse, cus);
}
void gemm_sym(hipStream_t s, bool ta, bool tb, int64_t M, int64_t K, const double *A, int64_t lda, const double *B,
              int64_t ldb, double *C, int64_t ldc, double *ws, int64_t wse, int cus) {
  gemm_impl<double, double, double>(s, ta, tb, M, M, K, 1.0, A, lda, B, ldb, 0.0, C, ldc, ws, wse, cus, nullptr,
                                    true);
}
void gemm_scaled(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K, const double *alpha_dev,
                 const double *A, int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc, double *ws,
                 int64_t wse, int cus) {
  gemm_impl<double, double, double>(s, ta, tb, M, N, K, 1.0, A, lda, B, ldb, 0.0, C, ldc, ws, wse, cus, alpha_dev);
}
// FP32 operands, FP64 result (Gram matrices of the mixed form).
void gemm_f32_f64out(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K,
                     const float *A, int64_t lda, const float *B, int64_t ldb, double *C,
                     int64_t ldc, float *ws, int64_t wse, int cus) {
  gemm_impl<float, float, double>(s, ta, tb, M, N, K, 1.0f, A, lda, B, ldb, 0.0f, C, ldc, ws, wse, cus);
}
// FP64 operands rounded to FP32 on the way into LDS, FP32 MFMA, FP64 result
// (the Gram products of the mixed form; ws holds FP32 split-K partials).
void gemm_f64_via_f32(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K,
                      const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
                      int64_t ldc, float *ws, int64_t wse, int cus) {
  gemm_impl<double, float, double>(s, ta, tb, M, N, K, 1.0f, A, lda, B, ldb, 0.0f, C, ldc, ws, wse, cus);
}

// ---------------------------------------------------------------------------
// Near-identity second pass of CholeskyQR2, fused with the split-K sum of its
// Gram product G = Q1^T Q1 = I + E (small E).  The upper factor R2 = I + U of
// G solves U + U^T + U^T U = E; U = Phi(E), Phi(X) = triu(X, 1) + diag(X)/2,
// misses it by O(|E|^2), and R2^{-1} = I - U to O(|U|^2), so Q = Q1 R2^{-1}
// is one GEMM; R2 is so close to I that the explicit inverse costs no
// accuracy.  The host accepts the pass only if max|E| is below the bound for
// which these truncations stay at rounding level (or at the form's error
// budget).  One kernel: G = sum_z P_z (fixed order, FP64), X = I - U,
// R2 = I + U, maxabs = max|E| (one atomic per workgroup; zeroed first).
// ---------------------------------------------------------------------------
template <class TP>
__global__ __launch_bounds__(256) void k_gram_nearid(int k, int S, const TP *__restrict__ P, double *__restrict__ X,
                                                     double *__restrict__ R2, double *maxabs) {
  __shared__ double red[4];
  const long total = (long)k * k;
  double mx = 0.0;
  for (long idx = blockIdx.x * 256L + threadIdx.x; idx < total; idx += (long)gridDim.x * 256L) {
    double g = 0.0;
    for (int z = 0; z < S; ++z) g += (double)P[(long)z * total + idx];
    const long c = idx / k, r = idx - c * k;
    const double id = (r == c) ? 1.0 : 0.0;
    const double e = g - id;
    const double u = (r < c) ? e : (r == c ? 0.5 * e : 0.0);
    X[idx] = id - u;
    R2[idx] = id + u;
    mx = fmax(mx, fabs(e));
  }
  mx = block_max(mx, red);
  if (threadIdx.x == 0 && mx > 0.0)  // non-negative doubles order like their bit patterns
    atomicMax(reinterpret_cast<unsigned long long *>(maxabs), (unsigned long long)__double_as_longlong(mx));
}

__global__ void k_zero_d(double *p, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0.0;
}

template <class T>
static void gram_nearid_impl(hipStream_t st, int64_t m, int64_t k, const double *A, double *G, double *X,
                             double *R2, double *maxabs, T *ws, int64_t ws_elems, int cus) {
  gemm_check_ld(m, m);
  // cleared by a kernel, not a memset node: the call may be captured in a graph
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, st, maxabs, 1);
  int S = splitk_count(k, k, m, cus);
  if (S > 1 && (ws == nullptr || (int64_t)S * k * k > ws_elems)) S = 1;
  const long total = (long)k * k;
  const int blocks = (int)std::min<long>((total + 1023) / 1024, 128);
  if (S > 1) {
    int kchunk = (int)(((m + S - 1) / S + GBK - 1) / GBK * GBK);
    S = (int)((m + kchunk - 1) / kchunk);
    const bool half = gemm_half_tiles(k, k, S, cus);
    const bool quarter = half && gemm_quarter_tiles(k, k, S, cus, false);
    const int bm = half ? 32 : GBM, bn = quarter ? 32 : GBN;
    dim3 grid((unsigned)((k + bm - 1) / bm), (unsigned)((k + bn - 1) / bn), (unsigned)S);
    {
      ProfScope prof(KC_GEMM, st, 2.0 * (double)k * k * m, 8.0 * (2.0 * m * k + k * k), true);
      if (quarter)
        TTG_TIMED_LAUNCH(prof, (k_gemm<double, T, double, true, false, 32, 32>), grid, dim3(256), 0, st, (int)k,
                         (int)k, (int)m, kchunk, T(1), A, (long)m, A, (long)m, T(0), G, (long)k, ws, (const T *)nullptr, 0);
      else if (half)
        TTG_TIMED_LAUNCH(prof, (k_gemm<double, T, double, true, false, 32>), grid, dim3(256), 0, st, (int)k, (int)k,
                         (int)m, kchunk, T(1), A, (long)m, A, (long)m, T(0), G, (long)k, ws, (const T *)nullptr, 0);
      else
        TTG_TIMED_LAUNCH(prof, (k_gemm<double, T, double, true, false, GBM>), grid, dim3(256), 0, st, (int)k, (int)k,
                         (int)m, kchunk, T(1), A, (long)m, A, (long)m, T(0), G, (long)k, ws, (const T *)nullptr, 0);
      TTG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL((k_gram_nearid<T>), dim3(blocks), dim3(256), 0, st, (int)k, S, (const T *)ws, X, R2, maxabs);
  } else {
    if constexpr (std::is_same<T, double>::value)
      gemm<double>(st, true, false, k, k, m, 1.0, A, m, A, m, 0.0, G, k, nullptr, 0, cus);
    else
      gemm_f64_via_f32(st, true, false, k, k, m, A, m, A, m, G, k, nullptr, 0, cus);
    hipLaunchKernelGGL((k_gram_nearid<double>), dim3(blocks), dim3(256), 0, st, (int)k, 1, (const double *)G, X, R2,
                       maxabs);
  }
  TTG_HIP(hipGetLastError());
}
void gram_nearid(hipStream_t st, int64_t m, int64_t k, const double *A, bool g32, double *G, double *X,
                 double *R2, double *maxabs, double *ws, int64_t ws_elems, int cus) {
  if (g32)
    gram_nearid_impl<float>(st, m, k, A, G, X, R2, maxabs, reinterpret_cast<float *>(ws), 2 * ws_elems, cus);
  else
    gram_nearid_impl<double>(st, m, k, A, G, X, R2, maxabs, ws, ws_elems, cus);
}

// ---------------------------------------------------------------------------
// elementwise helpers
// ---------------------------------------------------------------------------
template <class TO, class TI>
__global__ void k_convert(TO *y, const TI *x, long count) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
    y[i] = (TO)x[i];
}
template <class TO, class TI> void convert(hipStream_t s, TO *y, const TI *x, int64_t count) {
  if (count <= 0) return;
  const int blocks = (int)std::min<int64_t>((count + 255) / 256, 8192);
  hipLaunchKernelGGL((k_convert<TO, TI>), dim3(blocks), dim3(256), 0, s, y, x, (long)count);
  TTG_HIP(hipGetLastError());
}
template void convert<float, double>(hipStream_t, float *, const double *, int64_t);
template void convert<double, float>(hipStream_t, double *, const float *, int64_t);
template void convert<double, double>(hipStream_t, double *, const double *, int64_t);
template void convert<float, float>(hipStream_t, float *, const float *, int64_t);

// B (cols x rows) = A^T, A rows x cols; 32x32 LDS tiles.
template <class T>
__global__ void k_transpose(const T *__restrict__ A, int rows, int cols, T *__restrict__ B) {
  __shared__ T tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: ty 0..7
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + tx, c = c0 + j;
    if (r < rows && c < cols) tile[j][tx] = A[r + (long)c * rows];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + tx, r = r0 + j;
    if (r < rows && c < cols) B[c + (long)r * cols] = tile[tx][j];
  }
}
template <class T> void transpose(hipStream_t s, const T *A, int64_t rows, int64_t cols, T *B) {
  if (rows <= 0 || cols <= 0) return;
  dim3 grid((unsigned)((rows + 31) / 32), (unsigned)((cols + 31) / 32));
  hipLaunchKernelGGL((k_transpose<T>), grid, dim3(256), 0, s, A, (int)rows, (int)cols, B);
  TTG_HIP(hipGetLastError());
}
template void transpose<double>(hipStream_t, const double *, int64_t, int64_t, double *);
template void transpose<float>(hipStream_t, const float *, int64_t, int64_t, float *);

// Frobenius norm in one block (count is at most a few million here); 16
// loads in flight per thread, 16 partial sums combined in a fixed order.
template <class T>
__global__ __launch_bounds__(1024) void k_fro_norm(const T *__restrict__ x, long count, double *out) {
  __shared__ double red[16];
  constexpr int U = 16;
  double s[U];
#pragma unroll
  for (int u = 0; u < U; ++u) s[u] = 0.0;
  for (long i = threadIdx.x; i < count; i += (long)U * blockDim.x) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long idx = i + (long)u * blockDim.x;
      v[u] = idx < count ? (double)x[idx] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] += v[u] * v[u];
  }
#pragma unroll
  for (int w = U / 2; w >= 1; w >>= 1)
#pragma unroll
    for (int u = 0; u < w; ++u) s[u] += s[u + w];
  const double t = block_sum(s[0], red);
  if (threadIdx.x == 0) out[0] = sqrt(t);
}
template <class T> void fro_norm(hipStream_t s, const T *x, int64_t count, double *out) {
  hipLaunchKernelGGL((k_fro_norm<T>), dim3(1), dim3(1024), 0, s, x, (long)count, out);
  TTG_HIP(hipGetLastError());
}
template void fro_norm<double>(hipStream_t, const double *, int64_t, double *);
template void fro_norm<float>(hipStream_t, const float *, int64_t, double *);

template <class T>
__global__ void k_scale_by_inv(T *x, long count, const double *nrm) {
  const double v = nrm[0];
  if (v == 0.0) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
    x[i] = (T)((double)x[i] / v);
}
template <class T> void scale_by_inv(hipStream_t s, T *x, int64_t count, const double *nrm) {
  if (count <= 0) return;
  const int blocks = (int)std::min<int64_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL((k_scale_by_inv<T>), dim3(blocks), dim3(256), 0, s, x, (long)count, nrm);
  TTG_HIP(hipGetLastError());
}
template void scale_by_inv<double>(hipStream_t, double *, int64_t, const double *);
template void scale_by_inv<float>(hipStream_t, float *, int64_t, const double *);

__global__ void k_check_finite(const double *x, long count, int *flag) {
  int bad = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
    if (!isfinite(x[i])) bad = 1;
  if (bad) atomicOr(flag, 1);
}
void check_finite(hipStream_t s, const double *x, int64_t count, int *flag) {
  if (count <= 0) return;
  const int blocks = (int)std::min<int64_t>((count + 255) / 256, 2048);
  hipLaunchKernelGGL(k_check_finite, dim3(blocks), dim3(256), 0, s, x, (long)count, flag);
  TTG_HIP(hipGetLastError());
}
// y = x with a NaN / Inf check on the way (one read of x instead of two)
__global__ void k_copy_check_finite(double *__restrict__ y, const double *__restrict__ x, long count, int *flag) {
  int bad = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
    const double v = x[i];
    y[i] = v;
    if (!isfinite(v)) bad = 1;
  }
  if (bad) atomicOr(flag, 1);
}
void copy_check_finite(hipStream_t s, double *y, const double *x, int64_t count, int *flag) {
  if (count <= 0) return;
  const int blocks = (int)std::min<int64_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL(k_copy_check_finite, dim3(blocks), dim3(256), 0, s, y, x, (long)count, flag);
  TTG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Householder QR, one workgroup (robust path: rank-deficient / wide / small)
// ---------------------------------------------------------------------------
// LAPACK dlarfg: beta = -sign(alpha)*||x||, tau = (beta-alpha)/beta,
// v = x/(alpha-beta) with v0 = 1; tau = 0 when the sub-column is zero.
__global__ __launch_bounds__(1024) void k_householder_qr(int m, int n, double *__restrict__ A,
                                                         double *__restrict__ Q,
                                                         double *__restrict__ R,
                                                         double *__restrict__ tau) {
  __shared__ double red[16];
  __shared__ double sh_tau, sh_scale;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  const int kq = min(m, n);
  for (int j = 0; j < kq; ++j) {
    double *aj = A + (long)j * m;
    double ss = 0.0;
    for (int i = j + 1 + tid; i < m; i += blockDim.x) ss += aj[i] * aj[i];
    ss = block_sum(ss, red);
    if (tid == 0) {
      const double alpha = aj[j];
      if (ss == 0.0) {
        sh_tau = 0.0;
        sh_scale = 0.0;
      } else {
        const double nrm = sqrt(alpha * alpha + ss);
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        sh_tau = (beta - alpha) / beta;
        sh_scale = 1.0 / (alpha - beta);
        aj[j] = beta;
      }
      tau[j] = sh_tau;
    }
    __syncthreads();
    const double tj = sh_tau, sc = sh_scale;
    if (tj != 0.0)
      for (int i = j + 1 + tid; i < m; i += blockDim.x) aj[i] *= sc;
    __syncthreads();
    if (tj != 0.0) {
      for (int c = j + 1 + wv; c < n; c += nw) {
        double *ac = A + (long)c * m;
        double w = (lane == 0) ? ac[j] : 0.0;
        for (int i = j + 1 + lane; i < m; i += 64) w += aj[i] * ac[i];
        w = wave_sum(w) * tj;
        if (lane == 0) ac[j] -= w;
        for (int i = j + 1 + lane; i < m; i += 64) ac[i] -= w * aj[i];
      }
    }
    __syncthreads();
  }
  // R = upper trapezoid (kq x n)
  for (long idx = tid; idx < (long)kq * n; idx += blockDim.x) {
    const int i = (int)(idx % kq), c = (int)(idx / kq);
    R[idx] = (i <= c) ? A[i + (long)c * m] : 0.0;
  }
  __syncthreads();
  // Q = H_0 ... H_{kq-1} I(:,0:kq)  (dorg2r backward accumulation)
  for (int j = kq - 1; j >= 0; --j) {
    const double tj = tau[j];
    const double *v = A + (long)j * m;  // v[j] = 1 implicit, v[j+1..] stored
    if (j < kq - 1 && tj != 0.0) {
      for (int c = j + 1 + wv; c < kq; c += nw) {
        double *qc = Q + (long)c * m;
        double w = (lane == 0) ? qc[j] : 0.0;
        for (int i = j + 1 + lane; i < m; i += 64) w += v[i] * qc[i];
        w = wave_sum(w) * tj;
        if (lane == 0) qc[j] -= w;
        for (int i = j + 1 + lane; i < m; i += 64) qc[i] -= w * v[i];
      }
    }
    __syncthreads();
    double *qj = Q + (long)j * m;
    for (int i = tid; i < m; i += blockDim.x) {
      double val;
      if (i < j) val = 0.0;
      else if (i == j) val = 1.0 - tj;
      else val = -tj * v[i];
      qj[i] = val;
    }
    __syncthreads();
  }
}
// ---------------------------------------------------------------------------
// Blocked Householder QR (dgeqrf + dorgqr structure, compact WY): 32-column
// panels are factorised by one workgroup (k_hh_panel), which also writes the
// explicit reflector block V (unit diagonal, zeros above) and its triangular
// factor T (dlarft, forward, columnwise); the trailing matrix and the
// accumulation of Q are updated with the multi-CU GEMM:
//   A2 <- (I - V T V^T)^T A2 = A2 - V (T^T (V^T A2)),   Q2 <- Q2 - V (T (V^T Q2)).
// ---------------------------------------------------------------------------
constexpr int HB = 32;

__global__ __launch_bounds__(1024) void k_hh_panel(int m, int j0, int jb, double *__restrict__ A,
                                                   double *__restrict__ tau, double *__restrict__ V,
                                                   double *__restrict__ T) {
  __shared__ double red[16];
  __shared__ double sh_tau, sh_scale;
  __shared__ double Gv[HB][HB + 1];  // V^T V of the panel
  __shared__ double tl[HB];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  for (int c = 0; c < jb; ++c) {
    const int col = j0 + c;
    double *aj = A + (long)col * m;
    double ss = 0.0;
    for (int i = col + 1 + tid; i < m; i += blockDim.x) ss += aj[i] * aj[i];
    ss = block_sum(ss, red);
    if (tid == 0) {
      const double alpha = aj[col];
      if (ss == 0.0) {
        sh_tau = 0.0;
        sh_scale = 0.0;
      } else {
        const double nrm = sqrt(alpha * alpha + ss);
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        sh_tau = (beta - alpha) / beta;
        sh_scale = 1.0 / (alpha - beta);
        aj[col] = beta;
      }
      tau[col] = sh_tau;
      tl[c] = sh_tau;
    }
    __syncthreads();
    const double tj = sh_tau, sc = sh_scale;
    if (tj != 0.0)
      for (int i = col + 1 + tid; i < m; i += blockDim.x) aj[i] *= sc;
    __syncthreads();
    if (tj != 0.0) {
      for (int cc = col + 1 + wv; cc < j0 + jb; cc += nw) {
        double *ac = A + (long)cc * m;
        double w = (lane == 0) ? ac[col] : 0.0;
        for (int i = col + 1 + lane; i < m; i += 64) w += aj[i] * ac[i];
        w = wave_sum(w) * tj;
        if (lane == 0) ac[col] -= w;
        for (int i = col + 1 + lane; i < m; i += 64) ac[i] -= w * aj[i];
      }
    }
    __syncthreads();
  }
  // explicit V = panel reflectors (rows j0..m-1 of columns j0..j0+jb-1)
  const long rows = m - j0;
  for (long idx = tid; idx < rows * jb; idx += blockDim.x) {
    const int c = (int)(idx / rows), i = j0 + (int)(idx - (long)c * rows);
    const int col = j0 + c;
    V[i + (long)col * m] = (i < col) ? 0.0 : (i == col ? 1.0 : A[i + (long)col * m]);
  }
  __syncthreads();
  // V^T V (upper triangle): one wave per entry
  for (int e = wv; e < jb * (jb + 1) / 2; e += nw) {
    int b = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while (b * (b + 1) / 2 > e) --b;
    while ((b + 1) * (b + 2) / 2 <= e) ++b;
    const int a = e - b * (b + 1) / 2;  // a <= b
    const double *va = V + (long)(j0 + a) * m, *vb = V + (long)(j0 + b) * m;
    double g = 0.0;
    for (int i = j0 + b + lane; i < m; i += 64) g += va[i] * vb[i];
    g = wave_sum(g);
    if (lane == 0) Gv[a][b] = g;
  }
  __syncthreads();
  // T: T(i,i) = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) (V(:,0:i)^T v_i)   (dlarft),
  // built in LDS by one wave (lane r owns row r), then stored
  __shared__ double Tl[HB][HB + 1];
  if (wv == 0) {
    const int r = lane;
    for (int i = 0; i < jb; ++i) {
      double t = 0.0;
      if (r < i) {
        for (int l = r; l < i; ++l) t += Tl[r][l] * Gv[l][i];
        t *= -tl[i];
      } else if (r == i) {
        t = tl[i];
      }
      if (r < HB) Tl[r][i] = (r <= i) ? t : 0.0;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  for (int idx = tid; idx < HB * HB; idx += blockDim.x) {
    const int r = idx % HB, i = idx / HB;
    T[r + i * HB] = (r < jb && i < jb && r <= i) ? Tl[r][i] : 0.0;
  }
}

__global__ void k_triu_copy(int kq, int n, int m, const double *__restrict__ A, double *__restrict__ R) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < (long)kq * n;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % kq), c = (int)(idx / kq);
    R[idx] = (i <= c) ? A[i + (long)c * m] : 0.0;
  }
}
__global__ void k_eye(int m, int kq, double *__restrict__ Q) {
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < (long)m * kq;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % m), c = (int)(idx / m);
    Q[idx] = (i == c) ? 1.0 : 0.0;
  }
}

size_t householder_ws_elems(int64_t m, int64_t n, int num_cus) {
  const int64_t kq = std::min(m, n);
  const int64_t w = (int64_t)HB * std::max<int64_t>(n, 1);
  return (size_t)(m * kq + HB * kq + 2 * w) +
         (size_t)std::max<int64_t>(gemm_ws_elems(HB, n, m, num_cus), gemm_ws_elems(HB, kq, m, num_cus));
}

void householder_qr(hipStream_t s, int64_t m, int64_t n, double *A, double *Q, double *R, double *tau,
                    double *ws, int num_cus) {
  const int64_t kq = std::min(m, n);
  ProfScope prof(KC_HQR, s, 4.0 * m * n * kq - 2.0 * (m + n) * kq * kq + 4.0 / 3.0 * kq * kq * kq,
                 8.0 * ((double)m * n + m * kq + kq * n));
  if (kq <= HB || ws == nullptr) {
    hipLaunchKernelGGL(k_householder_qr, dim3(1), dim3(1024), 0, s, (int)m, (int)n, A, Q, R, tau);
    TTG_HIP(hipGetLastError());
    return;
  }
  double *V = ws;                       // m x kq
  double *T = V + m * kq;               // HB x kq (block j0 at column j0)
  double *W = T + HB * kq;              // HB x n
  double *W2 = W + HB * std::max<int64_t>(n, 1);
  double *gws = W2 + HB * std::max<int64_t>(n, 1);
  const int64_t gwe = std::max<int64_t>(gemm_ws_elems(HB, n, m, num_cus), gemm_ws_elems(HB, kq, m, num_cus));
  for (int64_t j0 = 0; j0 < kq; j0 += HB) {
    const int64_t jb = std::min<int64_t>(HB, kq - j0);
    hipLaunchKernelGGL(k_hh_panel, dim3(1), dim3(1024), 0, s, (int)m, (int)j0, (int)jb, A, tau, V, T + j0 * HB);
    TTG_HIP(hipGetLastError());
    const int64_t n2 = n - j0 - jb;
    if (n2 > 0) {
      const double *Vb = V + j0 + j0 * m;
      double *A2 = A + j0 + (j0 + jb) * m;
      gemm<double>(s, true, false, jb, n2, m - j0, 1.0, Vb, m, A2, m, 0.0, W, jb, gws, gwe, num_cus);
      gemm<double>(s, true, false, jb, n2, jb, 1.0, T + j0 * HB, HB, W, jb, 0.0, W2, jb, nullptr, 0, num_cus);
      gemm<double>(s, false, false, m - j0, n2, jb, -1.0, Vb, m, W2, jb, 1.0, A2, m, nullptr, 0, num_cus);
    }
  }
  const int blocks = (int)std::min<int64_t>((m * kq + 255) / 256, 2048);
  hipLaunchKernelGGL(k_triu_copy, dim3((unsigned)std::min<int64_t>((kq * n + 255) / 256, 2048)), dim3(256), 0, s,
                     (int)kq, (int)n, (int)m, A, R);
  hipLaunchKernelGGL(k_eye, dim3(blocks), dim3(256), 0, s, (int)m, (int)kq, Q);
  TTG_HIP(hipGetLastError());
  const int64_t last = (kq - 1) / HB * HB;
  for (int64_t j0 = last; j0 >= 0; j0 -= HB) {
    const int64_t jb = std::min<int64_t>(HB, kq - j0);
    const int64_t nc = kq - j0;
    const double *Vb = V + j0 + j0 * m;
    double *Q2 = Q + j0 + j0 * m;
    gemm<double>(s, true, false, jb, nc, m - j0, 1.0, Vb, m, Q2, m, 0.0, W, jb, gws, gwe, num_cus);
    gemm<double>(s, false, false, jb, nc, jb, 1.0, T + j0 * HB, HB, W, jb, 0.0, W2, jb, nullptr, 0, num_cus);
    gemm<double>(s, false, false, m - j0, nc, jb, -1.0, Vb, m, W2, jb, 1.0, Q2, m, nullptr, 0, num_cus);
  }
}

// ---------------------------------------------------------------------------
// singular values: column norms, descending rank sort, my_chop2
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_sort_chop(const double *__restrict__ Y, long ldy, int k,
                                                    int kpad, double ep, long rmax,
                                                    double *__restrict__ sv, int *__restrict__ perm,
                                                    int *out_i, double *out_d, const int *skip) {
  extern __shared__ __align__(16) double sig[];  // kpad
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  if (skip && *skip) {  // no truncation proven: keep all k columns in order
    for (int j = tid; j < k; j += blockDim.x) {
      perm[j] = j;
      sv[j] = 0.0;
    }
    if (tid == 0) {
      out_i[0] = k;
      out_d[0] = 0.0;
    }
    return;
  }
  for (int j = wv; j < kpad; j += nw) {
    const double *col = Y + (long)j * ldy;
    double s = 0.0;
    for (int i = lane; i < k; i += 64) s += col[i] * col[i];
    s = wave_sum(s);
    if (lane == 0) sig[j] = sqrt(s);
  }
  __syncthreads();
  for (int j = tid; j < kpad; j += blockDim.x) {
    const double v = sig[j];
    int pos = 0;
    for (int i = 0; i < kpad; ++i) {
      const double u = sig[i];
      pos += (u > v) || (u == v && i < j);
    }
    if (pos < k) {
      sv[pos] = v;
      perm[pos] = j;
    }
  }
  __syncthreads();
  if (tid == 0) {
    // norm(s) and my_chop2(s, norm(s)*ep)   (core/my_chop2.m:18-34)
    double ss = 0.0;
    for (int j = 0; j < k; ++j) ss += sv[j] * sv[j];
    const double nrm = sqrt(ss);
    const double e = nrm * ep;
    long r;
    if (nrm == 0.0) {
      r = 1;
    } else if (!(e > 0.0)) {
      r = k;
    } else {
      const double e2 = e * e;
      double c = 0.0;
      long K = 0;
      for (int j = k - 1; j >= 0; --j) {
        c += sv[j] * sv[j];
        if (c < e2) K = k - j;
      }
      r = k - K;
    }
    if (r > rmax) r = rmax;
    out_i[0] = (int)r;
    out_d[0] = nrm;
  }
}
void sort_chop(hipStream_t s, const double *Y, int64_t ldy, int64_t k, int64_t kpad, double ep,
               int64_t rmax, double *sv, int *perm, int *out_i, double *out_d, const int *skip) {
  hipLaunchKernelGGL(k_sort_chop, dim3(1), dim3(1024), (size_t)kpad * sizeof(double), s, Y,
                     (long)ldy, (int)k, (int)kpad, ep, (long)rmax, sv, perm, out_i, out_d, skip);
  TTG_HIP(hipGetLastError());
}

// No-truncation test of a bond: with W^T = Q R (k x k R), sigma_min(W) >=
// 1/||R^{-1}||_F, and my_chop2(s, ep ||s||) keeps all k values whenever
// sigma_min^2 >= (ep ||s||)^2 (core/my_chop2.m:25-31: the first reversed
// cumulative sum is sigma_min^2, compared with strict <).  skip = 1 if the
// bound clears the threshold with a factor 2, rmax does not bind and R != 0.
// skip[0] = 1 when every singular value of the bond clears my_chop2's
// threshold ep ||s|| (core/my_chop2.m:28-34): with a = ||R||_F = ||s|| and x
// an upper bound of ||R^{-1}||_2, sigma_min >= 1/x; the margin 1.05 covers
// R = R2 R1 (||R2 - I|| <= 5e-6) and the rounding of the norms.  x is the
// smaller of ||R^{-1}||_F (nrmX) and, when pw is given, ||(M^T M)^8||_F^(1/16)
// from the normalised squarings' norms pw[0..3]:
// ||B^8||_F = pw0^8 pw1^4 pw2^2 pw3 (B = M^T M).
__global__ void k_notrunc_decide(const double *nrmR, const double *nrmX, const double *pw, double ep, long k,
                                 long rmax, int *skip) {
  if (threadIdx.x == 0) {
    const double a = nrmR[0];
    double x = nrmX[0];
    if (pw && pw[0] > 0.0 && pw[1] > 0.0 && pw[2] > 0.0 && pw[3] > 0.0) {
      const double lb = (8.0 * log(pw[0]) + 4.0 * log(pw[1]) + 2.0 * log(pw[2]) + log(pw[3])) / 16.0;
      const double b = exp(lb);
      if (isfinite(b) && b > 0.0) x = fmin(x, b);
    }
    skip[0] = (k <= rmax && a > 0.0 && x > 0.0 && isfinite(x) && 1.0 / x >= 1.05 * ep * a) ? 1 : 0;
  }
}
void notrunc_decide(hipStream_t s, const double *nrmR, const double *nrmX, const double *pw, double ep,
                    int64_t k, int64_t rmax, int *skip) {
  hipLaunchKernelGGL(k_notrunc_decide, dim3(1), dim3(64), 0, s, nrmR, nrmX, pw, ep, (long)k, (long)rmax, skip);
  TTG_HIP(hipGetLastError());
}

template <class T>
__global__ void k_gather_cols(const double *__restrict__ Y, long ldy, long row0, long rows,
                              const int *__restrict__ perm, long r1, T *__restrict__ dst, long ldd,
                              int trans) {
  const long total = rows * r1;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long i = idx % rows, c = idx / rows;
    const double v = Y[row0 + i + (long)perm[c] * ldy];
    if (trans) dst[c + i * r1] = (T)v;
    else dst[i + c * ldd] = (T)v;
  }
}
template <class T>
void gather_cols(hipStream_t s, const double *Y, int64_t ldy, int64_t row0, int64_t rows,
                 const int *perm, int64_t r1, T *dst, int64_t ldd) {
  if (rows * r1 <= 0) return;
  const int blocks = (int)std::min<int64_t>((rows * r1 + 255) / 256, 4096);
  hipLaunchKernelGGL((k_gather_cols<T>), dim3(blocks), dim3(256), 0, s, Y, (long)ldy, (long)row0,
                     (long)rows, perm, (long)r1, dst, (long)ldd, 0);
  TTG_HIP(hipGetLastError());
}
template <class T>
void gather_cols_t(hipStream_t s, const double *Y, int64_t ldy, int64_t row0, int64_t rows,
                   const int *perm, int64_t r1, T *dst) {
  if (rows * r1 <= 0) return;
  const int blocks = (int)std::min<int64_t>((rows * r1 + 255) / 256, 4096);
  hipLaunchKernelGGL((k_gather_cols<T>), dim3(blocks), dim3(256), 0, s, Y, (long)ldy, (long)row0,
                     (long)rows, perm, (long)r1, dst, (long)r1, 1);
  TTG_HIP(hipGetLastError());
}
template void gather_cols<double>(hipStream_t, const double *, int64_t, int64_t, int64_t, const int *, int64_t, double *, int64_t);
template void gather_cols<float>(hipStream_t, const double *, int64_t, int64_t, int64_t, const int *, int64_t, float *, int64_t);
template void gather_cols_t<double>(hipStream_t, const double *, int64_t, int64_t, int64_t, const int *, int64_t, double *);
template void gather_cols_t<float>(hipStream_t, const double *, int64_t, int64_t, int64_t, const int *, int64_t, float *);

// ---------------------------------------------------------------------------
// final normalisation + balance (round.m:94-107) and packing
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_balance_factor(int d, double *nrm, double *scales) {
  __shared__ double red[16];
  (void)red;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < d; ++i) s += log(fabs(nrm[i]));
    const double nrm0 = exp(s / d);
    scales[0] = (nrm[0] != 0.0) ? nrm0 / nrm[0] : nrm0;
    scales[1] = nrm0;
  }
}
template <class TI>
__global__ void k_pack(int d, TI *const *src, const long *sizes, const long *ofs,
                       const double *scales, double *out) {
  const int kcore = blockIdx.y;
  const long sz = sizes[kcore];
  const TI *p = src[kcore];
  const double sc = (kcore == 0) ? scales[0] : scales[1];
  double *o = out + ofs[kcore];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < sz; i += (long)gridDim.x * blockDim.x)
    o[i] = sc * (double)p[i];
}
template <class TI>
void finalize_balance(hipStream_t s, int64_t d, TI *const *src_dev, const int64_t *sizes_dev,
                      const int64_t *ofs_dev, int64_t size0, double *nrm_dev, double *out,
                      int64_t total) {
  (void)total;
  // nrm[0] = ||core_0||_F is computed by the caller into nrm_dev[0]
  double *scales = nrm_dev + d;
  hipLaunchKernelGGL(k_balance_factor, dim3(1), dim3(64), 0, s, (int)d, nrm_dev, scales);
  TTG_HIP(hipGetLastError());
  (void)size0;
  hipLaunchKernelGGL((k_pack<TI>), dim3(256, (unsigned)d), dim3(256), 0, s, (int)d, src_dev,
                     (const long *)sizes_dev, (const long *)ofs_dev, scales, out);
  TTG_HIP(hipGetLastError());
}
template void finalize_balance<double>(hipStream_t, int64_t, double *const *, const int64_t *, const int64_t *, int64_t, double *, double *, int64_t);
template void finalize_balance<float>(hipStream_t, int64_t, float *const *, const int64_t *, const int64_t *, int64_t, double *, double *, int64_t);

// ---------------------------------------------------------------------------
// mtimes core (@tt_matrix/mtimes.m:38-65 layout)
// Y[(a + ra0*al) + ra0*rx0*(j + n*(b + ra1*be))] = sum_kk A[a,j,kk,b] X[al,kk,be]
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_mtimes_core(long ra0, long n, long m, long ra1, const double *__restrict__ A,
                              long rx0, long rx1, const double *__restrict__ X, T *__restrict__ Y) {
  const long r0 = ra0 * rx0, r1 = ra1 * rx1;
  const long total = r0 * n * r1;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long row = idx % r0;
    long rest = idx / r0;
    const long j = rest % n;
    const long col = rest / n;
    const long a = row % ra0, al = row / ra0;
    const long bb = col % ra1, be = col / ra1;
    double s = 0.0;
    for (long kk = 0; kk < m; ++kk)
      s += A[a + ra0 * (j + n * (kk + m * bb))] * X[al + rx0 * (kk + m * be)];
    Y[idx] = (T)s;
  }
}
template <class T>
void mtimes_core(hipStream_t s, int64_t ra0, int64_t n, int64_t m, int64_t ra1, const double *A,
                 int64_t rx0, int64_t rx1, const double *X, T *Y) {
  const int64_t total = ra0 * rx0 * n * ra1 * rx1;
  if (total <= 0) return;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 16384);
  hipLaunchKernelGGL((k_mtimes_core<T>), dim3(blocks), dim3(256), 0, s, (long)ra0, (long)n,
                     (long)m, (long)ra1, A, (long)rx0, (long)rx1, X, Y);
  TTG_HIP(hipGetLastError());
}
template void mtimes_core<double>(hipStream_t, int64_t, int64_t, int64_t, int64_t, const double *, int64_t, int64_t, const double *, double *);

// All d product cores in one launch (unfused A*x), y(row, j, b, be) =
// sum_q A(a, j, q, b) X(al, q, be) with row = a + ra0 al and output column
// j + n (b + ra1 be) (mtimes.m's order).  Block (x, y): core y, unit x =
// (group of MV_BE consecutive be, 64-row block); thread (row, g): one row and
// the columns (j + n b) = g mod 4.  For a fixed row and column the M entries
// A(a, j, :, b) are the same for every be, so the thread keeps them in
// registers (up to 32 per pass) and, per be, loads only its M values of X:
// about one load per output instead of M + 1.  With the mode size a
// compile-time constant the loads of a pass are issued together; 64
// consecutive rows per wave make each column's store one contiguous 512-byte
// run.  The table's pointers are global memory (global, not flat, loads and
// stores).  No barrier: threads past the last row leave at once.  Sizes are
// checked below 2^31 by the caller (32-bit indices).
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
constexpr int MV_BE = 8;  // be per block
template <int M>
__device__ __forceinline__ void mtimes_cols(gcdouble *A, gcdouble *X, gdouble *Y, unsigned ra0, unsigned nn,
                                            unsigned mm, unsigned rx0, unsigned r0, unsigned cnt, unsigned g,
                                            unsigned nbe, unsigned xs, size_t ys) {
  const unsigned qs = ra0 * nn;  // stride of the contracted index in A
  if (M > 0) {
    constexpr int MM = M > 0 ? M : 1;
    constexpr int NC = 32 / MM;  // columns per pass (A entries kept: NC M <= 32)
    for (unsigned cb = g; cb < cnt; cb += 4 * NC) {
      double av[NC][MM];
#pragma unroll
      for (int u = 0; u < NC; ++u) {
        const unsigned cl = min(cb + 4u * u, cnt - 1);  // past the end: repeat the last column, store nothing
        const unsigned j = cl % nn, b = cl / nn;
        gcdouble *ap = A + ra0 * (j + nn * mm * b);  // A(a, j, q, b) at ap[ra0 n q]
#pragma unroll
        for (int q = 0; q < M; ++q) av[u][q] = ap[qs * q];
      }
      for (unsigned e = 0; e < nbe; ++e) {
        double xv[MM];
#pragma unroll
        for (int q = 0; q < M; ++q) xv[q] = X[xs * e + rx0 * q];
#pragma unroll
        for (int u = 0; u < NC; ++u) {
          double acc = 0.0;
#pragma unroll
          for (int q = 0; q < M; ++q) acc = __builtin_fma(av[u][q], xv[q], acc);
          if (cb + 4u * u < cnt) Y[ys * e + (size_t)r0 * (cb + 4u * u)] = acc;
        }
      }
    }
  } else {
    for (unsigned e = 0; e < nbe; ++e)
      for (unsigned cl = g; cl < cnt; cl += 4) {
        const unsigned j = cl % nn, b = cl / nn;
        gcdouble *ap = A + ra0 * (j + nn * mm * b);
        double acc = 0.0;
        for (unsigned q = 0; q < mm; ++q) acc = __builtin_fma(ap[qs * q], X[xs * e + rx0 * q], acc);
        Y[ys * e + (size_t)r0 * cl] = acc;
      }
  }
}
__global__ __launch_bounds__(256) void k_mtimes_all(const MvCore *__restrict__ tab) {
  const MvCore c = tab[blockIdx.y];
  const unsigned ra0 = (unsigned)c.ra0, rx0 = (unsigned)c.rx0, nn = (unsigned)c.n, mm = (unsigned)c.m;
  const unsigned ra1 = (unsigned)c.ra1, rx1 = (unsigned)c.rx1;
  const unsigned r0 = ra0 * rx0, rb = (r0 + 63) / 64, nbg = (rx1 + MV_BE - 1) / MV_BE;
  if (blockIdx.x >= rb * nbg) return;
  const unsigned bg = blockIdx.x / rb;
  const unsigned row = (blockIdx.x - bg * rb) * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
  if (row >= r0) return;
  const unsigned be0 = bg * MV_BE, nbe = min((unsigned)MV_BE, rx1 - be0);
  const unsigned a = row % ra0, al = row / ra0;
  const unsigned cnt = nn * ra1;
  const unsigned xs = rx0 * mm;                   // X(al, q, be) at X[xs be + rx0 q]
  const size_t ys = (size_t)r0 * nn * ra1;        // column (j + n b) of be at Y[ys be + r0 (j + n b)]
  gcdouble *A = (gcdouble *)c.A + a;
  gcdouble *X = (gcdouble *)c.X + al + (size_t)xs * be0;
  gdouble *Y = (gdouble *)c.Y + row + ys * be0;
  switch (mm) {
    case 1: mtimes_cols<1>(A, X, Y, ra0, nn, mm, rx0, r0, cnt, g, nbe, xs, ys); break;
    case 2: mtimes_cols<2>(A, X, Y, ra0, nn, mm, rx0, r0, cnt, g, nbe, xs, ys); break;
    case 3: mtimes_cols<3>(A, X, Y, ra0, nn, mm, rx0, r0, cnt, g, nbe, xs, ys); break;
    case 4: mtimes_cols<4>(A, X, Y, ra0, nn, mm, rx0, r0, cnt, g, nbe, xs, ys); break;
    case 8: mtimes_cols<8>(A, X, Y, ra0, nn, mm, rx0, r0, cnt, g, nbe, xs, ys); break;
    default: mtimes_cols<0>(A, X, Y, ra0, nn, mm, rx0, r0, cnt, g, nbe, xs, ys); break;
  }
}
int64_t mtimes_units(int64_t r0, int64_t rx1) { return (r0 + 63) / 64 * ((rx1 + MV_BE - 1) / MV_BE); }
void mtimes_all(hipStream_t s, const MvCore *tab_dev, int ncores, int64_t max_units, double bytes) {
  if (ncores <= 0 || max_units <= 0) return;
  ProfScope prof(KC_MTIMES, s, 0.0, bytes, true);
  TTG_TIMED_LAUNCH(prof, k_mtimes_all, dim3((unsigned)max_units, (unsigned)ncores), dim3(256), 0, s, tab_dev);
  TTG_HIP(hipGetLastError());
}

// C (M x N) = R (M x K) * P, where P is the (ra0 rx0) x (n ra1 rx1) unfolding of
// the product core of TT-matrix core A (ra0 x n x m x ra1) and TT core X
// (rx0 x m x rx1), in the layout of @tt_matrix/mtimes.m:49-64 (row a + ra0*al,
// column j + n*(b + ra1*be)).  P is built one 32 x 64 tile at a time in LDS
// (m FMAs per entry, A and X stay in cache) and never exists in HBM.  FP64
// MFMA on the transposed product (R P)^T so that 16 lanes store 16
// consecutive rows of C.
constexpr int MP_TM = 32, MP_TN = 64, MP_TK = 32;
constexpr int MP_RLD = MP_TM + 16, MP_PLD = MP_TN + 16;  // lanes 16-31 of a read land on the other 32 banks
__global__ __launch_bounds__(256) void k_gemm_mtimes(int M, int N, int K, const double *__restrict__ R,
                                                     long ldr, int ra0, int nn, int mm, int ra1,
                                                     const double *__restrict__ A, int rx0,
                                                     const double *__restrict__ X, double *__restrict__ C,
                                                     long ldc) {
  __shared__ __align__(16) double Rs[MP_TK * MP_RLD];  // Rs[k][m] = R(m0 + m, k0 + k)
  __shared__ __align__(16) double Ps[MP_TK * MP_PLD];  // Ps[k][n] = P(k0 + k, n0 + n)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = wv & 1, wc = wv >> 1;  // wave tile: rows wr*16 +16, columns wc*32 +32
  const int m0 = blockIdx.x * MP_TM, n0 = blockIdx.y * MP_TN;
  const long qa = (long)ra0 * nn, qx = rx0;  // strides of the contracted index in A and X
  v4d acc[2] = {v4d{0, 0, 0, 0}, v4d{0, 0, 0, 0}};
  for (int k0 = 0; k0 < K; k0 += MP_TK) {
    for (int idx = tid; idx < MP_TM * MP_TK; idx += 256) {
      const int mi = idx % MP_TM, ki = idx / MP_TM;
      const int gm = m0 + mi, gk = k0 + ki;
      Rs[ki * MP_RLD + mi] = (gm < M && gk < K) ? R[gm + (long)gk * ldr] : 0.0;
    }
    for (int idx = tid; idx < MP_TK * MP_TN; idx += 256) {
      const int ni = idx % MP_TN, ki = idx / MP_TN;
      const int gk = k0 + ki, gn = n0 + ni;
      double v = 0.0;
      if (gk < K && gn < N) {
        const int a = gk % ra0, al = gk / ra0;
        const int j = gn % nn, rest = gn / nn;
        const int b = rest % ra1, be = rest / ra1;
        const double *ap = A + a + (long)ra0 * (j + (long)nn * mm * b);
        const double *xp = X + al + (long)rx0 * ((long)mm * be);
        for (int q = 0; q < mm; ++q) v = __builtin_fma(ap[qa * q], xp[qx * q], v);
      }
      Ps[ki * MP_PLD + ni] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < MP_TK; kk += 4) {
      const int kr = kk + (lane >> 4);
      const double rv = Rs[kr * MP_RLD + wr * 16 + (lane & 15)];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const double pv = Ps[kr * MP_PLD + wc * 32 + t * 16 + (lane & 15)];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(pv, rv, acc[t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // acc[t][v] of lane l = (R P)(m0 + wr*16 + l%16, n0 + wc*32 + t*16 + l/16 + 4v)
  const int gm = m0 + wr * 16 + (lane & 15);
  if (gm < M) {
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gn = n0 + wc * 32 + t * 16 + (lane >> 4) + 4 * v;
        if (gn < N) C[gm + (long)gn * ldc] = acc[t][v];
      }
  }
}

void gemm_mtimes(hipStream_t s, int64_t M, const double *R, int64_t ldr, int64_t ra0, int64_t n,
                 int64_t m, int64_t ra1, const double *A, int64_t rx0, int64_t rx1, const double *X,
                 double *C, int64_t ldc) {
  const int64_t K = ra0 * rx0, N = n * ra1 * rx1;
  if (M <= 0 || N <= 0) return;
  TTG_CHECK(K < (1LL << 31) && N < (1LL << 31) && M < (1LL << 31), TTG_BAD_ARG, "product core too large");
  TTG_CHECK(ldr >= M && ldc >= M, TTG_BAD_ARG, "bad leading dimension");
  const int64_t gy = (N + MP_TN - 1) / MP_TN;
  TTG_CHECK(gy <= 65535, TTG_BAD_ARG, "product core has too many columns");
  // algorithmic flops of R * P; bytes: R, C and the two factor cores (P is never stored)
  ProfScope prof(KC_GEMM, s, 2.0 * (double)M * N * K,
                 8.0 * ((double)M * K + (double)M * N + (double)ra0 * n * m * ra1 + (double)rx0 * m * rx1), true);
  TTG_TIMED_LAUNCH(prof, k_gemm_mtimes, dim3((unsigned)((M + MP_TM - 1) / MP_TM), (unsigned)gy), dim3(256), 0, s,
                   (int)M, (int)N, (int)K, R, (long)ldr, (int)ra0, (int)n, (int)m, (int)ra1, A, (int)rx0, X, C,
                   (long)ldc);
  TTG_HIP(hipGetLastError());
}
template void mtimes_core<float>(hipStream_t, int64_t, int64_t, int64_t, int64_t, const double *, int64_t, int64_t, const double *, float *);

}  // namespace ttg
