// R factor of the QR of a very tall, thin matrix A^T, with A (m x M,
// column-major, M >> m) stored as given: TSQR (a reduction tree of
// Householder QRs).  Used by the TT-SVD, whose unfoldings are wide,
// numerically low-rank and far too long for one-workgroup Householder sweeps.
//
// Level: the M columns of A are cut into blocks of B >= 2m (contiguous in
// memory); workgroup b copies block b transposed (B x m) into its slice of a
// scratch buffer, runs unblocked Householder (LAPACK dlarfg conventions) on
// it, and writes R_b^T into columns [b m, (b + 1) m) of the next level's
// A (m x nb m), so the next level sees [R_0; R_1; ...]^T in the same wide
// layout.  The last level has one block and leaves R^T (m x m).  Every
// level is backward stable (Householder), so R's singular values are A's to
// O(u ||A||), rank deficiency included.
#include "ttg_internal.h"
#include "ttg_device.h"

namespace ttg {
namespace {

__global__ __launch_bounds__(256) void k_tsqr_level(const double *__restrict__ A, int m, long M, int B,
                                                     double *__restrict__ Wall, double *__restrict__ out) {
  __shared__ double red[8];
  __shared__ double sh_tau, sh_scale;
  const long col0 = (long)blockIdx.x * B;
  const int Bb = (int)min((long)B, M - col0);
  double *W = Wall + (long)blockIdx.x * B * m;  // Bb x m, ld Bb
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  // W = A(:, col0 : col0 + Bb)^T (reads along A's columns, coalesced)
  for (long e = tid; e < (long)Bb * m; e += blockDim.x) {
    const int j = (int)(e % m), i = (int)(e / m);
    W[i + (long)Bb * j] = A[(long)j + (long)m * (col0 + i)];
  }
  __syncthreads();
  const int kq = min(Bb, m);
  for (int j = 0; j < kq; ++j) {
    double *aj = W + (long)j * Bb;
    double ss = 0.0;
    for (int i = j + 1 + tid; i < Bb; i += blockDim.x) ss += aj[i] * aj[i];
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
    }
    __syncthreads();
    const double tj = sh_tau, sc = sh_scale;
    if (tj != 0.0)
      for (int i = j + 1 + tid; i < Bb; i += blockDim.x) aj[i] *= sc;
    __syncthreads();
    if (tj != 0.0) {
      for (int c = j + 1 + wv; c < m; c += nw) {
        double *ac = W + (long)c * Bb;
        double w = (lane == 0) ? ac[j] : 0.0;
        for (int i = j + 1 + lane; i < Bb; i += 64) w += aj[i] * ac[i];
        w = wave_sum(w) * tj;
        if (lane == 0) ac[j] -= w;
        for (int i = j + 1 + lane; i < Bb; i += 64) ac[i] -= w * aj[i];
      }
    }
    __syncthreads();
  }
  // out(jj, b m + i) = R_b(i, jj): upper trapezoid, zero rows below kq
  for (long e = tid; e < (long)m * m; e += blockDim.x) {
    const int jj = (int)(e % m), i = (int)(e / m);
    out[jj + (long)m * ((long)blockIdx.x * m + i)] = (i <= jj && i < kq) ? W[i + (long)Bb * jj] : 0.0;
  }
}

int64_t tsqr_block(int64_t m) { return std::max<int64_t>(256, 2 * m); }

}  // namespace

int64_t tsqr_ws_elems(int64_t m, int64_t M, int64_t *level_elems) {
  const int64_t B = tsqr_block(m);
  const int64_t nb = (M + B - 1) / B;
  if (level_elems) *level_elems = std::max<int64_t>(nb * m * m, m * m);
  return nb * B * m;
}

void tsqr_rt(hipStream_t s, const double *A, int64_t m, int64_t M, double *W, double *Sa, double *Sb,
             double *Rt) {
  TTG_CHECK(m >= 1 && M >= 1 && m <= 4096, TTG_BAD_ARG, "tsqr: bad shape");
  const int64_t B = tsqr_block(m);
  const double *cur = A;
  int64_t curM = M;
  int flip = 0;
  for (;;) {
    const int64_t nb = (curM + B - 1) / B;
    TTG_CHECK(nb < (int64_t(1) << 31), TTG_BAD_ARG, "tsqr: too many blocks");
    double *out = nb == 1 ? Rt : (flip ? Sb : Sa);
    hipLaunchKernelGGL(k_tsqr_level, dim3((unsigned)nb), dim3(256), 0, s, cur, (int)m, (long)curM, (int)B, W,
                       out);
    TTG_HIP(hipGetLastError());
    if (nb == 1) break;
    cur = out;
    curM = nb * m;
    flip ^= 1;
  }
}

}  // namespace ttg
