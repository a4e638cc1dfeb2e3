/* MEX gateway: [ry, ycore] = ttg_mtimes_mex(n, m, ra, acore, rx, xcore)
 *              [ry, ycore] = ttg_mtimes_mex(n, m, ra, acore, rx, xcore, eps, rmax, prec)
 *
 * Replaces the core loop of @tt_matrix/mtimes.m:38-65 (A*x, no rounding:
 * ry = ra.*rx) and, with eps given, round(A*x, eps, rmax) as one call (the
 * product cores are built on chip inside the first rounding sweep).
 *   n, m: d x 1 (A.n, A.m), ra: (d+1) x 1 (A.tt.r), acore: A.tt.core,
 *   rx: (d+1) x 1, xcore: x.core (real doubles).
 * Build: make -C tt-toolbox_amd/matlab (needs MATLAB's mex and mex.h). */
#include "mex.h"
#include "ttgpu.h"

static ttg_ctx *g_ctx = NULL;
static void close_ctx(void) {
  if (g_ctx) {
    ttg_ctx_destroy(g_ctx);
    g_ctx = NULL;
  }
}

static int64_t *to_i64(const mxArray *a, mwSize count) {
  if (mxGetNumberOfElements(a) != count) mexErrMsgIdAndTxt("ttgpu:args", "size/rank vector of the wrong length");
  int64_t *v = (int64_t *)mxMalloc((count > 0 ? count : 1) * sizeof(int64_t));
  const double *p = mxGetPr(a);
  for (mwSize k = 0; k < count; ++k) v[k] = (int64_t)p[k];
  return v;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  if (nrhs != 6 && nrhs < 7)
    mexErrMsgIdAndTxt("ttgpu:args", "usage: ttg_mtimes_mex(n, m, ra, acore, rx, xcore[, eps, rmax, prec])");
  if (mxIsComplex(prhs[3]) || mxIsComplex(prhs[5])) mexErrMsgIdAndTxt("ttgpu:args", "complex cores are not supported");
  const mwSize d = mxGetNumberOfElements(prhs[0]);
  int64_t *n = to_i64(prhs[0], d), *m = to_i64(prhs[1], d), *ra = to_i64(prhs[2], d + 1),
          *rx = to_i64(prhs[4], d + 1);
  int64_t *ry = (int64_t *)mxMalloc((d + 1) * sizeof(int64_t)), *rmax = NULL;
  if (!g_ctx) {
    if (ttg_ctx_create(0, NULL, &g_ctx) != TTG_OK) mexErrMsgIdAndTxt("ttgpu:device", "no gfx950 device");
    mexAtExit(close_ctx);
  }
  int64_t cap = 0, len = 0;
  for (mwSize k = 0; k <= d; ++k) ry[k] = ra[k] * rx[k];
  ttg_status st;
  mxArray *out;
  if (nrhs == 6) {
    for (mwSize k = 0; k < d; ++k) cap += ry[k] * n[k] * ry[k + 1];
    out = mxCreateDoubleMatrix((mwSize)(cap > 0 ? cap : 1), 1, mxREAL);
    st = ttg_mtimes(g_ctx, (int64_t)d, n, m, ra, mxGetPr(prhs[3]), rx, mxGetPr(prhs[5]), ry, mxGetPr(out), cap,
                    &len);
  } else {
    const double eps = mxGetScalar(prhs[6]);
    const int prec = nrhs > 8 ? (int)mxGetScalar(prhs[8]) : TTG_PREC_FP64;
    if (nrhs > 7 && !mxIsEmpty(prhs[7])) {
      const mwSize nr = mxGetNumberOfElements(prhs[7]);
      const double *rm = mxGetPr(prhs[7]);
      if (nr != 1 && nr != d + 1) mexErrMsgIdAndTxt("ttgpu:args", "rmax must be a scalar or have d+1 entries");
      rmax = (int64_t *)mxMalloc((d + 1) * sizeof(int64_t));
      for (mwSize k = 0; k <= d; ++k) rmax[k] = (int64_t)(nr == 1 ? rm[0] : rm[k]);
    }
    if (ttg_round_capacity((int64_t)d, n, ry, &cap) != TTG_OK) mexErrMsgIdAndTxt("ttgpu:args", "bad sizes or ranks");
    out = mxCreateDoubleMatrix((mwSize)(cap > 0 ? cap : 1), 1, mxREAL);
    st = ttg_mtimes_round(g_ctx, (int64_t)d, n, m, ra, mxGetPr(prhs[3]), rx, mxGetPr(prhs[5]), eps, rmax, prec, ry,
                          mxGetPr(out), cap, &len, NULL);
  }
  if (st == TTG_NO_CONVERGENCE)
    mexWarnMsgIdAndTxt("ttgpu:svd", "%s", ttg_last_error(g_ctx));
  else if (st != TTG_OK)
    mexErrMsgIdAndTxt("ttgpu:mtimes", "%s", ttg_last_error(g_ctx));
  mxSetM(out, (mwSize)len);
  plhs[0] = mxCreateDoubleMatrix(d + 1, 1, mxREAL);
  for (mwSize k = 0; k <= d; ++k) mxGetPr(plhs[0])[k] = (double)ry[k];
  if (nlhs > 1)
    plhs[1] = out;
  else
    mxDestroyArray(out);
  mxFree(n);
  mxFree(m);
  mxFree(ra);
  mxFree(rx);
  mxFree(ry);
  if (rmax) mxFree(rmax);
}
