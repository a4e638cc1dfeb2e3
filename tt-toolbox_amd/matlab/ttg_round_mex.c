/* MEX gateway: [r_out, core_out] = ttg_round_mex(n, r, core, eps, rmax, prec)
 *
 * Replaces the body of @tt_tensor/round.m (with core/my_chop2.m) when the
 * compiled gateway is on the MATLAB path (dispatch guard shown in the
 * INTEGRATION notes; the pattern is the reference's own fmex/tt_read.c:105
 * gateway and the exist() test of cross/localcross.m:4-12).
 *   n: d x 1, r: (d+1) x 1 (doubles, as tt.n / tt.r), core: tt.core (real),
 *   eps: scalar, rmax: [] | scalar | (d+1) x 1 (round.m:30-32), prec: 0 plain
 *   FP64, 1 mixed (error budget 1e-6).
 * Outputs: the new ranks (d+1 x 1) and the packed core (column).
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

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  if (nrhs < 4) mexErrMsgIdAndTxt("ttgpu:args", "usage: ttg_round_mex(n, r, core, eps[, rmax, prec])");
  if (mxIsComplex(prhs[2])) mexErrMsgIdAndTxt("ttgpu:args", "complex cores are not supported");
  const mwSize d = mxGetNumberOfElements(prhs[0]);
  if (mxGetNumberOfElements(prhs[1]) != d + 1) mexErrMsgIdAndTxt("ttgpu:args", "r must have d+1 entries");
  const double *nd = mxGetPr(prhs[0]), *rd = mxGetPr(prhs[1]), *core = mxGetPr(prhs[2]);
  const double eps = mxGetScalar(prhs[3]);
  const int prec = (nrhs > 5) ? (int)mxGetScalar(prhs[5]) : TTG_PREC_FP64;
  int64_t *n = (int64_t *)mxMalloc(d * sizeof(int64_t));
  int64_t *r = (int64_t *)mxMalloc((d + 1) * sizeof(int64_t));
  int64_t *r_out = (int64_t *)mxMalloc((d + 1) * sizeof(int64_t));
  int64_t *rmax = NULL;
  for (mwSize k = 0; k < d; ++k) n[k] = (int64_t)nd[k];
  for (mwSize k = 0; k <= d; ++k) r[k] = (int64_t)rd[k];
  if (nrhs > 4 && !mxIsEmpty(prhs[4])) { /* round.m:30-32: a scalar becomes a d+1 vector */
    const mwSize nr = mxGetNumberOfElements(prhs[4]);
    const double *rm = mxGetPr(prhs[4]);
    if (nr != 1 && nr != d + 1) mexErrMsgIdAndTxt("ttgpu:args", "rmax must be a scalar or have d+1 entries");
    rmax = (int64_t *)mxMalloc((d + 1) * sizeof(int64_t));
    for (mwSize k = 0; k <= d; ++k) rmax[k] = (int64_t)(nr == 1 ? rm[0] : rm[k]);
  }
  if (!g_ctx) {
    if (ttg_ctx_create(0, NULL, &g_ctx) != TTG_OK) mexErrMsgIdAndTxt("ttgpu:device", "no gfx950 device");
    mexAtExit(close_ctx);
  }
  int64_t cap = 0, len = 0;
  if (ttg_round_capacity((int64_t)d, n, r, &cap) != TTG_OK) mexErrMsgIdAndTxt("ttgpu:args", "bad sizes or ranks");
  mxArray *out = mxCreateDoubleMatrix((mwSize)(cap > 0 ? cap : 1), 1, mxREAL);
  const ttg_status st = ttg_round(g_ctx, (int64_t)d, n, r, core, eps, rmax, prec, r_out, mxGetPr(out), cap, &len,
                                  NULL);
  if (st == TTG_NO_CONVERGENCE) /* round.m:64-74 warns (svd -> svds) and carries on */
    mexWarnMsgIdAndTxt("ttgpu:svd", "%s", ttg_last_error(g_ctx));
  else if (st != TTG_OK)
    mexErrMsgIdAndTxt("ttgpu:round", "%s", ttg_last_error(g_ctx));
  mxSetM(out, (mwSize)len); /* shrink to the used length */
  plhs[0] = mxCreateDoubleMatrix(d + 1, 1, mxREAL);
  for (mwSize k = 0; k <= d; ++k) mxGetPr(plhs[0])[k] = (double)r_out[k];
  if (nlhs > 1)
    plhs[1] = out;
  else
    mxDestroyArray(out);
  mxFree(n);
  mxFree(r);
  mxFree(r_out);
  if (rmax) mxFree(rmax);
}
