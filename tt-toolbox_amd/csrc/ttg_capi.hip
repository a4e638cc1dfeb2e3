// extern "C" entry points of libttgpu.so (declared in include/ttgpu.h).
#include "ttg_internal.h"
#include <cstring>
#include <algorithm>
#include <mutex>

namespace ttg {
ttg_status round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                     const double *dcore, double eps, const int64_t *rmax, int prec,
                     int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                     ttg_info *info);
ttg_status round_capacity(int64_t d, const int64_t *n, const int64_t *r, int64_t *cap);
// sizes/ranks of A*x: d >= 1, every mode size and rank >= 1 (throws TTG_BAD_ARG)
void validate_mv(int64_t d, const int64_t *n, const int64_t *m, const int64_t *ra, const int64_t *rx);
ttg_status mtimes_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                      const int64_t *ra, const double *dacore, const int64_t *rx,
                      const double *dxcore, int64_t *ry, double *dycore, int64_t cap,
                      int64_t *len_out);
ttg_status mtimes_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                            const int64_t *ra, const double *dacore, const int64_t *rx,
                            const double *dxcore, double eps, const int64_t *rmax, int prec,
                            int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                            ttg_info *info);
ttg_status round_batched_dev(ttg_ctx *ctx, int64_t count, const int64_t *d, const int64_t *nofs,
                             const int64_t *n, const int64_t *rofs, const int64_t *r,
                             const int64_t *cofs, const double *dcore, double eps, const int64_t *rmax,
                             int prec, int64_t *r_out, const int64_t *cofs_out, double *dcore_out,
                             ttg_info *info);
ttg_status plus_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx, double a,
                    const int64_t *ry, const double *dy, double b, int64_t *rz_out, double *dz, int64_t cap,
                    int64_t *len_out);
ttg_status plus_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx,
                          double a, const int64_t *ry, const double *dy, double b, double eps,
                          const int64_t *rmax, int prec, int64_t *r_out, double *dcore_out, int64_t cap,
                          int64_t *len_out, ttg_info *info);
ttg_status norm_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r, const double *dcore,
                    double *nrm);
ttg_status plus_norm_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx,
                         double a, const int64_t *ry, const double *dy, double b, double *nrm);
ttg_status dot_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx,
                   const int64_t *ry, const double *dy, double *value);
ttg_status full_to_tt_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const double *dfull, double eps,
                          const int64_t *rmax, int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                          ttg_info *info);
ttg_status sdv_info(const char *path, int64_t *d, int64_t *core_len, int *dtype);
ttg_status sdv_read(const char *path, int64_t dcap, int64_t *d, int64_t *n, int64_t *r, double *core, int64_t cap,
                    int64_t *len_out);
ttg_status sdv_read_dev(ttg_ctx *ctx, const char *path, int64_t dcap, int64_t *d, int64_t *n, int64_t *r,
                        double *dcore, int64_t cap, int64_t *len_out);
ttg_status sdv_write(const char *path, int64_t d, const int64_t *n, const int64_t *r, const double *core,
                     int precision);
ttg_status sdv_write_dev(ttg_ctx *ctx, const char *path, int64_t d, const int64_t *n, const int64_t *r,
                         const double *dcore, int precision);
template <class T>
ttg_status adjoint_completion(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                              const T *dcore, int64_t S, const int32_t *didx, const T *dvals,
                              double *loss, T *dgrad);
template <class T>
ttg_status adjoint_bilinear(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                            const int64_t *ry, const T *dycore, const int64_t *ra,
                            const T *dacore, const int64_t *rx, const T *dxcore, double *value,
                            T *dgx, T *dgy, void *ext_ws, size_t ext_bytes);

// Scratch of the no-allocation entry point: one copy of this array lives on
// every device the library runs on (until the process ends).  Host code never
// passes the array's name to a kernel (it is a host-side stand-in there); it
// uses the device address hipGetSymbolAddress returns, looked up once per
// device when the first context on that device is created.
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// device address of `scratch` on `device` (the device must be current)
static ttg_status scratch_lookup(int device) {
  if (device < 0 || device >= kMaxDevices) return TTG_NO_DEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return TTG_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    (void)hipGetLastError();
    return TTG_HIP_ERROR;
  }
  g_scratch_ptr[device] = p;
  return TTG_OK;
}
}  // namespace ttg

using namespace ttg;

// Makes ctx->prof visible to the launch helpers for the duration of a call.
struct ProfBind {
  ProfBind(ttg_ctx *c) { ttg::g_prof = c ? &c->prof : nullptr; }
  ~ProfBind() { ttg::g_prof = nullptr; }
};

#define TTG_GUARD_BEGIN \
  ProfBind prof_bind_(ctx); \
  try {
#define TTG_GUARD_END(ctx)                                                    \
  }                                                                           \
  catch (const ttg::Error &e) {                                               \
    if (ctx) (ctx)->last_error = e.what();                                    \
    if (e.status == TTG_HIP_ERROR && ctx) (void)hipGetLastError();            \
    return e.status;                                                          \
  }                                                                           \
  catch (const std::exception &e) {                                           \
    if (ctx) (ctx)->last_error = e.what();                                    \
    return TTG_HIP_ERROR;                                                     \
  }

static int64_t tt_len(int64_t d, const int64_t *n, const int64_t *r) {
  int64_t s = 0;
  for (int64_t i = 0; i < d; ++i) s += r[i] * n[i] * r[i + 1];
  return s;
}

extern "C" {

const char *ttg_version(void) { return "ttgpu 0.1 (gfx950)"; }

ttg_status ttg_ctx_create(int device, void *stream, ttg_ctx **out) {
  if (!out) return TTG_BAD_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return TTG_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return TTG_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return TTG_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return TTG_NO_DEVICE;
  // the no-allocation entry point's scratch address on this device
  if (ttg::scratch_lookup(device) != TTG_OK) return TTG_HIP_ERROR;
  ttg_ctx *ctx = new ttg_ctx();
  ctx->device = device;
  ctx->num_cus = prop.multiProcessorCount;
  if (stream) {
    ctx->stream = reinterpret_cast<hipStream_t>(stream);
    ctx->own_stream = false;
  } else {
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      return TTG_HIP_ERROR;
    }
    ctx->own_stream = true;
  }
  *out = ctx;
  return TTG_OK;
}

ttg_status ttg_ctx_destroy(ttg_ctx *ctx) {
  if (!ctx) return TTG_OK;
  ttg::release_ctx(ctx);
  delete ctx;
  return TTG_OK;
}

const char *ttg_last_error(const ttg_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }
void *ttg_ctx_stream(const ttg_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

ttg_status ttg_profile_begin(ttg_ctx *ctx, int kernel_class) {
  if (!ctx) return TTG_BAD_ARG;
  ctx->prof.kind = kernel_class;
  ctx->prof.used = 0;
  ctx->prof.flops.clear();
  ctx->prof.bytes.clear();
  return TTG_OK;
}

ttg_status ttg_profile_end(ttg_ctx *ctx, int64_t *launches, double *ms_total, double *flops_total,
                           double *bytes_total) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  TTG_HIP(hipStreamSynchronize(ctx->stream));
  double ms = 0.0, fl = 0.0, by = 0.0;
  for (size_t i = 0; i < ctx->prof.used; ++i) {
    float t = 0.f;
    TTG_HIP(hipEventElapsedTime(&t, ctx->prof.pool[2 * i], ctx->prof.pool[2 * i + 1]));
    ms += t;
    fl += ctx->prof.flops[i];
    by += ctx->prof.bytes[i];
  }
  if (launches) *launches = (int64_t)ctx->prof.used;
  if (ms_total) *ms_total = ms;
  if (flops_total) *flops_total = fl;
  if (bytes_total) *bytes_total = by;
  ctx->prof.kind = 0;
  ctx->prof.used = 0;
  return TTG_OK;
  TTG_GUARD_END(ctx)
}

ttg_status ttg_round_capacity(int64_t d, const int64_t *n, const int64_t *r, int64_t *cap) {
  ttg_ctx *ctx = nullptr;
  TTG_GUARD_BEGIN
  if (!cap) return TTG_BAD_ARG;
  return round_capacity(d, n, r, cap);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                         const double *dcore, double eps, const int64_t *rmax, int prec,
                         int64_t *r_out, double *dcore_out, int64_t cap, int64_t *len_out,
                         ttg_info *info) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_CHECK(r_out && dcore && dcore_out, TTG_BAD_ARG, "null pointer argument");
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return round_dev(ctx, d, n, r, dcore, eps, rmax, prec, r_out, dcore_out, cap, len_out, info);
  TTG_GUARD_END(ctx)
}

// Host-pointer form: stage through device buffers owned by this call.
ttg_status ttg_round(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r,
                     const double *core, double eps, const int64_t *rmax, int prec,
                     int64_t *r_out, double *core_out, int64_t cap, int64_t *len_out,
                     ttg_info *info) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_CHECK(core && core_out && r_out && n && r && d >= 1, TTG_BAD_ARG, "null pointer argument");
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  int64_t need = 0;
  ttg_status s = round_capacity(d, n, r, &need);
  if (s != TTG_OK) return s;
  if (cap < need) {
    if (len_out) *len_out = need;
    return TTG_NEED_SPACE;
  }
  const int64_t len = tt_len(d, n, r);
  double *din = nullptr, *dout = nullptr;
  TTG_HIP(hipMalloc(&din, std::max<int64_t>(len, 1) * sizeof(double)));
  hipError_t e2 = hipMalloc(&dout, std::max<int64_t>(need, 1) * sizeof(double));
  if (e2 != hipSuccess) {
    (void)hipFree(din);
    throw Error(TTG_HIP_ERROR, "hipMalloc failed for the output staging buffer");
  }
  ttg_status st = TTG_OK;
  try {
    TTG_HIP(hipMemcpyAsync(din, core, len * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int64_t used = 0;
    st = round_dev(ctx, d, n, r, din, eps, rmax, prec, r_out, dout, need, &used, info);
    if (st == TTG_OK || st == TTG_NO_CONVERGENCE) {
      TTG_HIP(hipMemcpyAsync(core_out, dout, used * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      TTG_HIP(hipStreamSynchronize(ctx->stream));
      if (len_out) *len_out = used;
    }
  } catch (...) {
    (void)hipFree(din);
    (void)hipFree(dout);
    throw;
  }
  (void)hipFree(din);
  (void)hipFree(dout);
  return st;
  TTG_GUARD_END(ctx)
}

ttg_status ttg_round_matrix(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                            const int64_t *r, const double *core, double eps,
                            const int64_t *rmax, int prec, int64_t *r_out, double *core_out,
                            int64_t cap, int64_t *len_out, ttg_info *info) {
  if (!ctx) return TTG_BAD_ARG;
  if (!n || !m || d < 1) return TTG_BAD_ARG;
  std::vector<int64_t> nm(d);
  for (int64_t i = 0; i < d; ++i) nm[i] = n[i] * m[i];
  return ttg_round(ctx, d, nm.data(), r, core, eps, rmax, prec, r_out, core_out, cap, len_out, info);
}

ttg_status ttg_mtimes_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                          const int64_t *ra, const double *dacore, const int64_t *rx,
                          const double *dxcore, int64_t *ry, double *dycore, int64_t cap,
                          int64_t *len_out) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_CHECK(n && m && ra && rx && ry && dacore && dxcore && dycore, TTG_BAD_ARG, "null pointer argument");
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return mtimes_dev(ctx, d, n, m, ra, dacore, rx, dxcore, ry, dycore, cap, len_out);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_mtimes(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                      const int64_t *ra, const double *acore, const int64_t *rx,
                      const double *xcore, int64_t *ry, double *ycore, int64_t cap,
                      int64_t *len_out) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_CHECK(n && m && ra && rx && ry && acore && xcore && ycore && d >= 1, TTG_BAD_ARG, "null pointer argument");
  // sizes are checked before any size arithmetic or staging allocation
  validate_mv(d, n, m, ra, rx);
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  int64_t la = 0, lx = 0, ly = 0;
  for (int64_t i = 0; i < d; ++i) {
    la += ra[i] * n[i] * m[i] * ra[i + 1];
    lx += rx[i] * m[i] * rx[i + 1];
    ly += ra[i] * rx[i] * n[i] * ra[i + 1] * rx[i + 1];
  }
  if (len_out) *len_out = ly;
  if (cap < ly) return TTG_NEED_SPACE;
  double *da = nullptr, *dx = nullptr, *dy = nullptr;
  TTG_HIP(hipMalloc(&da, std::max<int64_t>(la, 1) * sizeof(double)));
  TTG_HIP(hipMalloc(&dx, std::max<int64_t>(lx, 1) * sizeof(double)));
  TTG_HIP(hipMalloc(&dy, std::max<int64_t>(ly, 1) * sizeof(double)));
  ttg_status st = TTG_OK;
  try {
    TTG_HIP(hipMemcpyAsync(da, acore, la * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    TTG_HIP(hipMemcpyAsync(dx, xcore, lx * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    st = mtimes_dev(ctx, d, n, m, ra, da, rx, dx, ry, dy, ly, len_out);
    if (st == TTG_OK) {
      TTG_HIP(hipMemcpyAsync(ycore, dy, ly * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      TTG_HIP(hipStreamSynchronize(ctx->stream));
    }
  } catch (...) {
    (void)hipFree(da);
    (void)hipFree(dx);
    (void)hipFree(dy);
    throw;
  }
  (void)hipFree(da);
  (void)hipFree(dx);
  (void)hipFree(dy);
  return st;
  TTG_GUARD_END(ctx)
}

ttg_status ttg_mtimes_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                                const int64_t *ra, const double *dacore, const int64_t *rx,
                                const double *dxcore, double eps, const int64_t *rmax, int prec,
                                int64_t *r_out, double *dcore_out, int64_t cap,
                                int64_t *len_out, ttg_info *info) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_CHECK(n && m && ra && rx && r_out && dacore && dxcore && dcore_out, TTG_BAD_ARG, "null pointer argument");
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return mtimes_round_dev(ctx, d, n, m, ra, dacore, rx, dxcore, eps, rmax, prec, r_out, dcore_out,
                          cap, len_out, info);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_mtimes_round(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                            const int64_t *ra, const double *acore, const int64_t *rx,
                            const double *xcore, double eps, const int64_t *rmax, int prec,
                            int64_t *r_out, double *core_out, int64_t cap, int64_t *len_out,
                            ttg_info *info) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_CHECK(n && m && ra && rx && r_out && acore && xcore && core_out && d >= 1, TTG_BAD_ARG, "null pointer argument");
  validate_mv(d, n, m, ra, rx);
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  std::vector<int64_t> rp(d + 1);
  for (int64_t i = 0; i <= d; ++i) rp[i] = ra[i] * rx[i];
  int64_t need = 0;
  ttg_status s0 = round_capacity(d, n, rp.data(), &need);
  if (s0 != TTG_OK) return s0;
  if (cap < need) {
    if (len_out) *len_out = need;
    return TTG_NEED_SPACE;
  }
  int64_t la = 0, lx = 0;
  for (int64_t i = 0; i < d; ++i) {
    la += ra[i] * n[i] * m[i] * ra[i + 1];
    lx += rx[i] * m[i] * rx[i + 1];
  }
  double *da = nullptr, *dx = nullptr, *dy = nullptr;
  TTG_HIP(hipMalloc(&da, std::max<int64_t>(la, 1) * sizeof(double)));
  TTG_HIP(hipMalloc(&dx, std::max<int64_t>(lx, 1) * sizeof(double)));
  TTG_HIP(hipMalloc(&dy, std::max<int64_t>(need, 1) * sizeof(double)));
  ttg_status st = TTG_OK;
  try {
    TTG_HIP(hipMemcpyAsync(da, acore, la * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    TTG_HIP(hipMemcpyAsync(dx, xcore, lx * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int64_t used = 0;
    st = mtimes_round_dev(ctx, d, n, m, ra, da, rx, dx, eps, rmax, prec, r_out, dy, need, &used, info);
    if (st == TTG_OK || st == TTG_NO_CONVERGENCE) {
      TTG_HIP(hipMemcpyAsync(core_out, dy, used * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      TTG_HIP(hipStreamSynchronize(ctx->stream));
      if (len_out) *len_out = used;
    }
  } catch (...) {
    (void)hipFree(da);
    (void)hipFree(dx);
    (void)hipFree(dy);
    throw;
  }
  (void)hipFree(da);
  (void)hipFree(dx);
  (void)hipFree(dy);
  return st;
  TTG_GUARD_END(ctx)
}

ttg_status ttg_round_batched_dev(ttg_ctx *ctx, int64_t count, const int64_t *d,
                                 const int64_t *nofs, const int64_t *n, const int64_t *rofs,
                                 const int64_t *r, const int64_t *cofs, const double *dcore,
                                 double eps, const int64_t *rmax, int prec, int64_t *r_out,
                                 const int64_t *cofs_out, double *dcore_out, ttg_info *info) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_CHECK(d && nofs && n && rofs && r && cofs && dcore && r_out && cofs_out && dcore_out && count >= 0,
            TTG_BAD_ARG, "null pointer argument");
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return round_batched_dev(ctx, count, d, nofs, n, rofs, r, cofs, dcore, eps, rmax, prec, r_out, cofs_out,
                           dcore_out, info);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_adjoint_completion_f64(ttg_ctx *ctx, int64_t d, const int64_t *n,
                                      const int64_t *r, const double *dcore, int64_t S,
                                      const int32_t *didx, const double *dvals, double *loss,
                                      double *dgrad) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return adjoint_completion<double>(ctx, d, n, r, dcore, S, didx, dvals, loss, dgrad);
  TTG_GUARD_END(ctx)
}
ttg_status ttg_adjoint_completion_f32(ttg_ctx *ctx, int64_t d, const int64_t *n,
                                      const int64_t *r, const float *dcore, int64_t S,
                                      const int32_t *didx, const float *dvals, double *loss,
                                      float *dgrad) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return adjoint_completion<float>(ctx, d, n, r, dcore, S, didx, dvals, loss, dgrad);
  TTG_GUARD_END(ctx)
}
ttg_status ttg_adjoint_bilinear_f64(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                                    const int64_t *ry, const double *dycore, const int64_t *ra,
                                    const double *dacore, const int64_t *rx, const double *dxcore,
                                    double *value, double *dgx, double *dgy) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return adjoint_bilinear<double>(ctx, d, n, m, ry, dycore, ra, dacore, rx, dxcore, value, dgx, dgy, nullptr, 0);
  TTG_GUARD_END(ctx)
}
ttg_status ttg_adjoint_bilinear_f64_noalloc(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                                            const int64_t *ry, const double *dycore, const int64_t *ra,
                                            const double *dacore, const int64_t *rx, const double *dxcore,
                                            double *value, double *dgx, double *dgy) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  void *ws = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
  TTG_CHECK(ws != nullptr, TTG_HIP_ERROR, "no-allocation scratch not set up on this device");
  return adjoint_bilinear<double>(ctx, d, n, m, ry, dycore, ra, dacore, rx, dxcore, value, dgx, dgy, ws,
                                  kNoAllocScratchBytes);
  TTG_GUARD_END(ctx)
}
ttg_status ttg_adjoint_bilinear_f32(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *m,
                                    const int64_t *ry, const float *dycore, const int64_t *ra,
                                    const float *dacore, const int64_t *rx, const float *dxcore,
                                    double *value, float *dgx, float *dgy) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return adjoint_bilinear<float>(ctx, d, n, m, ry, dycore, ra, dacore, rx, dxcore, value, dgx, dgy, nullptr, 0);
  TTG_GUARD_END(ctx)
}

}  // extern "C"

// ---- sums, norms and inner products ---------------------------------------
ttg_status ttg_plus_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx, double a,
                        const int64_t *ry, const double *dy, double b, int64_t *rz, double *dz, int64_t cap,
                        int64_t *len_out) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_CHECK(dx && dy && rz && dz, TTG_BAD_ARG, "null pointer argument");
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return plus_dev(ctx, d, n, rx, dx, a, ry, dy, b, rz, dz, cap, len_out);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_plus_round_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx,
                              double a, const int64_t *ry, const double *dy, double b, double eps,
                              const int64_t *rmax, int prec, int64_t *r_out, double *dcore_out, int64_t cap,
                              int64_t *len_out, ttg_info *info) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return plus_round_dev(ctx, d, n, rx, dx, a, ry, dy, b, eps, rmax, prec, r_out, dcore_out, cap, len_out, info);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_norm_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r, const double *dcore,
                        double *nrm) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return norm_dev(ctx, d, n, r, dcore, nrm);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_plus_norm_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx,
                             double a, const int64_t *ry, const double *dy, double b, double *nrm) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return plus_norm_dev(ctx, d, n, rx, dx, a, ry, dy, b, nrm);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_dot_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *rx, const double *dx,
                       const int64_t *ry, const double *dy, double *value) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return dot_dev(ctx, d, n, rx, dx, ry, dy, value);
  TTG_GUARD_END(ctx)
}

// ---- SDV files ---------------------------------------------------------------
ttg_status ttg_sdv_info(const char *path, int64_t *d, int64_t *core_len, int *dtype) {
  ttg_ctx *ctx = nullptr;
  TTG_GUARD_BEGIN
  TTG_CHECK(d, TTG_BAD_ARG, "null pointer argument");
  return sdv_info(path, d, core_len, dtype);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_sdv_read(const char *path, int64_t dcap, int64_t *d, int64_t *n, int64_t *r, double *core,
                        int64_t cap, int64_t *len_out) {
  ttg_ctx *ctx = nullptr;
  TTG_GUARD_BEGIN
  return sdv_read(path, dcap, d, n, r, core, cap, len_out);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_sdv_read_dev(ttg_ctx *ctx, const char *path, int64_t dcap, int64_t *d, int64_t *n, int64_t *r,
                            double *dcore, int64_t cap, int64_t *len_out) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return sdv_read_dev(ctx, path, dcap, d, n, r, dcore, cap, len_out);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_sdv_write(const char *path, int64_t d, const int64_t *n, const int64_t *r, const double *core,
                         int precision) {
  ttg_ctx *ctx = nullptr;
  TTG_GUARD_BEGIN
  return sdv_write(path, d, n, r, core, precision);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_sdv_write_dev(ttg_ctx *ctx, const char *path, int64_t d, const int64_t *n, const int64_t *r,
                             const double *dcore, int precision) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return sdv_write_dev(ctx, path, d, n, r, dcore, precision);
  TTG_GUARD_END(ctx)
}

// ---- TT-SVD --------------------------------------------------------------------
ttg_status ttg_full_to_tt_dev(ttg_ctx *ctx, int64_t d, const int64_t *n, const double *dfull, double eps,
                              const int64_t *rmax, int64_t *r_out, double *dcore_out, int64_t cap,
                              int64_t *len_out, ttg_info *info) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return full_to_tt_dev(ctx, d, n, dfull, eps, rmax, r_out, dcore_out, cap, len_out, info);
  TTG_GUARD_END(ctx)
}

// ---- launch interface of the Gram-chain round ---------------------------------
ttg_status ttg_round_gram_prepare(ttg_ctx *ctx, int64_t d, const int64_t *n, const int64_t *r, const double *dcore,
                                  double eps, int prec, int64_t *r_out, double *dcore_out, int64_t cap,
                                  int64_t *len_out, int *proven) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return round_gram_prepare(ctx, d, n, r, dcore, eps, prec, r_out, dcore_out, cap, len_out, proven);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_round_gram_launch(ttg_ctx *ctx, int *dstatus) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  TTG_HIP(hipSetDevice(ctx->device));
  ctx->last_error.clear();
  return round_gram_launch(ctx, dstatus);
  TTG_GUARD_END(ctx)
}

ttg_status ttg_round_gram_graph(ttg_ctx *ctx, void **graph) {
  if (!ctx) return TTG_BAD_ARG;
  TTG_GUARD_BEGIN
  ctx->last_error.clear();
  return round_gram_graph(ctx, graph);
  TTG_GUARD_END(ctx)
}

