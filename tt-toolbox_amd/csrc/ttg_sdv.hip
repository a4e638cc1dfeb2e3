// SDV TT files: the on-disk format of fmex/tt_read.c:13-103 and
// fmex/tt_write.c:21-79 (@tt_tensor/tt_tensor.m:149-165 reads one into a
// tt_tensor, @tt_tensor/tt_save.m:22 writes one).
//
// Layout (little-endian, int32 fields):
//   128-byte header  char txt[8] ("TT" + 6 blanks), int32 ver[2] ({1, 0}),
//                    int32 inf[4] ({128, dtype, 0, 0}), char comment[64],
//                    int32 i[8] ({1, 0, d, 0, ...})
//   int32 l, m       first and last mode index, d = m - l + 1
//   int32 n[d], int32 r[d+1]
//   core             sum_k r[k] n[k] r[k+1] values; dtype 0 double, 2 float
//                    (1 and 3, the complex forms, are refused here: every
//                    kernel of this library is real)
//
// The _dev forms move the core between the file and device memory in 16 MiB
// chunks through two pinned buffers, so the disk read (write) of one chunk
// overlaps the host-device copy of the previous one; a float file is
// widened (narrowed) on the device, so only its 4-byte values cross PCIe.
#include "ttg_internal.h"

#include <cstdio>
#include <cstring>

namespace ttg {
namespace {

struct SdvHead {
  char txt[8];
  int32_t ver[2];
  int32_t inf[4];
  char comment[64];
  int32_t i[8];
};
static_assert(sizeof(SdvHead) == 128, "SDV header is 128 bytes");

constexpr size_t kChunk = size_t(16) << 20;

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

struct Pinned {
  void *p[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  Pinned() {
    for (int b = 0; b < 2; ++b) {
      TTG_HIP(hipHostMalloc(&p[b], kChunk, hipHostMallocDefault));
      TTG_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
    }
  }
  ~Pinned() {
    for (int b = 0; b < 2; ++b) {
      if (ev[b]) (void)hipEventSynchronize(ev[b]);
      if (ev[b]) (void)hipEventDestroy(ev[b]);
      if (p[b]) (void)hipHostFree(p[b]);
    }
  }
};

void read_exact(FILE *f, void *dst, size_t bytes, const char *what) {
  if (bytes && std::fread(dst, 1, bytes, f) != bytes)
    throw Error(TTG_IO_ERROR, std::string("SDV file ends inside the ") + what);
}

void write_exact(FILE *f, const void *src, size_t bytes) {
  if (bytes && std::fwrite(src, 1, bytes, f) != bytes) throw Error(TTG_IO_ERROR, "SDV file write failed");
}

// header, l/m, n and r (tt_read.c:29-68); returns the core length
int64_t read_prologue(FILE *f, int64_t dcap, int64_t *d, int64_t *n, int64_t *r, int *dtype) {
  SdvHead h;
  read_exact(f, &h, sizeof(h), "header");
  TTG_CHECK(std::strncmp(h.txt, "TT", 2) == 0, TTG_BAD_ARG, "not an SDV TT file");
  TTG_CHECK(h.ver[0] == 1, TTG_BAD_ARG, "wrong version of SDV TT file (current is 1)");
  *dtype = h.inf[1];
  TTG_CHECK(h.inf[1] >= 0 && h.inf[1] <= 3, TTG_BAD_ARG, "wrong SDV data type");
  int32_t lm[2];
  read_exact(f, lm, sizeof(lm), "mode range");
  const int64_t dd = (int64_t)lm[1] - lm[0] + 1;
  TTG_CHECK(dd >= 1 && dd <= (1 << 20), TTG_BAD_ARG, "SDV file has an invalid mode range");
  *d = dd;
  if (!n) return 0;  // size query
  TTG_CHECK(dcap >= dd, TTG_NEED_SPACE, "n/r arrays too short for the file's d");
  std::vector<int32_t> nn(dd), rr(dd + 1);
  read_exact(f, nn.data(), dd * sizeof(int32_t), "mode sizes");
  read_exact(f, rr.data(), (dd + 1) * sizeof(int32_t), "ranks");
  int64_t len = 0;
  for (int64_t k = 0; k < dd; ++k) {
    TTG_CHECK(nn[k] >= 1 && rr[k] >= 1 && rr[k + 1] >= 1, TTG_BAD_ARG, "SDV file has a size or rank below 1");
    n[k] = nn[k];
    r[k] = rr[k];
    len += (int64_t)rr[k] * nn[k] * rr[k + 1];
  }
  r[dd] = rr[dd];
  return len;
}

int64_t file_len_query(FILE *f, int64_t d) {
  std::vector<int32_t> nn(d), rr(d + 1);
  read_exact(f, nn.data(), d * sizeof(int32_t), "mode sizes");
  read_exact(f, rr.data(), (d + 1) * sizeof(int32_t), "ranks");
  int64_t len = 0;
  for (int64_t k = 0; k < d; ++k) len += (int64_t)rr[k] * nn[k] * rr[k + 1];
  return len;
}

FILE *open_or_throw(const char *path, const char *mode) {
  TTG_CHECK(path != nullptr, TTG_BAD_ARG, "null path");
  FILE *f = std::fopen(path, mode);
  if (!f) throw Error(TTG_IO_ERROR, std::string("the file ") + path + " can not be opened");
  return f;
}

// header + l/m + n + r as tt_write.c:40-54 writes them (mexFunction :97-136)
void write_prologue(FILE *f, int64_t d, const int64_t *n, const int64_t *r, int dtype) {
  TTG_CHECK(d >= 1 && d < (int64_t(1) << 31), TTG_BAD_ARG, "d out of range");
  for (int64_t k = 0; k < d; ++k)
    TTG_CHECK(n[k] >= 1 && n[k] < (int64_t(1) << 31), TTG_BAD_ARG, "mode size out of the file's int32 range");
  for (int64_t k = 0; k <= d; ++k)
    TTG_CHECK(r[k] >= 1 && r[k] < (int64_t(1) << 31), TTG_BAD_ARG, "rank out of the file's int32 range");
  SdvHead h;
  std::memset(&h, 0, sizeof(h));
  std::memcpy(h.txt, "TT      ", 8);
  h.ver[0] = 1;
  h.inf[0] = 128;
  h.inf[1] = dtype;
  h.i[0] = 1;
  h.i[2] = (int32_t)d;
  write_exact(f, &h, sizeof(h));
  const int32_t lm[2] = {1, (int32_t)d};
  write_exact(f, lm, sizeof(lm));
  std::vector<int32_t> nn(n, n + d), rr(r, r + d + 1);
  write_exact(f, nn.data(), d * sizeof(int32_t));
  write_exact(f, rr.data(), (d + 1) * sizeof(int32_t));
}

}  // namespace

ttg_status sdv_info(const char *path, int64_t *d, int64_t *core_len, int *dtype) {
  File fh;
  fh.f = open_or_throw(path, "rb");
  int dt = 0;
  read_prologue(fh.f, 0, d, nullptr, nullptr, &dt);
  if (core_len) *core_len = file_len_query(fh.f, *d);
  if (dtype) *dtype = dt;
  return TTG_OK;
}

ttg_status sdv_read(const char *path, int64_t dcap, int64_t *d, int64_t *n, int64_t *r, double *core, int64_t cap,
                    int64_t *len_out) {
  TTG_CHECK(d && n && r && core, TTG_BAD_ARG, "null pointer argument");
  File fh;
  fh.f = open_or_throw(path, "rb");
  int dtype = 0;
  const int64_t len = read_prologue(fh.f, dcap, d, n, r, &dtype);
  if (len_out) *len_out = len;
  TTG_CHECK(dtype == 0 || dtype == 2, TTG_BAD_ARG, "complex SDV cores are not supported");
  if (cap < len) return TTG_NEED_SPACE;
  if (dtype == 0) {
    read_exact(fh.f, core, len * sizeof(double), "core");
  } else {
    std::vector<float> tmp(std::min<int64_t>(len, kChunk / sizeof(float)));
    for (int64_t o = 0; o < len; o += (int64_t)tmp.size()) {
      const int64_t c = std::min<int64_t>((int64_t)tmp.size(), len - o);
      read_exact(fh.f, tmp.data(), c * sizeof(float), "core");
      for (int64_t i = 0; i < c; ++i) core[o + i] = tmp[i];
    }
  }
  return TTG_OK;
}

ttg_status sdv_read_dev(ttg_ctx *ctx, const char *path, int64_t dcap, int64_t *d, int64_t *n, int64_t *r,
                        double *dcore, int64_t cap, int64_t *len_out) {
  TTG_CHECK(d && n && r && dcore, TTG_BAD_ARG, "null pointer argument");
  File fh;
  fh.f = open_or_throw(path, "rb");
  int dtype = 0;
  const int64_t len = read_prologue(fh.f, dcap, d, n, r, &dtype);
  if (len_out) *len_out = len;
  TTG_CHECK(dtype == 0 || dtype == 2, TTG_BAD_ARG, "complex SDV cores are not supported");
  if (cap < len) return TTG_NEED_SPACE;
  hipStream_t st = ctx->stream;
  const size_t es = dtype == 0 ? sizeof(double) : sizeof(float);
  const int64_t per = (int64_t)(kChunk / es);
  // a float file's values land in a device staging buffer and are widened there
  float *stage = nullptr;
  if (dtype == 2) {
    ctx->arena.reserve(std::max<size_t>(len * sizeof(float), 256) + 4096);
    ctx->arena.reset();
    stage = ctx->arena.take<float>(std::max<int64_t>(len, 1));
  }
  Pinned pb;
  int b = 0;
  for (int64_t o = 0; o < len; o += per, b ^= 1) {
    const int64_t c = std::min<int64_t>(per, len - o);
    TTG_HIP(hipEventSynchronize(pb.ev[b]));  // the copy out of this buffer has finished
    read_exact(fh.f, pb.p[b], c * es, "core");
    void *dst = dtype == 0 ? static_cast<void *>(dcore + o) : static_cast<void *>(stage + o);
    TTG_HIP(hipMemcpyAsync(dst, pb.p[b], c * es, hipMemcpyHostToDevice, st));
    TTG_HIP(hipEventRecord(pb.ev[b], st));
  }
  if (dtype == 2) convert<double, float>(st, dcore, stage, len);
  TTG_HIP(hipStreamSynchronize(st));
  return TTG_OK;
}

ttg_status sdv_write(const char *path, int64_t d, const int64_t *n, const int64_t *r, const double *core,
                     int precision) {
  TTG_CHECK(n && r && core, TTG_BAD_ARG, "null pointer argument");
  TTG_CHECK(precision == 0 || precision == 1, TTG_BAD_ARG, "precision must be 0 (double) or 1 (float)");
  File fh;
  fh.f = open_or_throw(path, "wb");
  write_prologue(fh.f, d, n, r, 2 * precision);
  int64_t len = 0;
  for (int64_t k = 0; k < d; ++k) len += r[k] * n[k] * r[k + 1];
  if (precision == 0) {
    write_exact(fh.f, core, len * sizeof(double));
  } else {
    std::vector<float> tmp(std::min<int64_t>(len, kChunk / sizeof(float)));
    for (int64_t o = 0; o < len; o += (int64_t)tmp.size()) {
      const int64_t c = std::min<int64_t>((int64_t)tmp.size(), len - o);
      for (int64_t i = 0; i < c; ++i) tmp[i] = (float)core[o + i];
      write_exact(fh.f, tmp.data(), c * sizeof(float));
    }
  }
  TTG_CHECK(std::fflush(fh.f) == 0, TTG_IO_ERROR, "SDV file write failed");
  return TTG_OK;
}

ttg_status sdv_write_dev(ttg_ctx *ctx, const char *path, int64_t d, const int64_t *n, const int64_t *r,
                         const double *dcore, int precision) {
  TTG_CHECK(n && r && dcore, TTG_BAD_ARG, "null pointer argument");
  TTG_CHECK(precision == 0 || precision == 1, TTG_BAD_ARG, "precision must be 0 (double) or 1 (float)");
  File fh;
  fh.f = open_or_throw(path, "wb");
  write_prologue(fh.f, d, n, r, 2 * precision);
  int64_t len = 0;
  for (int64_t k = 0; k < d; ++k) len += r[k] * n[k] * r[k + 1];
  hipStream_t st = ctx->stream;
  const void *src = dcore;
  size_t es = sizeof(double);
  if (precision == 1) {
    ctx->arena.reserve(std::max<size_t>(len * sizeof(float), 256) + 4096);
    ctx->arena.reset();
    float *stage = ctx->arena.take<float>(std::max<int64_t>(len, 1));
    convert<float, double>(st, stage, dcore, len);
    src = stage;
    es = sizeof(float);
  }
  const int64_t per = (int64_t)(kChunk / es);
  Pinned pb;
  // chunk k is copied into buffer k % 2 while chunk k - 1 is written out
  const int64_t chunks = (len + per - 1) / per;
  auto issue = [&](int64_t k) {
    const int64_t o = k * per, c = std::min<int64_t>(per, len - o);
    TTG_HIP(hipMemcpyAsync(pb.p[k & 1], static_cast<const char *>(src) + o * es, c * es, hipMemcpyDeviceToHost,
                           st));
    TTG_HIP(hipEventRecord(pb.ev[k & 1], st));
  };
  if (chunks > 0) issue(0);
  for (int64_t k = 0; k < chunks; ++k) {
    if (k + 1 < chunks) issue(k + 1);
    TTG_HIP(hipEventSynchronize(pb.ev[k & 1]));
    const int64_t o = k * per, c = std::min<int64_t>(per, len - o);
    write_exact(fh.f, pb.p[k & 1], c * es);
    // buffer k & 1 is reused by chunk k + 2, issued after this write
  }
  TTG_CHECK(std::fflush(fh.f) == 0, TTG_IO_ERROR, "SDV file write failed");
  return TTG_OK;
}

}  // namespace ttg
