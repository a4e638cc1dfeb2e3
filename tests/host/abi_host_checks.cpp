// Host-side checks of libttgpu's C ABI that need no GPU: argument validation
// on every entry point, the size helpers, error strings, and the SDV reader
// and writer (round trips of the reference's SDV files, truncated, missing and
// undersized cases).  Built with the host code under AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py; any report
// aborts the run.
//   abi_host_checks <A.sdv> <b.sdv> <scratch dir>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ttgpu.h"

static int failures = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                                         \
    }                                                                     \
  } while (0)

static std::vector<char> slurp(const std::string &p) {
  std::vector<char> b;
  FILE *f = std::fopen(p.c_str(), "rb");
  if (!f) return b;
  char buf[4096];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + k);
  std::fclose(f);
  return b;
}

static void sdv_round_trip(const std::string &in, const std::string &out) {
  int64_t d = 0, len = 0;
  int dtype = -1;
  CHECK(ttg_sdv_info(in.c_str(), &d, &len, &dtype) == TTG_OK);
  CHECK(d >= 1 && len >= 1 && dtype == 0);
  std::vector<int64_t> n(d), r(d + 1);
  std::vector<double> core(len);
  int64_t d2 = 0, used = 0;
  // undersized n/r arrays and core capacity
  CHECK(ttg_sdv_read(in.c_str(), d - 1, &d2, n.data(), r.data(), core.data(), len, &used) == TTG_NEED_SPACE);
  CHECK(ttg_sdv_read(in.c_str(), d, &d2, n.data(), r.data(), core.data(), len - 1, &used) == TTG_NEED_SPACE);
  CHECK(used == len);
  CHECK(ttg_sdv_read(in.c_str(), d, &d2, n.data(), r.data(), core.data(), len, &used) == TTG_OK);
  CHECK(d2 == d && used == len);
  CHECK(ttg_sdv_write(out.c_str(), d, n.data(), r.data(), core.data(), 0) == TTG_OK);
  CHECK(slurp(in) == slurp(out));
  // float precision, then read back (widened)
  CHECK(ttg_sdv_write(out.c_str(), d, n.data(), r.data(), core.data(), 1) == TTG_OK);
  std::vector<double> back(len);
  CHECK(ttg_sdv_read(out.c_str(), d, &d2, n.data(), r.data(), back.data(), len, &used) == TTG_OK);
  for (int64_t i = 0; i < len; ++i) CHECK(back[i] == (double)(float)core[i]);
  CHECK(ttg_sdv_write(out.c_str(), d, n.data(), r.data(), core.data(), 2) == TTG_BAD_ARG);
  // truncated copy: the core ends early
  std::vector<char> raw = slurp(in);
  FILE *f = std::fopen(out.c_str(), "wb");
  std::fwrite(raw.data(), 1, raw.size() - 9, f);
  std::fclose(f);
  CHECK(ttg_sdv_read(out.c_str(), d, &d2, n.data(), r.data(), core.data(), len, &used) == TTG_IO_ERROR);
  // header only
  f = std::fopen(out.c_str(), "wb");
  std::fwrite(raw.data(), 1, 100, f);
  std::fclose(f);
  CHECK(ttg_sdv_info(out.c_str(), &d2, &used, &dtype) == TTG_IO_ERROR);
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: abi_host_checks A.sdv b.sdv scratch_dir\n");
    return 2;
  }
  const std::string dir = argv[3];
  CHECK(ttg_version() != nullptr && std::strlen(ttg_version()) > 0);
  CHECK(std::strcmp(ttg_last_error(nullptr), "no context") == 0);
  CHECK(ttg_ctx_stream(nullptr) == nullptr);
  CHECK(ttg_ctx_create(0, nullptr, nullptr) == TTG_BAD_ARG);
  ttg_ctx *ctx = nullptr;
  const ttg_status cs = ttg_ctx_create(0, nullptr, &ctx);
  CHECK(cs == TTG_NO_DEVICE || cs == TTG_OK);
  if (cs == TTG_OK) ttg_ctx_destroy(ctx);
  CHECK(ttg_ctx_create(-1, nullptr, &ctx) == TTG_NO_DEVICE && ctx == nullptr);
  CHECK(ttg_ctx_destroy(nullptr) == TTG_OK);  // like free(NULL)

  // size helper: validation and the QR-capped size
  const int64_t n[3] = {2, 3, 4}, r[4] = {1, 5, 7, 1}, bad_r[4] = {1, 0, 2, 1};
  int64_t cap = -1;
  CHECK(ttg_round_capacity(3, n, r, &cap) == TTG_OK);
  CHECK(cap == 1 * 2 * 2 + 2 * 3 * 6 + 6 * 4 * 1);  // caps r1 = min(2, 5), r2 = min(6, 7)
  CHECK(ttg_round_capacity(0, n, r, &cap) == TTG_BAD_ARG);
  CHECK(ttg_round_capacity(3, nullptr, r, &cap) == TTG_BAD_ARG);
  CHECK(ttg_round_capacity(3, n, bad_r, &cap) == TTG_BAD_ARG);
  CHECK(ttg_round_capacity(3, n, r, nullptr) == TTG_BAD_ARG);

  // every context-taking entry point refuses a null context before touching anything
  double v = 0.0;
  int64_t ro[4], len = 0;
  double buf[64];
  CHECK(ttg_round(nullptr, 3, n, r, buf, 1e-8, nullptr, 0, ro, buf, 64, &len, nullptr) == TTG_BAD_ARG);
  CHECK(ttg_round_dev(nullptr, 3, n, r, buf, 1e-8, nullptr, 0, ro, buf, 64, &len, nullptr) == TTG_BAD_ARG);
  CHECK(ttg_round_matrix(nullptr, 3, n, n, r, buf, 1e-8, nullptr, 0, ro, buf, 64, &len, nullptr) == TTG_BAD_ARG);
  CHECK(ttg_mtimes(nullptr, 3, n, n, r, buf, r, buf, ro, buf, 64, &len) == TTG_BAD_ARG);
  CHECK(ttg_mtimes_dev(nullptr, 3, n, n, r, buf, r, buf, ro, buf, 64, &len) == TTG_BAD_ARG);
  CHECK(ttg_mtimes_round(nullptr, 3, n, n, r, buf, r, buf, 1e-8, nullptr, 0, ro, buf, 64, &len, nullptr) ==
        TTG_BAD_ARG);
  CHECK(ttg_mtimes_round_dev(nullptr, 3, n, n, r, buf, r, buf, 1e-8, nullptr, 0, ro, buf, 64, &len, nullptr) ==
        TTG_BAD_ARG);
  const int64_t dd[1] = {3}, zero[1] = {0};
  CHECK(ttg_round_batched_dev(nullptr, 1, dd, zero, n, zero, r, zero, buf, 1e-8, nullptr, 0, ro, zero, buf,
                              nullptr) == TTG_BAD_ARG);
  CHECK(ttg_plus_dev(nullptr, 3, n, r, buf, 1.0, r, buf, 1.0, ro, buf, 64, &len) == TTG_BAD_ARG);
  CHECK(ttg_plus_round_dev(nullptr, 3, n, r, buf, 1.0, r, buf, 1.0, 1e-8, nullptr, 0, ro, buf, 64, &len,
                           nullptr) == TTG_BAD_ARG);
  CHECK(ttg_norm_dev(nullptr, 3, n, r, buf, &v) == TTG_BAD_ARG);
  CHECK(ttg_plus_norm_dev(nullptr, 3, n, r, buf, 1.0, r, buf, -1.0, &v) == TTG_BAD_ARG);
  CHECK(ttg_dot_dev(nullptr, 3, n, r, buf, r, buf, &v) == TTG_BAD_ARG);
  CHECK(ttg_full_to_tt_dev(nullptr, 3, n, buf, 1e-8, nullptr, ro, buf, 64, &len, nullptr) == TTG_BAD_ARG);
  CHECK(ttg_sdv_read_dev(nullptr, argv[1], 3, &len, ro, ro, buf, 64, &len) == TTG_BAD_ARG);
  CHECK(ttg_sdv_write_dev(nullptr, "x", 3, n, r, buf, 0) == TTG_BAD_ARG);
  int32_t idx[3] = {0, 0, 0};
  float fb[8];
  CHECK(ttg_adjoint_completion_f64(nullptr, 3, n, r, buf, 1, idx, buf, &v, buf) == TTG_BAD_ARG);
  CHECK(ttg_adjoint_completion_f32(nullptr, 3, n, r, fb, 1, idx, fb, &v, fb) == TTG_BAD_ARG);
  CHECK(ttg_adjoint_bilinear_f64(nullptr, 3, n, n, r, buf, r, buf, r, buf, &v, buf, buf) == TTG_BAD_ARG);
  CHECK(ttg_adjoint_bilinear_f32(nullptr, 3, n, n, r, fb, r, fb, r, fb, &v, fb, fb) == TTG_BAD_ARG);
  CHECK(ttg_adjoint_bilinear_f64_noalloc(nullptr, 3, n, n, r, buf, r, buf, r, buf, &v, buf, buf) == TTG_BAD_ARG);
  CHECK(ttg_profile_begin(nullptr, 1) == TTG_BAD_ARG);
  int flag = 0;
  void *graph = nullptr;
  CHECK(ttg_round_gram_prepare(nullptr, 3, n, r, buf, 1e-8, 0, ro, buf, 64, &len, &flag) == TTG_BAD_ARG);
  CHECK(ttg_round_gram_launch(nullptr, &flag) == TTG_BAD_ARG);
  CHECK(ttg_round_gram_graph(nullptr, &graph) == TTG_BAD_ARG && graph == nullptr);

  // SDV files
  sdv_round_trip(argv[1], dir + "/a_copy.sdv");
  sdv_round_trip(argv[2], dir + "/b_copy.sdv");
  int64_t d = 0;
  CHECK(ttg_sdv_info((dir + "/missing.sdv").c_str(), &d, &len, nullptr) == TTG_IO_ERROR);
  CHECK(ttg_sdv_info(nullptr, &d, &len, nullptr) == TTG_BAD_ARG);
  CHECK(ttg_sdv_write((dir + "/no/such/dir.sdv").c_str(), 3, n, r, buf, 0) == TTG_IO_ERROR);
  CHECK(ttg_sdv_write((dir + "/bad.sdv").c_str(), 3, n, bad_r, buf, 0) == TTG_BAD_ARG);

  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("host ABI checks passed\n");
  return 0;
}
