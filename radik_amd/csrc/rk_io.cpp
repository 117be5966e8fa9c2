// CSR batch and result files: the batched entry's layout (row offsets, per-row
// k, values; output offsets, values, indices) written to and read from disk
// with a 64-byte little-endian header.  Host memory only, no device work.
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/radik_amd.h"

namespace {

constexpr char BATCH_MAGIC[8] = {'R', 'K', 'B', 'A', 'T', 'C', 'H', '1'};
constexpr char RESULT_MAGIC[8] = {'R', 'K', 'R', 'E', 'S', 'L', 'T', '1'};

struct FileHeader {  // 64 bytes on disk, little-endian (the host order here)
  char magic[8];
  int32_t dtype;
  uint32_t flags;
  int64_t rows;
  int64_t total;
  uint8_t reserved[32];
};
static_assert(sizeof(FileHeader) == 64, "file header is 64 bytes");

int elem_bytes(int dtype) {
  if (dtype == RK_F32) return 4;
  if (dtype == RK_F16 || dtype == RK_BF16) return 2;
  return 0;
}

// Offsets of a CSR layout: start at 0, nondecreasing, end at total.
bool offsets_ok(const int64_t* off, int64_t rows, int64_t total) {
  if (off[0] != 0 || off[rows] != total) return false;
  for (int64_t r = 0; r < rows; ++r)
    if (off[r + 1] < off[r]) return false;
  return true;
}

class File {
 public:
  File(const char* path, const char* mode) : f_(path ? std::fopen(path, mode) : nullptr) {}
  ~File() {
    if (f_) std::fclose(f_);
  }
  bool ok() const { return f_ != nullptr; }
  bool write(const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f_) == n; }
  bool read(void* p, size_t n) { return n == 0 || std::fread(p, 1, n, f_) == n; }
  bool at_end() { return std::fgetc(f_) == EOF; }
  bool close() {
    const bool good = std::fflush(f_) == 0 && std::ferror(f_) == 0;
    const bool closed = std::fclose(f_) == 0;
    f_ = nullptr;
    return good && closed;
  }

 private:
  FILE* f_;
};

const uint8_t ZERO_PAD[8] = {0};

int read_header(File& f, const char (&magic)[8], FileHeader* h) {
  if (!f.ok()) return RK_E_IO;
  if (!f.read(h, sizeof(*h))) return RK_E_IO;
  if (std::memcmp(h->magic, magic, 8) != 0) return RK_E_IO;
  if (elem_bytes(h->dtype) == 0 || h->rows < 0 || h->rows > INT32_MAX || h->total < 0)
    return RK_E_IO;
  return RK_OK;
}

int info_of(const char* path, const char (&magic)[8], rk_file_info* info) {
  if (!path || !info) return RK_E_ARG;
  File f(path, "rb");
  FileHeader h;
  const int rc = read_header(f, magic, &h);
  if (rc != RK_OK) return rc;
  info->dtype = h.dtype;
  info->flags = h.flags;
  info->rows = h.rows;
  info->total = h.total;
  return RK_OK;
}

}  // namespace

extern "C" int rk_batch_write(const char* path, int dtype, uint32_t flags, int32_t rows,
                              const int64_t* row_off, const int32_t* row_k, const void* values) {
  if (!path || rows < 0 || !row_off || (rows && !row_k)) return RK_E_ARG;
  const int es = elem_bytes(dtype);
  if (es == 0) return RK_E_UNSUPPORTED;
  const int64_t total = row_off[rows];
  if (!offsets_ok(row_off, rows, total) || (total && !values)) return RK_E_ARG;
  for (int32_t r = 0; r < rows; ++r)
    if (row_k[r] < 0) return RK_E_ARG;
  FileHeader h{};
  std::memcpy(h.magic, BATCH_MAGIC, 8);
  h.dtype = dtype;
  h.flags = flags;
  h.rows = rows;
  h.total = total;
  File f(path, "wb");
  if (!f.ok()) return RK_E_IO;
  const size_t kbytes = (size_t)rows * 4;
  const bool good = f.write(&h, sizeof(h)) && f.write(row_off, ((size_t)rows + 1) * 8) &&
                    f.write(row_k, kbytes) && f.write(ZERO_PAD, (8 - kbytes % 8) % 8) &&
                    f.write(values, (size_t)total * es);
  return good && f.close() ? RK_OK : RK_E_IO;
}

extern "C" int rk_batch_info(const char* path, rk_file_info* info) {
  return info_of(path, BATCH_MAGIC, info);
}

extern "C" int rk_batch_read(const char* path, int dtype, int64_t max_rows, int64_t max_total,
                             int64_t* row_off, int32_t* row_k, void* values) {
  if (!path || !row_off || max_rows < 0 || max_total < 0) return RK_E_ARG;
  if (elem_bytes(dtype) == 0) return RK_E_UNSUPPORTED;
  File f(path, "rb");
  FileHeader h;
  int rc = read_header(f, BATCH_MAGIC, &h);
  if (rc != RK_OK) return rc;
  // The buffers were sized for `dtype`, max_rows and max_total (from
  // rk_batch_info): a file rewritten since with another element type or more
  // data is refused before anything is written.
  if (h.dtype != dtype) return RK_E_IO;
  if (h.rows > max_rows || h.total > max_total) return RK_E_IO;
  if ((h.rows && !row_k) || (h.total && !values)) return RK_E_ARG;
  const size_t kbytes = (size_t)h.rows * 4;
  uint8_t pad[8];
  if (!f.read(row_off, ((size_t)h.rows + 1) * 8) || !f.read(row_k, kbytes) ||
      !f.read(pad, (8 - kbytes % 8) % 8))
    return RK_E_IO;
  if (!offsets_ok(row_off, h.rows, h.total)) return RK_E_IO;
  for (int64_t r = 0; r < h.rows; ++r)
    if (row_k[r] < 0) return RK_E_IO;
  if (!f.read(values, (size_t)h.total * elem_bytes(h.dtype)) || !f.at_end()) return RK_E_IO;
  return RK_OK;
}

extern "C" int rk_result_write(const char* path, int dtype, int32_t rows, const int64_t* out_off,
                               const void* out_vals, const int32_t* out_idx) {
  if (!path || rows < 0 || !out_off) return RK_E_ARG;
  const int es = elem_bytes(dtype);
  if (es == 0) return RK_E_UNSUPPORTED;
  const int64_t n = out_off[rows];
  if (!offsets_ok(out_off, rows, n) || (n && (!out_vals || !out_idx))) return RK_E_ARG;
  FileHeader h{};
  std::memcpy(h.magic, RESULT_MAGIC, 8);
  h.dtype = dtype;
  h.rows = rows;
  h.total = n;
  File f(path, "wb");
  if (!f.ok()) return RK_E_IO;
  const bool good = f.write(&h, sizeof(h)) && f.write(out_off, ((size_t)rows + 1) * 8) &&
                    f.write(out_vals, (size_t)n * es) && f.write(out_idx, (size_t)n * 4);
  return good && f.close() ? RK_OK : RK_E_IO;
}

extern "C" int rk_result_info(const char* path, rk_file_info* info) {
  return info_of(path, RESULT_MAGIC, info);
}

extern "C" int rk_result_read(const char* path, int dtype, int64_t max_rows, int64_t max_total,
                              int64_t* out_off, void* out_vals, int32_t* out_idx) {
  if (!path || !out_off || max_rows < 0 || max_total < 0) return RK_E_ARG;
  if (elem_bytes(dtype) == 0) return RK_E_UNSUPPORTED;
  File f(path, "rb");
  FileHeader h;
  int rc = read_header(f, RESULT_MAGIC, &h);
  if (rc != RK_OK) return rc;
  if (h.dtype != dtype) return RK_E_IO;  // buffers sized for another element type
  if (h.rows > max_rows || h.total > max_total) return RK_E_IO;  // larger than the buffers
  if (h.total && (!out_vals || !out_idx)) return RK_E_ARG;
  if (!f.read(out_off, ((size_t)h.rows + 1) * 8)) return RK_E_IO;
  if (!offsets_ok(out_off, h.rows, h.total)) return RK_E_IO;
  if (!f.read(out_vals, (size_t)h.total * elem_bytes(h.dtype)) ||
      !f.read(out_idx, (size_t)h.total * 4) || !f.at_end())
    return RK_E_IO;
  return RK_OK;
}
