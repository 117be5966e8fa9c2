// Native benchmark / check harness for the C ABI, with the command line of
// the reference's harness (radik/test/test_radik.cu:225-300):
//
//   rk_harness BATCH LOG2N K [RANDOM_TASK SCALING DISTRIBUTION BATCHED PADDING SORTING]
//              [--largest] [--descend] [--dtype f32|f16|bf16] [--reps R] [--check]
//
//   RANDOM_TASK   0: every row has N elements; 1: lengths drawn from
//                 [0.9 N, 1.1 N] (seeded)            (test_radik.cu:31-39)
//   SCALING       accepted and ignored: the key-space scaling is exact and on
//   DISTRIBUTION  0 U[0,1)  1 U[0.6,0.7)  2 U[128.6,128.7)  3 Zipf(s=1.1)
//                 scaled by 1/len  4 all zero          (test_radik.cu:88-127)
//   BATCHED       1: one call for the batch; 0: one call per row
//   PADDING       accepted and ignored (no padded layout is needed)
//   SORTING       needSorting
//
// Defaults follow the reference harness: smallest-k, ascending
// (test_radik.cu:18-20).  Values are generated on the device by the
// library's seeded generators (rk_gen_fill).  Calls go through the
// reference-surface entry (rk_topk_compat: host taskLen[], one K,
// [BATCH][K] output).  Prints "elapsed: X ms" for the timed call(s) (the
// mean of --reps repetitions after one warm-up), as the reference does
// (test_radik.cu:184); --check compares values and indices bit for bit
// with the CPU oracle and exits non-zero on a mismatch.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "radik_amd.h"
#include "radik_amd_harness.h"

extern "C" int orc_topk(int dt, const void* vals, const int64_t* row_off, const int32_t* row_k,
                        int32_t K, int32_t num_rows, uint32_t flags, const int32_t* idx_in,
                        void* out_vals, int32_t* out_idx, const int64_t* out_off,
                        int32_t nthreads);
extern "C" int orc_hardware_threads(void);

#define HIP_OK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                   __LINE__);                                                         \
      std::exit(2);                                                                   \
    }                                                                                 \
  } while (0)
#define RK_OK_(x)                                                                   \
  do {                                                                              \
    int rc_ = (x);                                                                  \
    if (rc_ != RK_OK) {                                                             \
      std::fprintf(stderr, "radik_amd error %s at %s:%d\n", rk_strerror(rc_), __FILE__, \
                   __LINE__);                                                       \
      std::exit(2);                                                                 \
    }                                                                               \
  } while (0)

namespace {

// SplitMix64: the row lengths' seeded stream
uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int usage() {
  std::fprintf(stderr,
               "usage: rk_harness BATCH LOG2N K [RANDOM_TASK SCALING DISTRIBUTION BATCHED "
               "PADDING SORTING] [--largest] [--descend] [--dtype f32|f16|bf16] [--reps R] "
               "[--check]\n");
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  bool largest = false, ascend = true, check = false;
  int dtype = RK_F32, reps = 1;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--largest") largest = true;
    else if (a == "--descend") ascend = false;
    else if (a == "--check") check = true;
    else if (a == "--reps" && i + 1 < argc) reps = std::atoi(argv[++i]);
    else if (a == "--dtype" && i + 1 < argc) {
      const std::string d = argv[++i];
      dtype = d == "f32" ? RK_F32 : d == "f16" ? RK_F16 : d == "bf16" ? RK_BF16 : -1;
      if (dtype < 0) return usage();
    } else if (!a.empty() && a[0] == '-') return usage();
    else pos.push_back(a);
  }
  if (pos.size() < 3) return usage();
  const int B = std::atoi(pos[0].c_str());
  const int log2n = std::atoi(pos[1].c_str());
  const int K = std::atoi(pos[2].c_str());
  auto arg = [&](size_t i, int dflt) { return pos.size() > i ? std::atoi(pos[i].c_str()) : dflt; };
  const int random_task = arg(3, 0);
  const int distribution = arg(5, 0);
  const int batched = arg(6, 1);
  const int sorting = arg(8, 1);
  if (B <= 0 || log2n < 0 || log2n > 29 || K < 0 || distribution < 0 || distribution > 4 ||
      reps < 1)
    return usage();
  const int N = 1 << log2n;

  // rows
  std::vector<int32_t> len(B, N);
  if (random_task) {
    uint64_t s = 1;
    const int lo = N * 9 / 10, hi = N * 11 / 10;
    for (int i = 0; i < B; ++i)
      len[i] = lo + (int)((double)(splitmix(s) >> 11) * 0x1.0p-53 * (hi - lo));
  }
  std::vector<int64_t> off(B + 1, 0);
  int32_t longest = 0;
  for (int i = 0; i < B; ++i) {
    off[i + 1] = off[i] + len[i];
    longest = len[i] > longest ? len[i] : longest;
  }
  const int64_t total = off[B];
  std::printf("batchSize: %d, K: %d, total: %lld, dtype: %s, %s, %s\n", B, K, (long long)total,
              dtype == RK_F32 ? "f32" : dtype == RK_F16 ? "f16" : "bf16",
              largest ? "largest" : "smallest", batched ? "batched" : "per-row loop");

  const size_t es = dtype == RK_F32 ? 4 : 2;
  void* vals = nullptr;
  HIP_OK(hipMalloc(&vals, (size_t)total * es + 16));
  for (int i = 0; i < B; ++i) {
    void* row = static_cast<char*>(vals) + off[i] * es;
    switch (distribution) {
      case 0: RK_OK_(rk_gen_fill(dtype, RK_GEN_UNIFORM, 1234, i, len[i], len[i], 0, 1, 0, row, nullptr)); break;
      case 1: RK_OK_(rk_gen_fill(dtype, RK_GEN_RANGE, 1, i, len[i], len[i], 0.6, 0.7, 0, row, nullptr)); break;
      case 2: RK_OK_(rk_gen_fill(dtype, RK_GEN_RANGE, 1, i, len[i], len[i], 128.6, 128.7, 0, row, nullptr)); break;
      case 3: RK_OK_(rk_gen_fill(dtype, RK_GEN_ZIPF, 1, i, len[i], len[i], 1.1, (double)len[i], 0, row, nullptr)); break;
      default: RK_OK_(rk_gen_fill(dtype, RK_GEN_CONST, 1, i, len[i], len[i], 0, 0, 0, row, nullptr)); break;
    }
  }
  void* out_v = nullptr;
  int32_t* out_i = nullptr;
  const size_t n_out = (size_t)B * (size_t)K;
  HIP_OK(hipMalloc(&out_v, n_out * es + 16));
  HIP_OK(hipMalloc(reinterpret_cast<void**>(&out_i), n_out * 4 + 16));
  const uint32_t flags = (largest ? RK_LARGEST : 0u) | (ascend ? RK_ASCEND : 0u) |
                         (sorting ? RK_SORTED : 0u);
  size_t ws_bytes = 0;
  RK_OK_(rk_compat_workspace_size(dtype, K, longest, batched ? B : 1, flags,
                                  &ws_bytes));
  void* ws = nullptr;
  HIP_OK(hipMalloc(&ws, ws_bytes));
  hipStream_t stream;
  HIP_OK(hipStreamCreate(&stream));

  auto run = [&]() {
    if (batched) {
      RK_OK_(rk_topk_compat(dtype, vals, nullptr, out_v, out_i, ws, ws_bytes, len.data(), B, K,
                            flags, stream));
    } else {
      for (int i = 0; i < B; ++i)
        RK_OK_(rk_topk_compat(dtype, static_cast<char*>(vals) + off[i] * es, nullptr,
                              static_cast<char*>(out_v) + (size_t)i * K * es, out_i + (size_t)i * K,
                              ws, ws_bytes, len.data() + i, 1, K, flags, stream));
    }
  };
  run();  // warm-up
  HIP_OK(hipStreamSynchronize(stream));
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, stream));
  for (int r = 0; r < reps; ++r) run();
  HIP_OK(hipEventRecord(e1, stream));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  std::printf("elapsed: %f ms\n", ms / reps);
  uint32_t status = 0;
  RK_OK_(rk_topk_status(ws, stream, &status, nullptr, nullptr));

  int rc = 0;
  if (check) {
    std::vector<uint8_t> hv((size_t)total * es + 16), gv(n_out * es + 16), ov(n_out * es + 16);
    std::vector<int32_t> gi(n_out + 1), oi(n_out + 1);
    HIP_OK(hipMemcpy(hv.data(), vals, (size_t)total * es, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(gv.data(), out_v, n_out * es, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(gi.data(), out_i, n_out * 4, hipMemcpyDeviceToHost));
    if (orc_topk(dtype, hv.data(), off.data(), nullptr, K, B, flags, nullptr, ov.data(), oi.data(),
                 nullptr, orc_hardware_threads()) != 0) {
      std::fprintf(stderr, "oracle: bad argument\n");
      return 2;
    }
    size_t bad = 0;
    for (size_t j = 0; j < n_out; ++j) {
      if (gi[j] != oi[j] || std::memcmp(&gv[j * es], &ov[j * es], es) != 0) {
        if (bad < 5)
          std::printf("mismatch at row %zu slot %zu: gpu index %d, oracle index %d\n",
                      j / (size_t)K, j % (size_t)K, gi[j], oi[j]);
        ++bad;
      }
    }
    std::printf("check: %s (%zu of %zu outputs differ)\n", bad ? "FAILED" : "ok", bad, n_out);
    rc = bad ? 1 : 0;
  }
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
  HIP_OK(hipStreamDestroy(stream));
  HIP_OK(hipFree(ws));
  HIP_OK(hipFree(out_i));
  HIP_OK(hipFree(out_v));
  HIP_OK(hipFree(vals));
  return rc;
}
