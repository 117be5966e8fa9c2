// Autotuner and hipGraph helpers.
//
// The autotuner times the batched pipeline over a grid of tunables (tile
// size, persistent grid, loads in flight, sample shape) on the caller's own
// buffers and stores the fastest per shape bucket
//   (dtype, exact?, ceil log2 total_len, ceil log2 rows, ceil log2 max_k)
// in a process-wide cache that rk_topk_batched / rk_topk_rows consult.  The
// cache can be saved to and loaded from a text file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/radik_amd.h"
#include "rk_pipeline.hpp"

namespace rk {
int timing_enable(int on);
}

namespace {

using Key = std::tuple<int, int, int, int, int>;

std::mutex g_cache_mu;
std::map<Key, rk::Tune>& cache() {
  static std::map<Key, rk::Tune> c;
  return c;
}

int lg2ceil(int64_t x) {
  int l = 0;
  while (l < 62 && (int64_t(1) << l) < x) ++l;
  return l;
}

Key bucket(int dtype, uint32_t flags, int32_t rows, int64_t total, int32_t max_k) {
  return Key{dtype, (flags & RK_EXACT) ? 1 : 0, lg2ceil(total), lg2ceil(rows), lg2ceil(max_k)};
}

rk::Tune from_c(const rk_tune& t) {
  return rk::Tune{t.tile, t.grid, t.sample_groups, t.group_elems, t.hist_bits, t.unroll};
}

}  // namespace

namespace rk {
// Tunables cached for this shape bucket, if any (used by the launch path).
bool tune_lookup(int dtype, uint32_t flags, int32_t rows, int64_t total, int32_t max_k, Tune* t) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto it = cache().find(bucket(dtype, flags, rows, total, max_k));
  if (it == cache().end()) return false;
  *t = it->second;
  return true;
}
}  // namespace rk

extern "C" {

int rk_autotune_batched(int dtype, const void* vals, const int64_t* row_off,
                        const int32_t* row_k, int32_t num_rows, int64_t total_len, int32_t max_k,
                        uint32_t flags, void* out_vals, int32_t* out_idx, const int64_t* out_off,
                        void* workspace, size_t ws_bytes, void* stream, int32_t reps,
                        rk_tune* best, float* best_ms) {
  if (reps <= 0) reps = 3;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return RK_E_HIP;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return RK_E_HIP;
  rk_tune saved;
  rk_get_tune(&saved);
  // Drop a cached entry for this bucket while measuring (the candidates are
  // applied through the default tunables).
  const Key key = bucket(dtype, flags, num_rows, total_len, max_k);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    cache().erase(key);
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return RK_E_HIP;
  // Candidates: tile size x persistent grid (blocks per CU) x loads in
  // flight, and per path the parameter that shapes its passes: the sample
  // size for the sampled path, the digit width (hence the pass count, e.g.
  // 3 x 11 or 4 x 8 bits for 32-bit keys) for the exact path.
  std::vector<rk_tune> cands;
  const bool exact = (flags & RK_EXACT) != 0;
  const std::vector<int> gms = exact ? std::vector<int>{1, 2, 4} : std::vector<int>{2, 3, 4, 8};
  for (int tile : {4096, 16384, 65536})
    for (int gm : gms)
      for (int un : {2, 4}) {
        if (exact) {
          for (int hb : {8, 11}) cands.push_back(rk_tune{tile, gm * cus, 1024, 16, hb, un});
        } else {
          for (int sg : {512, 2048})
            cands.push_back(rk_tune{tile, gm * cus, sg, 16, saved.hist_bits, un});
        }
      }
  float best_t = 1e30f;
  rk_tune best_c = saved;
  int rc = RK_OK;
  for (const rk_tune& c : cands) {
    rk_set_tune(&c);
    rc = rk_topk_batched(dtype, vals, row_off, row_k, num_rows, total_len, max_k, flags, nullptr,
                         out_vals, out_idx, out_off, workspace, ws_bytes, stream);  // warm-up
    if (rc != RK_OK) break;
    std::vector<float> ts;
    for (int r = 0; r < reps; ++r) {
      (void)hipEventRecord(e0, s);
      rc = rk_topk_batched(dtype, vals, row_off, row_k, num_rows, total_len, max_k, flags,
                           nullptr, out_vals, out_idx, out_off, workspace, ws_bytes, stream);
      (void)hipEventRecord(e1, s);
      if (rc != RK_OK) break;
      if (hipEventSynchronize(e1) != hipSuccess) {
        rc = RK_E_HIP;
        break;
      }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ts.push_back(ms);
    }
    if (rc != RK_OK) break;
    std::sort(ts.begin(), ts.end());
    const float med = ts[ts.size() / 2];
    if (med < best_t) {
      best_t = med;
      best_c = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  rk_set_tune(&saved);
  if (rc != RK_OK) return rc;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    cache()[key] = from_c(best_c);
  }
  if (best) *best = best_c;
  if (best_ms) *best_ms = best_t;
  return RK_OK;
}

int rk_tune_cache_save(const char* path) {
  if (!path) return RK_E_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return RK_E_ARG;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  std::fprintf(f, "# dtype exact lg_total lg_rows lg_k tile grid sample_groups group_elems "
                  "hist_bits unroll\n");
  for (const auto& kv : cache()) {
    const Key& k = kv.first;
    const rk::Tune& t = kv.second;
    std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %d\n", std::get<0>(k), std::get<1>(k),
                 std::get<2>(k), std::get<3>(k), std::get<4>(k), t.tile, t.grid, t.sample_groups,
                 t.group_elems, t.hist_bits, t.unroll);
  }
  std::fclose(f);
  return RK_OK;
}

int rk_tune_cache_load(const char* path) {
  if (!path) return RK_E_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) return RK_E_ARG;
  char line[512];
  int n = 0;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int v[11];
    if (std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4],
                    &v[5], &v[6], &v[7], &v[8], &v[9], &v[10]) == 11) {
      cache()[Key{v[0], v[1], v[2], v[3], v[4]}] = rk::Tune{v[5], v[6], v[7], v[8], v[9], v[10]};
      ++n;
    }
  }
  std::fclose(f);
  return n > 0 ? RK_OK : RK_E_ARG;
}

int rk_tune_cache_clear(void) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  cache().clear();
  return RK_OK;
}

int rk_tune_cache_size(void) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  return (int)cache().size();
}

// ---------------------------------------------------------------------------
// hipGraph helpers: capture any sequence of radik_amd calls on a stream.
// ---------------------------------------------------------------------------
struct rk_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

int rk_graph_begin(void* stream) {
  if (!stream) return RK_E_ARG;  // the null stream cannot be captured
  return hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal) ==
                 hipSuccess
             ? RK_OK
             : RK_E_HIP;
}

int rk_graph_end(void* stream, rk_graph** out) {
  if (!stream || !out) return RK_E_ARG;
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture(static_cast<hipStream_t>(stream), &g) != hipSuccess || !g)
    return RK_E_HIP;
  hipGraphExec_t e = nullptr;
  if (hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return RK_E_HIP;
  }
  *out = new rk_graph{g, e};
  return RK_OK;
}

int rk_graph_launch(rk_graph* g, void* stream) {
  if (!g) return RK_E_ARG;
  return hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)) == hipSuccess ? RK_OK
                                                                                 : RK_E_HIP;
}

int rk_graph_destroy(rk_graph* g) {
  if (!g) return RK_E_ARG;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return RK_OK;
}

}  // extern "C"
