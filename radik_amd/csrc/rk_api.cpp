// C-ABI layer: argument validation, workspace layout, device limits and the
// reference-surface compat entry.  Launches go through rk::run_pipeline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/radik_amd.h"
#include "rk_pipeline.hpp"

namespace rk {
hipError_t run_pipeline(int dtype, const KParams& p, hipStream_t stream, int row_grid, int grid,
                        int part, uint32_t* publish);
int lds_keys_max();
int timing_enable(int on);
int timing_read(double* ms, int32_t* launches);
bool tune_lookup(int dtype, uint32_t flags, int32_t rows, int64_t total, int32_t max_k, Tune* t);
}  // namespace rk

namespace {

using rk::KParams;
using rk::Tune;

std::mutex g_mu;
Tune g_tune{0, 0, 0, 0, 0, 0};  // 0: the default of each field (current_tune)

struct DevInfo {
  int cus = 0;
  int lds_keys = 0;
  bool ok = false;
};

DevInfo device_info() {
  static std::mutex mu;
  static std::vector<DevInfo> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return DevInfo{};
  std::lock_guard<std::mutex> lk(mu);
  if ((int)cache.size() <= dev) cache.resize(dev + 1);
  DevInfo& d = cache[dev];
  if (!d.ok) {
    int cus = 0, lds = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return DevInfo{};
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      return DevInfo{};
    // LDS of the select kernels: the key buffer plus ~10 KB of static arrays.
    int keys = rk::lds_keys_max();
    while (keys > 1024 && (size_t)keys * 8 + 10240 > (size_t)lds) keys >>= 1;
    d.cus = cus;
    d.lds_keys = keys;
    d.ok = true;
  }
  return d;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct Layout {
  size_t hdr, desc, st, tile_start, cand_base, cand_cnt, order, ghist, tile_cnt, chunk_row,
      cand, xrows, xst, total;
  uint64_t tile_cap, cand_cap;
};

// Candidate slots of a row: min(len, 4k + ceil(len/64) + 1024) (set on the
// device by k_plan_rows); summed over rows this is bounded by the expression
// below, which needs only the batch totals.
Layout make_layout(int32_t R, int64_t total_len, int32_t max_k) {
  Layout L{};
  const uint64_t Ru = (uint64_t)R;
  const uint64_t E = (uint64_t)total_len;
  L.tile_cap = E / rk::TILE_MIN + Ru + 1;
  // rows with k beyond the LDS select take no candidate slots
  const uint64_t kc = (uint64_t)std::min(max_k, rk::lds_keys_max());
  // rows with k beyond it (unsorted) bracket their k-th key: len/16 + 4096
  // slots each
  const uint64_t big = (uint64_t)max_k > kc ? E / 16 + 4096 * Ru : 0;
  const uint64_t cand_bound = 4 * Ru * kc + E / 64 + 1025 * Ru + big;
  L.cand_cap = std::min<uint64_t>(E, cand_bound);
  size_t o = 0;
  L.hdr = o; o = align256(o + sizeof(rk::Header));
  L.desc = o; o = align256(o + Ru * sizeof(rk::RowDesc));
  L.st = o; o = align256(o + Ru * sizeof(rk::RowState));
  L.tile_start = o; o = align256(o + (Ru + 1) * sizeof(uint32_t));
  L.cand_base = o; o = align256(o + (Ru + 1) * sizeof(uint64_t));
  L.cand_cnt = o; o = align256(o + Ru * sizeof(uint32_t));
  L.order = o; o = align256(o + 2 * Ru * sizeof(uint32_t));
  L.ghist = o; o = align256(o + Ru * rk::HIST_MAX_BINS * sizeof(uint32_t));
  L.tile_cnt = o; o = align256(o + L.tile_cap * sizeof(uint2));
  L.chunk_row = o; o = align256(o + L.tile_cap * sizeof(uint32_t));
  L.cand = o; o = align256(o + L.cand_cap * sizeof(uint64_t));
  L.xrows = o; o = align256(o + Ru * sizeof(uint32_t));
  L.xst = o; o = align256(o + Ru * sizeof(rk::XRow));
  L.total = o;
  return L;
}

int elem_size(int dtype) { return dtype == RK_F32 ? 4 : 2; }

bool valid_dtype(int dtype) { return dtype == RK_F32 || dtype == RK_F16 || dtype == RK_BF16; }

Tune current_tune(const DevInfo& di, int dtype, uint32_t flags, int32_t rows, int64_t total,
                  int32_t max_k) {
  Tune t;
  if (!rk::tune_lookup(dtype, flags, rows, total, max_k, &t)) {
    std::lock_guard<std::mutex> lk(g_mu);
    t = g_tune;
  }
  if (t.tile <= 0) {
    // Default 16384 keys; a call too small to give every CU two tiles gets
    // smaller tiles (down to the minimum) so that its streaming kernels
    // still spread over the chip (a single serving request, a short batch).
    int64_t tl = 16384;
    while (tl > rk::TILE_MIN && total / tl < 2 * (int64_t)di.cus) tl >>= 1;
    t.tile = (int32_t)tl;
  }
  if (t.tile < rk::TILE_MIN) t.tile = rk::TILE_MIN;
  t.tile = (t.tile / 64) * 64;
  // Default grid: the sampled path's filter keeps two load sets in flight
  // per wave and runs best at 4 workgroups per CU; the exact-pass executor
  // at 1 (2 from 2^26 keys), where its chunks' claims, histogram merges and
  // completions (device-scope round trips) contend least.
  const bool exact = (flags & RK_EXACT) != 0;
  if (t.grid <= 0) t.grid = di.cus * (exact ? rk::exact_grid_per_cu(total) : 4);
  if (t.sample_groups <= 0) t.sample_groups = 1024;
  if (t.group_elems <= 0) t.group_elems = 16;
  t.group_elems = std::min(256, std::max(8, (t.group_elems / 8) * 8));
  if (t.hist_bits < 8 || t.hist_bits > 11) t.hist_bits = 11;
  if (t.unroll != 2 && t.unroll != 4 && t.unroll != 8) t.unroll = exact ? 4 : 2;
  return t;
}

// Shared body of the batched entry points.
struct Outputs {
  void* vals = nullptr;
  void* idx = nullptr;  // int32, or int64 when idx64
  const int64_t* off = nullptr;
  void* const* val_ptrs = nullptr;
  int32_t* const* idx_ptrs = nullptr;
};

int launch(int dtype, const void* vals, const int64_t* row_off, const void* const* row_ptrs,
           const int64_t* row_len, const int32_t* row_k, int32_t K, int32_t num_rows,
           int64_t total_len, int32_t max_k, uint32_t flags, const void* idx_in,
           const Outputs& out, void* workspace, size_t ws_bytes, void* stream,
           int part = rk::PART_ALL, uint32_t* publish = nullptr, bool idx64 = false) {
  if (!valid_dtype(dtype)) return RK_E_UNSUPPORTED;
  if (num_rows < 0 || total_len < 0 || max_k < 0) return RK_E_ARG;
  if (num_rows == 0) return RK_OK;
  if (out.val_ptrs ? !out.idx_ptrs : (!out.vals || !out.idx)) return RK_E_ARG;
  void* out_vals = out.vals;
  if (row_ptrs ? !row_len : (!row_off || (!vals && total_len > 0))) return RK_E_ARG;
  const DevInfo di = device_info();
  if (!di.ok) return RK_E_HIP;
  // k beyond the LDS select: unsorted output only (exact passes + ordered emit)
  if (max_k > di.lds_keys && (flags & RK_SORTED)) return RK_E_UNSUPPORTED;
  const int es = elem_size(dtype);
  if (((uintptr_t)vals % es) || ((uintptr_t)out_vals % es)) return RK_E_ALIGN;
  const size_t isz = idx64 ? 8 : 4;
  if (((uintptr_t)idx_in % isz) || ((uintptr_t)out.idx % isz)) return RK_E_ALIGN;
  if (!workspace) return RK_E_WORKSPACE;
  if ((uintptr_t)workspace % 256) return RK_E_ALIGN;
  const Layout L = make_layout(num_rows, total_len, max_k);
  if (ws_bytes < L.total) return RK_E_WORKSPACE;

  char* ws = static_cast<char*>(workspace);
  KParams p{};
  p.vals = vals;
  p.row_off = row_off;
  p.row_k = row_k;
  p.out_off = out.off;
  p.idx_in = idx_in;
  p.row_ptrs = row_ptrs;
  p.row_len = row_len;
  p.out_val_ptrs = out.val_ptrs;
  p.out_idx_ptrs = out.idx_ptrs;
  p.out_vals = out.vals;
  p.out_idx = out.idx;
  p.idx64 = idx64 ? 1 : 0;
  p.K = K;
  p.R = num_rows;
  p.max_k = max_k;
  p.flags = flags;
  p.total_len = total_len;
  p.hdr = reinterpret_cast<rk::Header*>(ws + L.hdr);
  p.desc = reinterpret_cast<rk::RowDesc*>(ws + L.desc);
  p.st = reinterpret_cast<rk::RowState*>(ws + L.st);
  p.tile_start = reinterpret_cast<uint32_t*>(ws + L.tile_start);
  p.cand_base = reinterpret_cast<uint64_t*>(ws + L.cand_base);
  p.cand_cnt = reinterpret_cast<uint32_t*>(ws + L.cand_cnt);
  p.order = reinterpret_cast<uint32_t*>(ws + L.order);
  p.ghist = reinterpret_cast<uint32_t*>(ws + L.ghist);
  p.tile_cnt = reinterpret_cast<uint2*>(ws + L.tile_cnt);
  p.chunk_row = reinterpret_cast<uint32_t*>(ws + L.chunk_row);
  p.cand = reinterpret_cast<uint64_t*>(ws + L.cand);
  p.xrows = reinterpret_cast<uint32_t*>(ws + L.xrows);
  p.xst = reinterpret_cast<rk::XRow*>(ws + L.xst);
  p.tile_capacity = L.tile_cap;
  p.cand_capacity = L.cand_cap;
  // LDS of the select kernels sized to this call: the k selected keys (and
  // rows up to that length) live in LDS; larger candidate sets are selected
  // from global memory.  At least 8192 keys (64 KB): the select kernel's
  // registers hold it at two workgroups per CU, which leave 80 KB of LDS
  // each, so the smaller buffer would buy no occupancy.
  int lk = 8192;
  while (lk < max_k && lk < di.lds_keys) lk <<= 1;
  p.lds_keys = std::min(lk, di.lds_keys);
  p.small_max = p.lds_keys;
  p.tune = current_tune(di, dtype, flags, num_rows, total_len, max_k);
  // at least 64 runs fit (the sample's minimum), at most 16384 keys of LDS
  p.sample_keys = (int32_t)std::min<int64_t>(
      16384, (int64_t)std::max(p.tune.sample_groups, 64) * p.tune.group_elems);
  p.round = 0;
  const int row_grid = std::min(num_rows, 8192);
  const uint64_t tiles_bound = (uint64_t)total_len / (uint64_t)p.tune.tile + (uint64_t)num_rows;
  const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)p.tune.grid, tiles_bound));
  // The filter's workgroups take chunks of tiles from a queue.  Chunks of
  // about total / (8 x grid) bytes, within [256 KB, 2 MB]: large enough to
  // amortise a chunk's start-up latency, small enough that all workgroups
  // finish within about one chunk of each other (the last quarter of the
  // tiles goes in quarter-size chunks).  Small calls use smaller chunks so
  // that every workgroup still gets two.
  int64_t chunk_bytes = int64_t(256) << 10;
  while (chunk_bytes < (int64_t(2) << 20) && chunk_bytes * 8 * grid < total_len * es)
    chunk_bytes <<= 1;
  int64_t ct = std::max<int64_t>(1, chunk_bytes / ((int64_t)p.tune.tile * es));
  while (ct > 1 && (int64_t)tiles_bound / ct < 2 * (int64_t)grid) ct >>= 1;
  p.chunk_tiles = (int32_t)ct;
  // The exact-pass executor hands rows out in chunks of about two per
  // workgroup of the call's whole work, but of at least 16K keys (or 1/64
  // of a small call), within [4K, 1M] keys: each chunk pays a few
  // device-scope round trips (claim, histogram merge, completion), which
  // larger chunks amortise, and a single row still splits into many
  // (measured on MI355X: 16K-key chunks beat 4K ones from 2^20 keys, 4K
  // beat 16K on a 2^17-key row).
  {
    const int64_t xg = (flags & RK_EXACT)
                           ? std::max<int64_t>(1, std::min<int64_t>(grid, (int64_t)di.cus * 4))
                           : (int64_t)di.cus * rk::exact_grid_per_cu(total_len);
    const int64_t want = std::max<int64_t>(total_len / (2 * xg),
                                           std::min<int64_t>(16384, total_len / 64));
    const int64_t xe = std::min<int64_t>(int64_t(1) << 20, std::max<int64_t>(4096, want));
    p.x_chunk_tiles = (int32_t)std::max<int64_t>(1, (xe + p.tune.tile - 1) / p.tune.tile);
  }
  const hipError_t e =
      rk::run_pipeline(dtype, p, static_cast<hipStream_t>(stream), row_grid, grid, part, publish);
  return e == hipSuccess ? RK_OK : RK_E_HIP;
}

}  // namespace

namespace rk {

// rk_topk_rows issuing one part of the launch sequence (serving loop).  The
// two parts of one call must be issued with identical arguments, in order,
// on one stream; the workspace header (its first 256 bytes) tells between
// them whether PART_FALLBACK has work.  With PART_MAIN, `publish` (device
// address of mapped host memory, or null) receives the header's first 8
// words when the part completes.
int topk_rows_part(int dtype, const void* const* row_ptrs, const int64_t* row_len,
                   const int32_t* row_k, int32_t num_rows, int64_t total_len, int32_t max_k,
                   uint32_t flags, void* const* out_val_ptrs, int32_t* const* out_idx_ptrs,
                   void* workspace, size_t ws_bytes, void* stream, int part,
                   uint32_t* publish) {
  if (!valid_dtype(dtype)) return RK_E_UNSUPPORTED;
  if (!row_ptrs || !row_len || !row_k || !out_val_ptrs || !out_idx_ptrs) return RK_E_ARG;
  if (part != PART_ALL && (flags & RK_EXACT)) return RK_E_ARG;
  Outputs o;
  o.val_ptrs = out_val_ptrs;
  o.idx_ptrs = out_idx_ptrs;
  return launch(dtype, nullptr, nullptr, row_ptrs, row_len, row_k, 0, num_rows, total_len,
                max_k, flags, nullptr, o, workspace, ws_bytes, stream, part,
                part == PART_MAIN ? publish : nullptr);
}

}  // namespace rk

// ---------------------------------------------------------------------------
// Allocation-free entry.  Its workspace is this file-scope device array, one
// copy per GPU (the code object's storage on that GPU, held until the
// process ends).  In host code the array's name is only a host-side stand-in:
// the device address is taken once per device with hipGetSymbolAddress, with
// that device current, and only that address is ever passed to a kernel.
// Nothing in it must start at zero: the plan kernel writes every header
// field, row state, counter and table entry a later kernel reads, so the
// previous call's contents are harmless.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[RK_NOALLOC_SCRATCH_BYTES];

namespace {

// Device address of `scratch` on the current device (null on a HIP error).
void* scratch_address() {
  static std::mutex mu;
  static std::vector<void*> addr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
  if (!addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    addr[dev] = p;
  }
  return addr[dev];
}

}  // namespace

extern "C" {

int32_t rk_max_k(void) {
  const DevInfo di = device_info();
  return di.ok ? di.lds_keys : rk::lds_keys_max();
}

int rk_topk_workspace_size(int dtype, int32_t num_rows, int64_t total_len, int32_t max_k,
                           uint32_t flags, size_t* bytes) {
  (void)flags;
  if (!bytes) return RK_E_ARG;
  if (!valid_dtype(dtype)) return RK_E_UNSUPPORTED;
  if (num_rows < 0 || total_len < 0 || max_k < 0) return RK_E_ARG;
  *bytes = make_layout(num_rows, total_len, max_k).total;
  return RK_OK;
}

int rk_topk_batched(int dtype, const void* vals, const int64_t* row_off, const int32_t* row_k,
                    int32_t num_rows, int64_t total_len, int32_t max_k, uint32_t flags,
                    const int32_t* idx_in, void* out_vals, int32_t* out_idx,
                    const int64_t* out_off, void* workspace, size_t ws_bytes, void* stream) {
  if (!valid_dtype(dtype)) return RK_E_UNSUPPORTED;
  if (!row_k || !out_off) return RK_E_ARG;
  Outputs o;
  o.vals = out_vals;
  o.idx = out_idx;
  o.off = out_off;
  return launch(dtype, vals, row_off, nullptr, nullptr, row_k, 0, num_rows, total_len, max_k,
                flags, idx_in, o, workspace, ws_bytes, stream);
}

int rk_topk_batched_i64(int dtype, const void* vals, const int64_t* row_off,
                        const int32_t* row_k, int32_t num_rows, int64_t total_len,
                        int32_t max_k, uint32_t flags, const int64_t* idx_in, void* out_vals,
                        int64_t* out_idx, const int64_t* out_off, void* workspace,
                        size_t ws_bytes, void* stream) {
  if (!valid_dtype(dtype)) return RK_E_UNSUPPORTED;
  if (!row_k || !out_off) return RK_E_ARG;
  Outputs o;
  o.vals = out_vals;
  o.idx = out_idx;
  o.off = out_off;
  return launch(dtype, vals, row_off, nullptr, nullptr, row_k, 0, num_rows, total_len, max_k,
                flags, idx_in, o, workspace, ws_bytes, stream, rk::PART_ALL, nullptr, true);
}

int rk_topk_rows(int dtype, const void* const* row_ptrs, const int64_t* row_len,
                 const int32_t* row_k, int32_t num_rows, int64_t total_len, int32_t max_k,
                 uint32_t flags, void* const* out_val_ptrs, int32_t* const* out_idx_ptrs,
                 void* workspace, size_t ws_bytes, void* stream) {
  if (!valid_dtype(dtype)) return RK_E_UNSUPPORTED;
  if (!row_ptrs || !row_len || !row_k || !out_val_ptrs || !out_idx_ptrs) return RK_E_ARG;
  Outputs o;
  o.val_ptrs = out_val_ptrs;
  o.idx_ptrs = out_idx_ptrs;
  return launch(dtype, nullptr, nullptr, row_ptrs, row_len, row_k, 0, num_rows, total_len,
                max_k, flags, nullptr, o, workspace, ws_bytes, stream);
}

int rk_compat_workspace_size(int dtype, int32_t K, int32_t n, int32_t taskNum, uint32_t flags,
                             size_t* bytes) {
  if (!bytes) return RK_E_ARG;
  if (K < 0 || n < 0 || taskNum < 0) return RK_E_ARG;
  size_t b = 0;
  const int rc = rk_topk_workspace_size(dtype, taskNum, (int64_t)n * taskNum, K, flags, &b);
  if (rc != RK_OK) return rc;
  *bytes = b + align256(((size_t)taskNum + 1) * sizeof(int64_t));
  return RK_OK;
}

}  // extern "C"

namespace {

int topk_compat(int dtype, const void* valIn, const void* idxIn, void* valOut, void* idxOut,
                void* workspace, size_t ws_bytes, const int32_t* taskLen, int32_t taskNum,
                int32_t K, uint32_t flags, void* stream, bool idx64) {
  if (!taskLen || taskNum < 0 || K < 0) return RK_E_ARG;
  if (!workspace) return RK_E_WORKSPACE;  // topk_radixselect.h:77-79
  if (taskNum == 0) return RK_OK;
  std::vector<int64_t> off((size_t)taskNum + 1, 0);
  int32_t longest = 0;
  for (int32_t i = 0; i < taskNum; ++i) {
    if (taskLen[i] < 0) return RK_E_ARG;
    off[i + 1] = off[i] + taskLen[i];
    longest = std::max(longest, taskLen[i]);
  }
  size_t need = 0;
  int rc = rk_compat_workspace_size(dtype, K, longest, taskNum, flags, &need);
  if (rc != RK_OK) return rc;
  if (ws_bytes < need) return RK_E_WORKSPACE;
  const size_t off_bytes = align256(((size_t)taskNum + 1) * sizeof(int64_t));
  char* ws = static_cast<char*>(workspace);
  int64_t* d_off = reinterpret_cast<int64_t*>(ws);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(d_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice, s) !=
      hipSuccess)
    return RK_E_HIP;
  Outputs o;
  o.vals = valOut;
  o.idx = idxOut;
  return launch(dtype, valIn, d_off, nullptr, nullptr, nullptr, K, taskNum,
                (int64_t)longest * taskNum, K, flags, idxIn, o, ws + off_bytes,
                ws_bytes - off_bytes, stream, rk::PART_ALL, nullptr, idx64);
}

}  // namespace

extern "C" {

int rk_topk_compat(int dtype, const void* valIn, const int32_t* idxIn, void* valOut,
                   int32_t* idxOut, void* workspace, size_t ws_bytes, const int32_t* taskLen,
                   int32_t taskNum, int32_t K, uint32_t flags, void* stream) {
  return topk_compat(dtype, valIn, idxIn, valOut, idxOut, workspace, ws_bytes, taskLen, taskNum,
                     K, flags, stream, false);
}

int rk_topk_compat_i64(int dtype, const void* valIn, const int64_t* idxIn, void* valOut,
                       int64_t* idxOut, void* workspace, size_t ws_bytes, const int32_t* taskLen,
                       int32_t taskNum, int32_t K, uint32_t flags, void* stream) {
  return topk_compat(dtype, valIn, idxIn, valOut, idxOut, workspace, ws_bytes, taskLen, taskNum,
                     K, flags, stream, true);
}

int rk_topk_batched_noalloc(int dtype, const void* vals, const int64_t* row_off,
                            const int32_t* row_k, int32_t num_rows, int64_t total_len,
                            int32_t max_k, uint32_t flags, const int32_t* idx_in, void* out_vals,
                            int32_t* out_idx, const int64_t* out_off, void* stream) {
  if (!valid_dtype(dtype)) return RK_E_UNSUPPORTED;
  if (!row_k || !out_off) return RK_E_ARG;
  if (num_rows < 0 || total_len < 0 || max_k < 0) return RK_E_ARG;
  if (make_layout(num_rows, total_len, max_k).total > (size_t)RK_NOALLOC_SCRATCH_BYTES)
    return RK_E_SCRATCH;
  void* ws = scratch_address();
  if (!ws) return RK_E_HIP;
  Outputs o;
  o.vals = out_vals;
  o.idx = out_idx;
  o.off = out_off;
  return launch(dtype, vals, row_off, nullptr, nullptr, row_k, 0, num_rows, total_len, max_k,
                flags, idx_in, o, ws, (size_t)RK_NOALLOC_SCRATCH_BYTES, stream);
}

int rk_topk_status(const void* workspace, void* stream, uint32_t* status, uint32_t* fallback_rows,
                   uint32_t* ordered_rows) {
  if (!workspace) return RK_E_WORKSPACE;
  rk::Header h{};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (hipMemcpyAsync(&h, workspace, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess)
    return RK_E_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return RK_E_HIP;
  if (status) *status = h.status;
  if (fallback_rows) *fallback_rows = h.fallback_rows;
  if (ordered_rows) *ordered_rows = h.ordered_rows;
  return h.status ? RK_E_DEVICE : RK_OK;
}

const char* rk_strerror(int code) {
  switch (code) {
    case RK_OK: return "ok";
    case RK_E_ARG: return "bad argument";
    case RK_E_WORKSPACE: return "workspace missing or too small";
    case RK_E_UNSUPPORTED: return "unsupported dtype, or sorted output with k above rk_max_k()";
    case RK_E_ALIGN: return "pointer not aligned to its element size (workspace: 256 bytes)";
    case RK_E_HIP: return "HIP runtime error";
    case RK_E_DEVICE: return "device-side validation failed";
    case RK_E_IO: return "batch or result file missing, unreadable or malformed";
    case RK_E_SCRATCH:
      return "call needs more workspace than the allocation-free entry's "
             "RK_NOALLOC_SCRATCH_BYTES";
    default: return "unknown error";
  }
}

int rk_timing_enable(int on) { return rk::timing_enable(on) == 0 ? RK_OK : RK_E_HIP; }

int rk_timing_read(double* filter_ms, int32_t* launches) {
  return rk::timing_read(filter_ms, launches) == 0 ? RK_OK : RK_E_HIP;
}

int rk_get_tune(rk_tune* out) {
  if (!out) return RK_E_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  out->tile = g_tune.tile;
  out->grid = g_tune.grid;
  out->sample_groups = g_tune.sample_groups;
  out->group_elems = g_tune.group_elems;
  out->hist_bits = g_tune.hist_bits;
  out->unroll = g_tune.unroll;
  return RK_OK;
}

int rk_set_tune(const rk_tune* in) {
  if (!in) return RK_E_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  // < 0 keeps the current value; 0 selects the default
  if (in->tile >= 0) g_tune.tile = in->tile;
  if (in->grid >= 0) g_tune.grid = in->grid;
  if (in->sample_groups >= 0) g_tune.sample_groups = in->sample_groups;
  if (in->group_elems >= 0) g_tune.group_elems = in->group_elems;
  if (in->hist_bits >= 0) g_tune.hist_bits = in->hist_bits;
  if (in->unroll >= 0) g_tune.unroll = in->unroll;
  return RK_OK;
}

}  // extern "C"
