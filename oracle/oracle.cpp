// ORACLE — test infrastructure only, never part of the product path.
//
// CPU restatement of RadiK's batched radix top-k semantics, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
// Nothing in radik_amd/ links, loads or calls this file.
//
// What it restates (reference files, relative to the RadiK repository):
//   * order-preserving float -> unsigned key flip:
//       radik/RadixSelect/utils.cuh:329-334 (getBinId's mask/xor)
//   * NaN never selected (histogram skips NaN):
//       radik/RadixSelect/radixselect_l.cuh:189-194; utils.cuh:17-31 (NaN doc)
//   * per-row top-k, copy-all when len <= K, tail fill idx=-1 / val=0:
//       radik/RadixSelect/topk_radixselect.h:59-323 (esp. :312-320),
//       radixselect_l.cuh:518-565 (len <= K copies the whole row)
//   * optional sort of the k survivors, ASCEND / DESCEND:
//       radik/RadixSelect/bitonic_sort.cuh:195-233, topk_radixselect.h:263-307
//   * the sort-based CPU verifier this mirrors:
//       radik/test/test_radik.cu:188-213 (std::sort of (value, index) pairs)
// Refinements the reference leaves undefined (it is racy there,
// radixselect_l.cuh:417-421): ties at the k-th key are broken by the lower
// row position; sorted output is ordered by key, ties by position ascending;
// unsorted output is in position order; NaN ranks below every non-NaN value
// in both directions (it is selected only when fewer than k non-NaN exist).
//
// Parity pinning: tests/test_oracle.py checks every function here against
// the committed brute-force fixtures in tests/golden/ (written by a
// pure-Python float-comparison implementation that shares no code with this
// file), the key-flip known answers, an exhaustive 2^16 check of the 16-bit
// flips against numpy's float order, and closed-form special cases
// (k = 1 -> argmax/argmin, k = len -> full stable sort, all-equal rows).
//
// The multithreaded orc_topk is also the CPU baseline (std::nth_element +
// std::sort over (key, index) pairs, one row per worker thread).

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

enum { DT_F32 = 0, DT_F16 = 1, DT_BF16 = 2 };
enum : uint32_t { F_LARGEST = 1u, F_SORTED = 2u, F_ASCEND = 4u };

// Order-preserving flip of an IEEE bit pattern to an unsigned key
// (utils.cuh:329-334): positive values get the sign bit set, negative values
// are bit-inverted, so unsigned order equals numeric order with -0 < +0.
inline uint32_t flip_bits(int dt, uint32_t u) {
  if (dt == DT_F32) {
    uint32_t mask = (u & 0x80000000u) ? 0xffffffffu : 0x80000000u;
    return u ^ mask;
  }
  u &= 0xffffu;
  uint32_t mask = (u & 0x8000u) ? 0xffffu : 0x8000u;
  return (u ^ mask) & 0xffffu;
}

inline bool is_nan_bits(int dt, uint32_t u) {
  if (dt == DT_F32) return (u & 0x7fffffffu) > 0x7f800000u;
  if (dt == DT_F16) return (u & 0x7fffu) > 0x7c00u;
  return (u & 0x7fffu) > 0x7f80u;  // bf16
}

inline int key_bits(int dt) { return dt == DT_F32 ? 32 : 16; }

// Selection key: larger is better in both directions; NaN maps to 0 (worst).
inline uint32_t sel_key(int dt, uint32_t u, bool largest) {
  if (is_nan_bits(dt, u)) return 0u;
  uint32_t f = flip_bits(dt, u);
  if (!largest) f = (dt == DT_F32) ? ~f : (f ^ 0xffffu);
  return f;
}

inline uint32_t load_bits(int dt, const void* vals, int64_t i) {
  if (dt == DT_F32) return static_cast<const uint32_t*>(vals)[i];
  return static_cast<const uint16_t*>(vals)[i];
}

inline void store_bits(int dt, void* out, int64_t i, uint32_t u) {
  if (dt == DT_F32)
    static_cast<uint32_t*>(out)[i] = u;
  else
    static_cast<uint16_t*>(out)[i] = static_cast<uint16_t>(u);
}

struct Item {
  uint32_t key;  // selection key (larger = better)
  uint32_t pos;  // row position
};

// One row: select the min(len, k) best (key desc, pos asc), then order them.
void topk_row(int dt, const void* vals, int64_t off, int64_t len, int64_t k,
              uint32_t flags, const int32_t* idx_in, void* out_vals,
              int32_t* out_idx, int64_t out_pos, int64_t slots,
              std::vector<Item>& buf) {
  const bool largest = flags & F_LARGEST;
  const bool sorted = flags & F_SORTED;
  const bool ascend = flags & F_ASCEND;
  const int64_t m = std::max<int64_t>(0, std::min(len, k));
  buf.resize(static_cast<size_t>(len));
  for (int64_t i = 0; i < len; ++i)
    buf[i] = Item{sel_key(dt, load_bits(dt, vals, off + i), largest),
                  static_cast<uint32_t>(i)};
  auto better = [](const Item& a, const Item& b) {
    return a.key != b.key ? a.key > b.key : a.pos < b.pos;
  };
  if (m < len && m > 0)
    std::nth_element(buf.begin(), buf.begin() + (m - 1), buf.end(), better);
  // buf[0, m) now holds exactly the m best (nth_element partitions around
  // the m-th element under a strict total order, since positions are unique).
  if (sorted) {
    // best-first when (largest && descending) or (smallest && ascending).
    const bool best_first = (largest != ascend);
    if (best_first) {
      std::sort(buf.begin(), buf.begin() + m, better);
    } else {
      std::sort(buf.begin(), buf.begin() + m, [](const Item& a, const Item& b) {
        return a.key != b.key ? a.key < b.key : a.pos < b.pos;
      });
    }
  } else {
    std::sort(buf.begin(), buf.begin() + m,
              [](const Item& a, const Item& b) { return a.pos < b.pos; });
  }
  for (int64_t i = 0; i < m; ++i) {
    const uint32_t p = buf[i].pos;
    store_bits(dt, out_vals, out_pos + i, load_bits(dt, vals, off + p));
    out_idx[out_pos + i] = idx_in ? idx_in[off + p] : static_cast<int32_t>(p);
  }
  // Tail fill for rows shorter than their slot count (topk_radixselect.h:312-320).
  for (int64_t i = m; i < slots; ++i) {
    store_bits(dt, out_vals, out_pos + i, 0u);
    out_idx[out_pos + i] = -1;
  }
}

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

uint32_t orc_flip(int dt, uint32_t bits) { return flip_bits(dt, bits); }

uint32_t orc_sel_key(int dt, uint32_t bits, int largest) {
  return sel_key(dt, bits, largest != 0);
}

// Batched top-k.  Row r is vals[row_off[r] .. row_off[r+1]), its k is
// row_k[r] (or K when row_k is null).  Output row r occupies
// [out_off[r], out_off[r+1]) (or [r*K, (r+1)*K) when out_off is null); the
// first min(len, k) slots receive the selection, the rest are tail-filled.
// Returns 0, or -1 on a bad argument.
int orc_topk(int dt, const void* vals, const int64_t* row_off,
             const int32_t* row_k, int32_t K, int32_t num_rows, uint32_t flags,
             const int32_t* idx_in, void* out_vals, int32_t* out_idx,
             const int64_t* out_off, int32_t nthreads) {
  if (dt < 0 || dt > 2 || num_rows < 0 || !row_off) return -1;
  if (!row_k && K < 0) return -1;
  if (!out_off && K < 0) return -1;
  const int nt = std::max(1, std::min<int>(nthreads, num_rows > 0 ? num_rows : 1));
  std::atomic<int32_t> next{0};
  auto worker = [&]() {
    std::vector<Item> buf;
    for (;;) {
      int32_t r = next.fetch_add(1);
      if (r >= num_rows) break;
      const int64_t off = row_off[r];
      const int64_t len = row_off[r + 1] - off;
      const int64_t k = row_k ? row_k[r] : K;
      const int64_t op = out_off ? out_off[r] : static_cast<int64_t>(r) * K;
      const int64_t slots = out_off ? out_off[r + 1] - out_off[r] : K;
      topk_row(dt, vals, off, len, k, flags, idx_in, out_vals, out_idx, op,
               slots, buf);
    }
  };
  if (nt == 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int i = 0; i < nt; ++i) th.emplace_back(worker);
    for (auto& t : th) t.join();
  }
  return 0;
}

int orc_hardware_threads(void) {
  unsigned n = std::thread::hardware_concurrency();
  return n ? static_cast<int>(n) : 1;
}

}  // extern "C"
