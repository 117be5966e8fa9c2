// gfx950 kernels of the batched radix top-k and their launch sequence.
//
// The pipeline replaces radik/RadixSelect/topk_radixselect.h:59-323
// (topKRadixSelectL) with a fixed launch sequence whose per-row decisions
// live in device memory (RowState.phase), so the whole call is stream
// ordered, has no host round trip and can be captured in a hipGraph:
//
//   plan_rows, plan_scan            row descriptors, tile list, candidate slots
//                                   (one-row sampled calls: done by the sample
//                                   kernel's one workgroup instead)
//   sample                          long rows: threshold T from a sample
//   filter                          one streaming read: compact keys >= T
//   select_finalize                 per row: sort (or radix select, then
//                                   sort) of the candidates; write
//   merge_sort, merge_rank          (max_k > 4096) rows with k > 4096: the
//                                   sort split over workgroups
//   [exact_init], exact             the exact RadiK passes for rows the
//                                   sampled path could not settle (and the
//                                   plain path) in one persistent executor:
//                                   digit histograms, then a filter read or
//                                   the ordered emit's count and write reads
//   select_finalize                 second round for those rows, and the
//                                   ordered rows' index mapping / tail fill
//
// Selection order everywhere: the 64-bit composite key
//   (sel_key << 32) | ~row_position
// which is unique per element, so "the k largest composite keys" is the
// deterministic lowest-position tie-break that radixselect_l.cuh:417-421
// leaves to an atomic race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "rk_device.hpp"
#include "rk_pipeline.hpp"

namespace rk {

// Optional phase timing for tuning builds (-DRK_PHASE_TIMING): wall-clock
// stamps per row into that row's (otherwise idle) histogram slot.
#ifdef RK_PHASE_TIMING
#define RK_STAMP(p, r, slot)                                                     \
  do {                                                                           \
    if (threadIdx.x == 0) (p).ghist[(size_t)(r) * HIST_MAX_BINS + 64 + (slot)] = \
        (uint32_t)wall_clock64();                                                \
  } while (0)
#else
#define RK_STAMP(p, r, slot) do {} while (0)
#endif

constexpr int SEL_NT = 512;          // threads of the per-row select kernels
constexpr int LDS_KEYS_MAX = 16384;  // compile-time bound of the LDS key buffer
constexpr int STREAM_NT = 256;       // threads of the streaming kernels


// ---------------------------------------------------------------------------
// Work partition of the exact-pass and ordered-emit kernels: the tiles of
// all rows form one flat list (tile_start is its per-row prefix).
// Workgroup w takes the contiguous range [T*w/G, T*(w+1)/G), so its tiles of
// one row form one contiguous element range (one merge of its histogram per
// row), and the partition is identical in every launch with the same grid.
// ---------------------------------------------------------------------------
template <typename F>
RK_DEV void for_each_segment(const KParams& p, F&& f) {
  const uint32_t T = p.tile_start[p.R];
  const uint64_t G = gridDim.x, w = blockIdx.x;
  const uint32_t t0 = (uint32_t)((uint64_t)T * w / G);
  const uint32_t t1 = (uint32_t)((uint64_t)T * (w + 1) / G);
  if (t0 >= t1) return;
  int lo = 0, hi = p.R;  // invariant: tile_start[lo] <= t0 < tile_start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (p.tile_start[mid] <= t0) lo = mid;
    else hi = mid;
  }
  int r = lo;
  uint32_t t = t0;
  while (t < t1 && r < p.R) {
    const uint32_t rs = p.tile_start[r], re = p.tile_start[r + 1];
    if (re <= t) {
      ++r;
      continue;
    }
    const uint32_t se = re < t1 ? re : t1;
    f(r, t - rs, se - rs);
    t = se;
    ++r;
  }
}

// Row holding tile t: the last r in [0, R) with tile_start[r] <= t (rows
// without tiles never qualify for t < tile_start[R]).  A 64-way search by
// one whole wave: each step loads 64 evenly spaced prefix entries and keeps
// the span between the last one <= t and the next, so R rows take
// ceil(log64 R) dependent loads (2 for 2048 rows).
RK_DEV int find_row_wave(const uint32_t* tile_start, int R, uint32_t t) {
  const int lane = lane_id();
  int lo = 0, hi = R;
  while (hi - lo > 1) {
    const int64_t span = hi - lo;
    const int pos = lo + (int)((span * lane) / 64);
    const uint64_t m = __ballot(ldg(tile_start + pos) <= t);  // a prefix of the lanes
    const int last = 63 - __builtin_clzll(m);
    const int nlo = lo + (int)((span * last) / 64);
    hi = last == 63 ? hi : lo + (int)((span * (last + 1)) / 64);
    lo = nlo;
  }
  return lo;
}

// The filter's partition of the same flat tile list into chunks of
// chunk_tiles tiles, guided: the first ~3/4 of the tiles in chunks of CT,
// the rest in chunks of CT/4, so that the last chunks (whose end sets the
// kernel's) are short.
struct ChunkGrid {
  uint32_t T, CT, CS, n_big, t_s, nch;
  RK_DEV ChunkGrid(uint32_t T_, uint32_t CT_) : T(T_), CT(CT_) {
    CS = CT >= 4 ? CT / 4 : CT;
    n_big = CS < CT ? (uint32_t)(((uint64_t)T * 3 / 4) / CT) : (T + CT - 1) / CT;
    t_s = n_big * CT < T ? n_big * CT : T;
    nch = n_big + (T - t_s + CS - 1) / CS;
  }
  RK_DEV uint32_t start(uint32_t c) const { return c < n_big ? c * CT : t_s + (c - n_big) * CS; }
  RK_DEV uint32_t size(uint32_t c) const { return c < n_big ? CT : CS; }
  // first chunk whose first tile is >= t (t <= T)
  RK_DEV uint32_t first_ge(uint32_t t) const {
    return t <= t_s ? (t + CT - 1) / CT : n_big + (t - t_s + CS - 1) / CS;
  }
};
constexpr uint32_t NO_ROW = 0xffffffffu;

// Records row r as the row of every chunk that starts inside it (called by
// a whole workgroup of NT threads for each sampled row).
template <int NT>
RK_DEV void record_chunk_rows(const KParams& p, int r) {
  const ChunkGrid g(p.tile_start[p.R], (uint32_t)p.chunk_tiles);
  const uint32_t c0 = g.first_ge(p.tile_start[r]), c1 = g.first_ge(p.tile_start[r + 1]);
  for (uint32_t c = c0 + threadIdx.x % NT; c < c1; c += NT) p.chunk_row[c] = (uint32_t)r;
}

// Chunks are taken from a queue (one atomic per chunk): workgroups that
// stream faster take more of them, so all finish within about one chunk of
// each other (a static slice per workgroup, or static round-robin chunks,
// measured up to 15% slower on MI355X, the slowest workgroup setting the
// kernel's end).  A chunk's start-up would otherwise wait on dependent
// loads at the memory latency of a saturated chip: a workgroup claims its
// chunks two ahead and loads the next chunk's row from chunk_row while it
// reads the current one.  Chunks without a recorded row (rows that took the
// exact passes without sampling) find it with the 64-way search.  When
// there are no more chunks than workgroups (small calls), chunk =
// workgroup index.
template <typename F>
RK_DEV void for_each_chunk_queued(const KParams& p, uint32_t* q, F&& f) {
  __shared__ uint32_t c_sh, cn_sh, rn_sh;
  __shared__ int row_sh;
  const ChunkGrid g(p.tile_start[p.R], (uint32_t)p.chunk_tiles);
  const uint32_t nch = g.nch;
  const bool direct = nch <= gridDim.x;
  const bool wave0 = threadIdx.x < 64;
  uint32_t c = blockIdx.x, cn = nch;
  if (!direct) {
    if (threadIdx.x == 0) {
      const uint32_t a = atomicAdd(q, 2u);  // the first two chunks
      c_sh = a;
      cn_sh = a + 1;
    }
    __syncthreads();
    c = c_sh;
    cn = cn_sh;
  }
  if (c >= nch) return;  // uniform: c is the same for the whole workgroup
  if (wave0) {
    uint32_t r0 = ldg(p.chunk_row + c);
    if (r0 == NO_ROW) r0 = (uint32_t)find_row_wave(p.tile_start, p.R, g.start(c));
    if (threadIdx.x == 0) row_sh = (int)r0;
  }
  __syncthreads();
  for (;;) {
    int r = row_sh;
    uint32_t cnn = nch, rn = NO_ROW;
    if (!direct && threadIdx.x == 0) {
      // issued now, consumed after this chunk: they complete while it streams
      cnn = atomicAdd(q, 1u);
      if (cn < nch) rn = ldg(p.chunk_row + cn);
    }
    const uint32_t t0 = g.start(c);
    const uint32_t t1 = (g.T - t0) < g.size(c) ? g.T : t0 + g.size(c);
    uint32_t t = t0;
    while (t < t1 && r < p.R) {
      const uint32_t rs = p.tile_start[r], re = p.tile_start[r + 1];
      if (re <= t) {
        ++r;
        continue;
      }
      const uint32_t se = re < t1 ? re : t1;
      f(r, t - rs, se - rs);
      t = se;
      ++r;
    }
    if (direct) break;
    if (threadIdx.x == 0) {
      c_sh = cn;
      cn_sh = cnn;
      rn_sh = rn;
    }
    __syncthreads();
    c = c_sh;
    cn = cn_sh;
    if (c >= nch) break;  // uniform
    if (wave0) {
      uint32_t r1 = rn_sh;
      if (r1 == NO_ROW) r1 = (uint32_t)find_row_wave(p.tile_start, p.R, g.start(c));
      if (threadIdx.x == 0) row_sh = (int)r1;
    }
    __syncthreads();
  }
}

// First element index >= es whose address is 16-byte aligned.
template <int DT>
RK_DEV int64_t aligned_start(const typename Traits<DT>::U* rb, int64_t es) {
  constexpr int ES = sizeof(typename Traits<DT>::U);
  const uint64_t addr = (uint64_t)(rb + es);
  return es + (int64_t)(((16u - (addr & 15u)) & 15u) / ES);
}

// Output indices are int32 or int64 (the reference's IdxType,
// topk_radixselect.h:59-63), chosen per call (KParams.idx64): the caller's
// index of row position pos, and the stores into the row's index slots.
RK_DEV int64_t caller_idx(const KParams& p, const RowDesc& d, uint32_t pos) {
  if (!d.idx_in) return (int64_t)pos;
  return p.idx64 ? ldg(static_cast<const int64_t*>(d.idx_in) + pos)
                 : (int64_t)ldg(static_cast<const int32_t*>(d.idx_in) + pos);
}
RK_DEV void put_idx(const KParams& p, const RowDesc& d, int64_t o, int64_t v) {
  if (p.idx64) static_cast<int64_t*>(d.out_i)[o] = v;
  else static_cast<int32_t*>(d.out_i)[o] = (int32_t)v;
}
RK_DEV int64_t get_idx(const KParams& p, const RowDesc& d, int64_t o) {
  return p.idx64 ? static_cast<const int64_t*>(d.out_i)[o]
                 : (int64_t) static_cast<const int32_t*>(d.out_i)[o];
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
// What the scan of k_plan_scan needs of a planned row (kept in registers by
// the fused plan, so it makes no second round trip to memory).
struct RowPlan {
  int64_t len;
  int32_t k;
  uint32_t tiles, cap, phase;
  bool bad;
};

template <int DT>
RK_DEV RowPlan plan_row(const KParams& p, int r) {
  using U = typename Traits<DT>::U;
  const U* base;
  const void* idx = nullptr;
  const int64_t isz = p.idx64 ? 8 : 4;  // bytes per index
  int64_t len;
  if (p.row_ptrs) {
    base = static_cast<const U*>(p.row_ptrs[r]);
    len = p.row_len[r];
  } else {
    const int64_t off = p.row_off[r];
    len = p.row_off[r + 1] - off;
    base = static_cast<const U*>(p.vals) + off;
    if (p.idx_in) idx = static_cast<const char*>(p.idx_in) + off * isz;
  }
  const int64_t k = p.row_k ? p.row_k[r] : p.K;
  int64_t slots;
  RowDesc d;
  if (p.out_val_ptrs) {  // pointer form: one output block of k slots per row
    slots = k;
    d.out_v = p.out_val_ptrs[r];
    d.out_i = p.out_idx_ptrs[r];
  } else {
    const int64_t out_pos = p.out_off ? p.out_off[r] : (int64_t)r * p.K;
    slots = p.out_off ? p.out_off[r + 1] - p.out_off[r] : p.K;
    d.out_v = static_cast<U*>(p.out_vals) + out_pos;
    d.out_i = static_cast<char*>(p.out_idx) + out_pos * isz;
  }
  const int64_t m = len < k ? len : k;
  d.base = base;
  d.idx_in = idx;
  d.len = len;
  d.k = (int32_t)k;
  d.slots = (int32_t)slots;
  RowState s;
  s.flags = 0;
  s.T = s.prefix = s.resolved = s.rem = s.V = 0;
  s.cap = 0;
  s.why = FB_NONE;
  uint32_t tiles = 0;
  const bool bad = len < 0 || k < 0 || k > p.max_k || m > slots || slots > 0x7fffffff ||
                   len > 0x7fffffffLL;
  if (bad) {
    s.phase = PH_DONE;
    s.flags = 0x80000000u;  // reported by plan_scan
  } else if (len <= k || k == 0) {
    s.phase = PH_COPYALL;  // m = min(len, k) elements (none when k == 0)
  } else if (len <= p.small_max) {
    s.phase = PH_SMALL;
  } else {
    // k beyond the LDS select (unsorted output only; the host rejects sorted
    // calls with such k): the ordered emit writes position order directly,
    // after a sampled bracket of the k-th key (slots for the keys inside
    // it) or, on the exact path, the digit passes (no slots).
    const bool bigk = k > p.lds_keys;
    s.phase = (p.flags & FL_EXACT) ? PH_HIST : PH_SAMPLE;
    if (s.phase == PH_HIST) s.flags = RF_NEED_INIT;
    tiles = (uint32_t)((len + p.tune.tile - 1) / p.tune.tile);
    const int64_t cap = !bigk ? 4 * k + (len + 63) / 64 + 1024
                              : ((p.flags & FL_EXACT) ? 0 : len / 16 + 4096);
    s.cap = (uint32_t)(cap < len ? cap : len);
    // A row whose slots hold all of it gains nothing from a threshold: every
    // key becomes a candidate (T = 0) and the select works on the whole row.
    if (s.phase == PH_SAMPLE && !bigk && cap >= len) {
      s.phase = PH_FILTER;
      s.T = 0;
    }
  }
  p.desc[r] = d;
  p.st[r] = s;
  p.tile_start[r] = tiles;
  p.cand_base[r] = s.cap;
  p.cand_cnt[r] = 0;
  return RowPlan{len, (int32_t)k, tiles, s.cap, s.phase, bad};
}

template <int DT>
__global__ __launch_bounds__(256) void k_plan_rows(KParams p) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r < p.R) plan_row<DT>(p, r);
}

// Exclusive scans of the per-row tile counts and candidate capacities (one
// workgroup; rows are processed in chunks of 1024 with a running carry).
// FUSED (at most 1024 rows): the same workgroup plans the rows first, which
// saves a launch on small batches.
// NT threads (1024, or one wave for calls of at most 64 rows, where the
// workgroup barriers of 16 waves would be most of its latency).
template <int DT, bool FUSED, int NT>
RK_DEV void plan_scan(const KParams& p) {
  // FUSED: thread t plans row t and keeps its plan in registers; otherwise
  // the plans are read back from what k_plan_rows wrote.
  RowPlan mine{0, 0, 0u, 0u, (uint32_t)PH_DONE, false};
  uint32_t my_t0 = 0;
  if (FUSED && (int)threadIdx.x < p.R) mine = plan_row<DT>(p, threadIdx.x);
  auto plan_of = [&](int r) -> RowPlan {
    if (FUSED) return mine;  // r == threadIdx.x in every loop below
    const RowState st = p.st[r];
    return RowPlan{p.desc[r].len, p.desc[r].k, p.tile_start[r], st.cap, st.phase,
                   (st.flags & 0x80000000u) != 0};
  };
  __shared__ uint32_t scratch[16];
  __shared__ uint64_t scratch64[16];
  __shared__ uint32_t bad_sh;
  __shared__ uint32_t hist_sh;
  __shared__ uint32_t status_sh;
  const int tid = threadIdx.x;
  if (tid == 0) {
    bad_sh = 0;
    hist_sh = 0;
  }
  __syncthreads();
  uint32_t carry_t = 0;
  uint64_t carry_c = 0;
  for (int base = 0; base < p.R; base += NT) {
    const int r = base + tid;
    RowPlan rp{0, 0, 0u, 0u, (uint32_t)PH_DONE, false};
    if (r < p.R) rp = plan_of(r);
    const uint32_t t = rp.tiles;
    const uint64_t c = rp.cap;
    if (rp.bad) atomicOr(&bad_sh, 1u);
    if (rp.phase == PH_HIST) atomicAdd(&hist_sh, 1u);
    uint32_t tot_t;
    const uint32_t ex_t = block_excl_scan<NT>(t, scratch, tot_t);
    // 64-bit scan of the capacities: wave scan + per-wave totals.
    const uint64_t inc = wave_incl_scan64(c);
    const int lane = lane_id(), wid = tid >> 6;
    if (lane == 63) scratch64[wid] = inc;
    __syncthreads();
    uint64_t pre = 0, tot_c = 0;
    for (int w = 0; w < NT / 64; ++w) {
      const uint64_t v = scratch64[w];
      pre += (w < wid) ? v : 0ull;
      tot_c += v;
    }
    __syncthreads();
    if (r < p.R) {
      p.tile_start[r] = carry_t + ex_t;
      p.cand_base[r] = carry_c + pre + inc - c;
    }
    if (FUSED) my_t0 = carry_t + ex_t;  // this row's first tile
    carry_t += tot_t;
    carry_c += tot_c;
  }
  if (tid == 0) {
    p.tile_start[p.R] = carry_t;
    p.cand_base[p.R] = carry_c;
    uint32_t status = bad_sh ? ST_BAD_ROW : 0u;
    Header* h = p.hdr;
    h->total_tiles = carry_t;
    h->total_cand = carry_c;
    h->fallback_rows = hist_sh;  // rows that start in the exact passes
    h->ordered_rows = 0;
    h->q_sample = 0;
    h->q_select[0] = 0;
    h->q_select[1] = 0;
    h->sel_done = 0;
    h->q_chunk[0] = 0;
    h->q_chunk[1] = 0;
    h->x_rows = 0;
    h->x_left = 0;
#ifdef RK_PHASE_TIMING
    for (int i = 30; i < 50; ++i) h->pad[i] = 0;
#endif
    if (carry_t > p.tile_capacity) status |= ST_TILES;
    if (carry_c > p.cand_capacity) status |= ST_CANDIDATES;
    h->status = status;
    status_sh = status;
    if (status & (ST_TILES | ST_CANDIDATES)) p.tile_start[p.R] = 0;  // disable streaming
  }
  __syncthreads();
  // Row orders of the per-row kernels, largest first, by counting sorts:
  // order[0, R) by floor(log2 k) for the select (the longest selections and
  // sorts start first), order[R, 2R) by floor(log2 len) for the sample,
  // whose cost grows with the row length.
  {
    __shared__ uint32_t bucket[2][32];
    if (tid < 64) bucket[tid >> 5][tid & 31] = 0;
    __syncthreads();
    auto classes = [&](int r, uint32_t& kc, uint32_t& lc) {
      const RowPlan rp = FUSED ? mine : RowPlan{p.desc[r].len, p.desc[r].k, 0u, 0u, 0u, false};
      const uint32_t k = (uint32_t)max(rp.k, 1);
      const uint64_t len = (uint64_t)max<int64_t>(rp.len, 1);
      kc = (uint32_t)min(31 - __builtin_clz(k), 15);
      lc = (uint32_t)min(63 - __builtin_clzll(len), 31);
    };
    for (int r = tid; r < p.R; r += NT) {
      uint32_t kc, lc;
      classes(r, kc, lc);
      atomicAdd(&bucket[0][kc], 1u);
      atomicAdd(&bucket[1][lc], 1u);
    }
    __syncthreads();
    if (tid < 2) {
      uint32_t acc = 0;
      for (int c = 31; c >= 0; --c) {
        const uint32_t n = bucket[tid][c];
        bucket[tid][c] = acc;
        acc += n;
      }
    }
    __syncthreads();
    for (int r = tid; r < p.R; r += NT) {
      uint32_t kc, lc;
      classes(r, kc, lc);
      p.order[atomicAdd(&bucket[0][kc], 1u)] = (uint32_t)r;
      p.order[p.R + atomicAdd(&bucket[1][lc], 1u)] = (uint32_t)r;
    }
  }
  // On a capacity overflow no long row is processed (the status reports it).
  if (status_sh & (ST_TILES | ST_CANDIDATES)) {
    for (int r = tid; r < p.R; r += NT) {
      const uint32_t ph = FUSED ? mine.phase : p.st[r].phase;
      if (ph == PH_SAMPLE || ph == PH_HIST) p.st[r].phase = PH_DONE;
      p.tile_start[r] = 0;
    }
    return;
  }
  // Chunk -> row table of the filter: cleared, then filled for the rows that
  // skip sampling (the sample kernel fills it for the sampled rows).
  const ChunkGrid g(carry_t, (uint32_t)p.chunk_tiles);  // carry_t: total tiles (uniform)
  for (uint32_t c = tid; c < g.nch; c += NT) p.chunk_row[c] = NO_ROW;
  __syncthreads();
  for (int r = tid; r < p.R; r += NT) {
    if ((FUSED ? mine.phase : p.st[r].phase) != PH_FILTER) continue;
    const uint32_t t0 = FUSED ? my_t0 : p.tile_start[r];
    const uint32_t t1 = FUSED ? my_t0 + mine.tiles : p.tile_start[r + 1];
    const uint32_t c0 = g.first_ge(t0), c1 = g.first_ge(t1);
    for (uint32_t c = c0; c < c1; ++c) p.chunk_row[c] = (uint32_t)r;
  }
}

template <int DT, bool FUSED, int NT = 1024>
__global__ __launch_bounds__(NT) void k_plan_scan(KParams p) {
  plan_scan<DT, FUSED, NT>(p);
}

// Sends row r to the exact passes from its first digit: histogram cleared,
// nothing resolved, all of its k still to take, and the row added to the
// exact-pass executor's list.  Called by a whole workgroup of NT threads of
// the kernel that makes the decision (after every thread has read the row's
// state), so the fallback round needs no initialisation launch of its own.
template <int NT>
RK_DEV void exact_start_row(const KParams& p, int r, uint32_t why, uint32_t phase = PH_HIST) {
  __syncthreads();  // the row's state is read before it changes
  uint32_t* h = p.ghist + (size_t)r * HIST_MAX_BINS;
  for (int i = threadIdx.x; i < HIST_MAX_BINS; i += NT) h[i] = 0;
  if (threadIdx.x == 0) {
    RowState& w = p.st[r];
    w.prefix = 0;
    w.resolved = 0;
    w.rem = (uint32_t)p.desc[r].k;
    w.flags = 0;
    w.why = why;
    w.phase = phase;
    p.cand_cnt[r] = 0;
    // chunks of x_chunk_tiles tiles (sized by the host from the call's work)
    const uint32_t tiles = p.tile_start[r + 1] - p.tile_start[r];
    uint32_t ct = (uint32_t)p.x_chunk_tiles;
    ct = ct > tiles ? tiles : ct;
    ct = ct < 1 ? 1 : ct;
    XRow& x = p.xst[r];
    x.claim = 0;
    x.done = 0;
    x.kor = 0u;
    x.kand = 0xffffffffu;
    x.ct = ct;
    x.nch = (tiles + ct - 1) / ct;
    const uint32_t slot = atomicAdd(&p.hdr->x_rows, 1u);
    if (slot < (uint32_t)p.R) {  // a row enters at most once per call
      p.xrows[slot] = (uint32_t)r;
      atomicAdd(&p.hdr->x_left, 1u);
    } else {
      atomicOr(&p.hdr->status, (uint32_t)ST_INTERNAL);
    }
  }
}

// ---------------------------------------------------------------------------
// sample: threshold T for a long row from sample_groups evenly spaced runs of
// group_elems contiguous elements.  With s sampled keys and mu = s*k/len, T is
// the r-th largest sampled key, r = ceil(mu + 4 sqrt(mu) + 2); #(row >= T) >=
// k then holds unless the sample is unrepresentative, which the select kernel
// detects exactly (it counts) and sends to the exact passes.
// ---------------------------------------------------------------------------
// PLAN (calls of one row, launched as one workgroup): the workgroup plans
// the call first (plan_scan), which saves the plan's launch; its barriers
// make the plan visible to the workgroup, and later launches see it anyway.
template <int DT, bool LARGEST, bool PLAN = false>
__global__ __launch_bounds__(SEL_NT) void k_sample(KParams p) {
  using U = typename Traits<DT>::U;
  extern __shared__ uint64_t dyn64[];
  uint32_t* keys = reinterpret_cast<uint32_t*>(dyn64);
  __shared__ uint32_t hist[HIST_MAX_BINS];
  __shared__ uint32_t scratch[16];
  __shared__ uint64_t scratch64[16];
  __shared__ uint32_t outv[4];
  __shared__ uint32_t next_sh;
  if constexpr (PLAN) {
    plan_scan<DT, true, SEL_NT>(p);
    __syncthreads();
  }
  for (;;) {
    if (threadIdx.x == 0) next_sh = atomicAdd(&p.hdr->q_sample, 1u);
    __syncthreads();
    const uint32_t qi = __builtin_amdgcn_readfirstlane(next_sh);  // wave-uniform
    __syncthreads();
    if (qi >= (uint32_t)p.R) break;
    const int r = __builtin_amdgcn_readfirstlane((int)p.order[p.R + qi]);
    const RowState st0 = p.st[r];
    if (st0.phase == PH_HIST && (st0.flags & RF_NEED_INIT)) {
      exact_start_row<SEL_NT>(p, r, st0.why);  // planned straight into the exact passes
      __syncthreads();
      continue;
    }
    if (st0.phase != PH_SAMPLE) continue;
    record_chunk_rows<SEL_NT>(p, r);
    RK_STAMP(p, r, 0);
    const RowDesc d = p.desc[r];
    const U* rb = static_cast<const U*>(d.base);
    const int64_t len = d.len;
    const int64_t ge = p.tune.group_elems;  // a multiple of 8, so of PACK
    int64_t ng = p.tune.sample_groups;
    const int64_t ng_cap = p.sample_keys / ge;      // runs that fit the LDS keys
    // read about 1/64 of the row (a bracketed row, k above the LDS select,
    // takes the full sample: its bracket narrows with the sample),
    if (d.k <= p.lds_keys && ng * ge * 64 > len) ng = len / (64 * ge);
    const int64_t ng_min = len / (2 * ge) < 64 ? len / (2 * ge) : 64;
    if (ng < ng_min) ng = ng_min;                   // but at least 64 runs,
    if (ng > ng_cap) ng = ng_cap;                   // and never beyond the LDS
    if (ng < 1) ng = 1;
    const uint32_t s = (uint32_t)(ng * ge);
    // Runs on the row's 16-byte grid, read as whole vectors: run g starts at
    // a0 + floor(g * (len - a0 - ge) / (ng - 1)) rounded down to a vector
    // (in double: no 64-bit integer division per run), so the last run ends
    // inside the row.  Sampled rows are far longer than a run.
    constexpr int P = Traits<DT>::PACK;
    const int64_t a0 = aligned_start<DT>(rb, 0);
    const double step = ng > 1 ? (double)(len - a0 - ge) / (double)(ng - 1) : 0.0;
    const uint32_t vpr = (uint32_t)(ge / P);  // vectors per run
    const uint32_t nvec = (uint32_t)ng * vpr;
    constexpr int UN = 8;  // 16-byte loads in flight per thread
    // the bits that vary among the sampled keys: OR ^ AND, gathered while
    // loading, so the radix select skips its own pass over them
    uint32_t kor = 0, kand = 0xffffffffu;
    for (uint32_t v0 = threadIdx.x; v0 < nvec; v0 += SEL_NT * UN) {
      uint4 x[UN];
#pragma unroll
      for (int q = 0; q < UN; ++q) {
        const uint32_t v = v0 + q * SEL_NT;
        const uint32_t g = v / vpr, j = v - g * vpr;
        const int64_t start = a0 + ((int64_t)((double)g * step) / P) * P;
        x[q] = v < nvec ? ldg16<false>(rb + start + (int64_t)j * P) : make_uint4(0u, 0u, 0u, 0u);
      }
#pragma unroll
      for (int q = 0; q < UN; ++q) {
        const uint32_t v = v0 + q * SEL_NT;
        if (v >= nvec) continue;
        const Vec<DT> e = unpack<DT>(x[q]);
#pragma unroll
        for (int c = 0; c < P; ++c) {
          const uint32_t kk = sel_key<DT, LARGEST>(e.e[c]);
          keys[v * P + c] = kk;
          kor |= kk;
          kand &= kk;
        }
      }
    }
    {
      const uint64_t both = block_or64<SEL_NT>(((uint64_t)(~kand) << 32) | kor, scratch64);
      kand = ~(uint32_t)(both >> 32);
      kor = (uint32_t)both;
    }  // its barriers also complete the keys
    RK_STAMP(p, r, 1);
    if (d.k > p.lds_keys) {
      // k above the LDS select: bracket the k-th key between the sampled
      // keys of ranks mu -+ (4 sigma + 4) (mu its expected rank, sigma the
      // binomial spread), so that it lies in [T_lo, T_hi] unless the sample
      // is unrepresentative, which the bracket pass detects exactly.
      const float mu = (float)s * (float)d.k / (float)len;
      const float dl = 4.0f * sqrtf(fmaxf(mu * (1.0f - mu / (float)s), 0.0f)) + 4.0f;
      const float fh = floorf(mu - dl), fl = ceilf(mu + dl);
      const uint32_t r_hi = fh < 1.0f ? 1u : (fh > (float)s ? s : (uint32_t)fh);
      const uint32_t r_lo = fl < 1.0f ? 1u : (fl > (float)s ? s : (uint32_t)fl);
      const uint32_t t_hi = block_radix_select_from<SEL_NT, 11, uint32_t, false>(
          keys, s, r_hi, kor ^ kand, kand, hist, scratch, outv);
      const uint32_t ge_hi = outv[3];
      __syncthreads();
      const uint32_t t_lo = block_radix_select_from<SEL_NT, 11, uint32_t, false>(
          keys, s, r_lo, kor ^ kand, kand, hist, scratch, outv);
      const uint32_t ge_lo = outv[3];
      // sampled keys in [T_lo, T_hi] (at least those of equal keys), scaled.
      // The bracket pays while its finisher's one-workgroup select stays
      // short: up to about 64K keys inside (measured on MI355X: 2^22-key
      // rows with k = N/4 and N/2, ~130K keys inside, were 2-5 % slower
      // than the digit passes; k = N/100 there, ~30K inside, 1.6x faster).
      const double est = (double)(ge_lo - ge_hi + 1u) * (double)len / (double)s;
      const bool fits = est * 1.25 <= (double)st0.cap && est <= 65536.0;
      exact_start_row<SEL_NT>(p, r, fits ? FB_NONE : FB_SAMPLE_TOO_MANY,
                              fits ? (uint32_t)PH_BRACKET : (uint32_t)PH_HIST);
      if (threadIdx.x == 0) {
        if (fits) {
          RowState& w = p.st[r];
          w.T = t_lo;
          w.V = t_hi;
        }
        atomicAdd(&p.hdr->fallback_rows, 1u);  // the fallback round finishes it
      }
      __syncthreads();
      continue;
    }
    const float mu = (float)s * (float)d.k / (float)len;
    float rf = ceilf(mu + 4.0f * sqrtf(mu) + 2.0f);
    uint32_t rank = rf >= (float)s ? s : (uint32_t)rf;
    if (rank < 1) rank = 1;
    uint32_t T = 0, ge_cnt = s;
    if (rank < s) {
      T = block_radix_select_from<SEL_NT, 11, uint32_t, false>(keys, s, rank, kor ^ kand, kand,
                                                              hist, scratch, outv);
      ge_cnt = outv[3];  // sampled keys >= T
    }
    RK_STAMP(p, r, 2);
    // Expected candidates: the sampled fraction at or above T, scaled.
    const double est = (double)ge_cnt * (double)len / (double)s;
    const uint32_t why = (T == 0 || rank >= s) ? FB_SAMPLE_DEGENERATE
                         : est * 1.25 > (double)st0.cap ? FB_SAMPLE_TOO_MANY : FB_NONE;
    if (why != FB_NONE) {  // uniform
      exact_start_row<SEL_NT>(p, r, why);
      if (threadIdx.x == 0) atomicAdd(&p.hdr->fallback_rows, 1u);
    } else if (threadIdx.x == 0) {
      RowState& st = p.st[r];
      st.T = T;
      st.phase = PH_FILTER;
      st.flags = 0;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// filter: the streaming pass.  Every element of every PH_FILTER row is read
// once with 16-byte loads; keys >= T are appended (as composite keys) to the
// row's candidate slots, staged in LDS and flushed with one global atomic
// per segment.  Each 16-byte vector is first tested as a whole in integer
// ord space (one compare per vector); only vectors that pass are tested per
// element on the selection key, so exactly the keys >= T are admitted.
// ---------------------------------------------------------------------------
// Candidate staging of the streaming filter.  Each wave owns a private LDS
// stage of WSTAGE composite keys; only that wave writes or reads it (LDS
// operations of one wave complete in order), so appends need no atomics and
// a full stage is flushed by the wave alone with one global atomic on the
// row's counter: no workgroup barrier anywhere in the streaming loop.
constexpr int WSTAGE = 512;
// Second level: a vector that the prefilter passes is staged whole (16 bytes
// and the row position of its first element) in a per-wave LDS stage of
// VSTAGE vectors; when it is full (and at the segment's end) each lane takes
// one staged vector, computes its exact element mask and appends the keys.
// The streaming loop then spends a few instructions per hit vector instead
// of the per-element masks and ballots with most lanes idle (hits are
// sparse: ~0.1-0.5 % of vectors on the batched workloads).
constexpr int VSTAGE = 64;

struct AppendCtx {
  uint64_t* stage;     // this wave's LDS stage (WSTAGE entries)
  uint64_t* cand;      // global candidate slots of the row
  uint32_t* cand_cnt;  // global candidate counter of the row
  uint32_t cap;        // candidate slots of the row
  bool coh = false;    // device-coherent stores (read back in the same launch)
};

// Moves the wave's staged keys to the row's candidate slots.  `full` is set
// (wave-uniform) once the row's candidates exceed its slots: the row then
// takes the exact passes, and the wave stops streaming it.
RK_DEV void wave_flush(const AppendCtx& ac, uint32_t& n, bool& full) {
  if (n == 0) return;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  uint32_t gb = 0;
  if (lane_id() == 0) gb = atomicAdd(ac.cand_cnt, n);
  gb = __shfl(gb, 0, 64);
  for (uint32_t i = lane_id(); i < n; i += 64) {
    if (gb + i >= ac.cap) continue;
    if (ac.coh) st_dev(ac.cand + gb + i, ac.stage[i]);
    else ac.cand[gb + i] = ac.stage[i];
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (gb + n > ac.cap) full = true;
  n = 0;
}

// Wave-wide append of the masked elements of one 16-byte vector per lane
// (called by the whole wave, under a wave-uniform branch).  Slots come from
// one ballot per element position: position c of lane l goes to
//   sum over c' < c of popcount(ballot_c') + popcount(ballot_c below lane l).
template <int DT, bool LARGEST>
RK_DEV void append_vec(const AppendCtx& ac, uint32_t& n, bool& full, uint4 x, uint32_t mask,
                       int64_t first) {
  constexpr int P = Traits<DT>::PACK;
  const Vec<DT> v = unpack<DT>(x);
  uint64_t bal[P];
  uint32_t tot = 0;
#pragma unroll
  for (int c = 0; c < P; ++c) {
    bal[c] = __ballot((mask >> c) & 1u);
    tot += (uint32_t)__popcll(bal[c]);
  }
  if (tot == 0) return;
  if (n + tot > (uint32_t)WSTAGE) wave_flush(ac, n, full);  // tot <= 64 * P <= WSTAGE
  uint32_t base = n;
#pragma unroll
  for (int c = 0; c < P; ++c) {
    if ((mask >> c) & 1u) {
      const uint32_t pos = (uint32_t)(first + c);
      ac.stage[base + mbcnt(bal[c])] =
          ((uint64_t)sel_key<DT, LARGEST>(v.e[c]) << 32) | (uint32_t)(~pos);
    }
    base += (uint32_t)__popcll(bal[c]);
  }
  n += tot;
}

// The vector grid of a row, shared by every streaming pass: 16-byte vectors
// counted from a0, the row's first element on a 128-byte boundary (an L2
// line), so a wave's 1 KB load covers 8 lines, not 9 (rows start anywhere;
// measured on C4: +2 % fp32, +3 % bf16).  A segment of tiles [ta, tb) owns
// vectors [ta * tile / P, tb * tile / P) of that grid; the row's first
// segment also owns the head elements [0, a0) (fewer than 128 bytes) and its
// last segment the tail elements after the last whole vector (fewer than P).
struct SegGrid {
  int64_t a0;     // the row's first grid element
  int64_t al;     // first element of the segment's vectors
  int64_t nv;     // vectors of the segment
  int64_t tail0;  // first element after the row's last whole vector
  bool first_seg, last_seg;
};
template <int DT>
RK_DEV SegGrid seg_grid(const void* base, int64_t len, int tile, uint32_t ta, uint32_t tb) {
  constexpr int P = Traits<DT>::PACK;
  constexpr int ESZ = (int)sizeof(typename Traits<DT>::U);
  SegGrid g;
  const int64_t a_line = (int64_t)(((128u - ((uint64_t)base & 127u)) & 127u) / ESZ);
  g.a0 = a_line < len ? a_line : len;
  const int64_t nv_row = (len - g.a0) / P;
  const int64_t v_lo = (int64_t)ta * (tile / P);
  int64_t v_hi = (int64_t)tb * (tile / P);
  if (v_hi > nv_row) v_hi = nv_row;
  g.al = g.a0 + (v_lo < v_hi ? v_lo : v_hi) * P;
  g.nv = v_hi > v_lo ? v_hi - v_lo : 0;
  g.tail0 = g.a0 + nv_row * P;
  g.first_seg = (ta == 0);
  g.last_seg = (int64_t)tb * tile >= len;
  return g;
}

// LDS of one streaming workgroup's candidate staging.
struct FilterLds {
  uint64_t stage[STREAM_NT / 64][WSTAGE];
  uint4 vstage[STREAM_NT / 64][VSTAGE];
  uint32_t vpos[STREAM_NT / 64][VSTAGE];
  uint32_t flush_n[STREAM_NT / 64], flush_base[STREAM_NT / 64];
};

// Tiles [ta, tb) of row r (a PH_FILTER row): every key >= T appended to the
// row's candidate slots.  Called by a whole workgroup of STREAM_NT threads.
template <int DT, bool LARGEST, int U>
RK_DEV void filter_segment(const KParams& p, FilterLds& lds, const RowState& st, const RowDesc& d,
                           uint64_t cbase, int r, uint32_t ta, uint32_t tb) {
  using UT = typename Traits<DT>::U;
  constexpr int P = Traits<DT>::PACK;
  auto& stage = lds.stage;
  auto& vstage = lds.vstage;
  auto& vpos = lds.vpos;
  auto& flush_n = lds.flush_n;
  auto& flush_base = lds.flush_base;
  const int lane = lane_id();
  // wave-uniform (readfirstlane tells the compiler), so the streaming loop's
  // control flow stays scalar and its loads stay unconditional
  const int wave_base = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63);
  {
    if (st.phase != PH_FILTER) return;  // uniform across the block
    const uint32_t T = st.T;
    const uint32_t cap = st.cap;

    const bool all = (T == 0);
    const uint32_t xm = prefilter_mask<DT, LARGEST>(T);
    const UT* rb = static_cast<const UT*>(d.base);
    const SegGrid g = seg_grid<DT>(d.base, d.len, p.tune.tile, ta, tb);
    const int64_t a0 = g.a0, al = g.al, nv = g.nv, tail0 = g.tail0;
    const bool first_seg = g.first_seg, last_seg = g.last_seg;

    auto pred = [&](uint32_t u) -> bool { return all || sel_key<DT, LARGEST>(u) >= T; };
    AppendCtx ac;
    ac.stage = stage[wave_base >> 6];
    ac.cand = p.cand + cbase;
    ac.cand_cnt = &p.cand_cnt[r];
    ac.cap = cap;
    uint32_t staged = 0;  // keys in this wave's stage (wave-uniform)
    bool full = false;    // the row's slots overflowed (wave-uniform)
    uint4* vst = vstage[wave_base >> 6];
    uint32_t* vps = vpos[wave_base >> 6];
    uint32_t vstaged = 0;  // vectors in this wave's vector stage (wave-uniform)
    // Exact masks of the staged vectors, one per lane, and their keys
    // appended (called by the whole wave).
    auto flush_vecs = [&]() {
      if (vstaged == 0) return;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      uint4 xv = make_uint4(0u, 0u, 0u, 0u);
      uint32_t first = 0, mm = 0;
      if ((uint32_t)lane < vstaged) {
        xv = vst[lane];
        first = vps[lane];
        const Vec<DT> v = unpack<DT>(xv);
#pragma unroll
        for (int c = 0; c < P; ++c) mm |= (sel_key<DT, LARGEST>(v.e[c]) >= T ? 1u : 0u) << c;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (__ballot(mm != 0)) append_vec<DT, LARGEST>(ac, staged, full, xv, mm, (int64_t)first);
      vstaged = 0;
    };

    // The row's head and tail elements outside the vector grid, one per
    // lane: wave 0 of the first segment and the last wave of the last.
    if ((wave_base == 0 && first_seg) || (wave_base == STREAM_NT - 64 && last_seg)) {
      int64_t pos = -1;
      if (wave_base == 0 && lane < a0) pos = lane;
      if (wave_base != 0 && lane < P && tail0 + lane < d.len) pos = tail0 + lane;
      uint32_t mask = 0;
      uint4 x = make_uint4(0u, 0u, 0u, 0u);
      if (pos >= 0) {
        const uint32_t u0 = (uint32_t)rb[pos];
        x.x = u0;  // element 0 of the vector (16-bit types: low half)
        mask = pred(u0) ? 1u : 0u;
      }
      if (__ballot(mask != 0)) append_vec<DT, LARGEST>(ac, staged, full, x, mask, pos);
    }
    // Aligned body.
    const uint4* vp = reinterpret_cast<const uint4*>(rb + al);
    if (all) {
      // Threshold admits every key (kept simple: such rows are rare).
      for (int64_t jb = wave_base; jb < nv && !full; jb += STREAM_NT) {
        const int64_t j = jb + lane;
        uint4 x = make_uint4(0u, 0u, 0u, 0u);
        uint32_t m = 0;
        if (j < nv) {
          x = ldg16<false>(vp + j);
          m = (1u << P) - 1u;
        }
        append_vec<DT, LARGEST>(ac, staged, full, x, m, al + j * P);
      }
    } else {
      // Buffer loads, U x 16 bytes per lane per step.  Ranges are cut into
      // chunks of at most 2^30 bytes for the 32-bit offsets.
      constexpr uint32_t STEP = (uint32_t)STREAM_NT * U;
      constexpr int64_t CHUNK = int64_t(1) << 26;  // vectors per chunk
      for (int64_t c0 = 0; c0 < nv; c0 += CHUNK) {
        const uint32_t nvc = (uint32_t)((nv - c0) < CHUNK ? (nv - c0) : CHUNK);
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(vp + c0, nvc * 16u);
        auto load_set = [&](uint4 (&x)[U], uint32_t jb) {
#pragma unroll
          for (int u = 0; u < U; ++u) x[u] = bload16_nt(rs, (jb + lane + u * STREAM_NT) * 16u);
        };
        auto test_set = [&](const uint4 (&x)[U], uint32_t jb) {
          uint32_t hb = 0;  // bit u: this lane's vector u may hold a candidate
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint32_t j = jb + lane + u * STREAM_NT;
            hb |= ((vec_max_xm<DT>(x[u], xm) >= T && j < nvc) ? 1u : 0u) << u;
          }
          if (__ballot(hb != 0)) {  // rare: stage the vectors that passed
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const bool hit = (hb >> u) & 1u;
              const uint64_t bm = __ballot(hit);
              if (!bm) continue;
              const uint32_t nb = (uint32_t)__popcll(bm);
              if (vstaged + nb > (uint32_t)VSTAGE) flush_vecs();
              if (hit) {
                const uint32_t sl = vstaged + mbcnt(bm);
                vst[sl] = x[u];
                vps[sl] = (uint32_t)(al + (int64_t)(c0 + jb + lane + u * STREAM_NT) * P);
              }
              vstaged += nb;
            }
          }
        };
        // Software pipelined: the next step's U loads are issued before this
        // step's keys are tested, so a wave always has loads in flight and
        // 4 workgroups per CU reach the bandwidth (measured on MI355X: 6.27
        // -> 6.60 TB/s fp32, 5.83 -> 6.14 bf16; a third stage added nothing).
        // Loads past the range return zeros (masked above, or never tested),
        // so they are issued unconditionally.
        uint4 x[U], y[U];
        uint32_t jb = (uint32_t)wave_base;
        if (jb < nvc) load_set(x, jb);
        for (; jb < nvc; jb += STEP) {  // wave-uniform
          load_set(y, jb + STEP);
          test_set(x, jb);
#pragma unroll
          for (int u = 0; u < U; ++u) x[u] = y[u];
          if (full) break;
        }
        if (full) break;
      }
    }
    flush_vecs();
    // End of the segment: one atomic per workgroup reserves the slots of
    // all four waves' staged keys (per-wave atomics on the row's counter,
    // from every workgroup reading that row, serialized at the counter).
    if (lane == 0) flush_n[wave_base >> 6] = staged;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (int w = 0; w < STREAM_NT / 64; ++w) tot += flush_n[w];
      uint32_t b = tot ? atomicAdd(ac.cand_cnt, tot) : 0u;
      for (int w = 0; w < STREAM_NT / 64; ++w) {
        flush_base[w] = b;
        b += flush_n[w];
      }
    }
    __syncthreads();
    const uint32_t fb = flush_base[wave_base >> 6];
    for (uint32_t i = lane; i < staged; i += 64)
      if (fb + i < ac.cap) ac.cand[fb + i] = ac.stage[i];
    staged = 0;
  }
}

// The sampled round's streaming pass over every PH_FILTER row.  At most 72
// VGPRs (7 waves per SIMD) for 2 loads in flight and 128 (4 waves) for 4:
// the grids the autotuner offers keep at most 4 workgroups (4 waves per
// SIMD) per CU, and at 80 or 96 registers the 4-load form spills to scratch
// memory (gfx950: 16-84 bytes per lane).
template <int DT, bool LARGEST, int U>
__global__ __launch_bounds__(STREAM_NT)
__attribute__((amdgpu_waves_per_eu(U >= 8 ? 1 : (U == 4 ? 4 : 7))))
void k_filter(KParams p) {
  __shared__ FilterLds lds;
#ifdef RK_PHASE_TIMING
  // per-workgroup start, end and segment count (tuning builds), in the
  // histogram slots of row blockIdx.x when the call has that many rows
  uint32_t nseg = 0;
  const uint64_t t_start = wall_clock64();
#endif
  for_each_chunk_queued(p, &p.hdr->q_chunk[0], [&](int r, uint32_t ta, uint32_t tb) {
    // the row's metadata, loaded together (one round trip)
    const RowState st = p.st[r];
    const RowDesc d = p.desc[r];
    const uint64_t cbase = p.cand_base[r];
    filter_segment<DT, LARGEST, U>(p, lds, st, d, cbase, r, ta, tb);
#ifdef RK_PHASE_TIMING
    ++nseg;
#endif
  });
#ifdef RK_PHASE_TIMING
  if (threadIdx.x == 0 && (int)blockIdx.x < p.R) {
    uint32_t* w = p.ghist + (size_t)blockIdx.x * HIST_MAX_BINS + 1024;
    w[0] = (uint32_t)t_start;
    w[1] = (uint32_t)wall_clock64();
    w[2] = nseg;
  }
#endif
}

// ---------------------------------------------------------------------------
// Output ordering.  The k selected composite keys are re-keyed so that one
// descending bitonic sort yields the requested order:
//   position order (unsorted):  (~pos << 32) | key
//   best first:                 (key << 32) | ~pos
//   worst first:                (~key << 32) | ~pos    (ties: position ascending)
// All re-keyed values are non-zero, so 0 pads the sort to a power of two.
// ---------------------------------------------------------------------------
enum OrderMode : int { ORD_POS = 0, ORD_BEST = 1, ORD_WORST = 2 };

RK_DEV int order_mode(uint32_t flags) {
  if (!(flags & FL_SORTED)) return ORD_POS;
  const bool largest = flags & FL_LARGEST, ascend = flags & FL_ASCEND;
  return (largest != ascend) ? ORD_BEST : ORD_WORST;
}

template <int DT>
RK_DEV uint64_t order_key(int mode, uint32_t sk, uint32_t pos) {
  constexpr uint32_t KM = Traits<DT>::KEY_MASK;
  if (mode == ORD_POS) return ((uint64_t)(~pos) << 32) | sk;
  if (mode == ORD_BEST) return ((uint64_t)sk << 32) | (uint32_t)(~pos);
  return ((uint64_t)(~sk & KM) << 32) | (uint32_t)(~pos);
}

template <int DT>
RK_DEV void order_decode(int mode, uint64_t key, uint32_t& sk, uint32_t& pos) {
  constexpr uint32_t KM = Traits<DT>::KEY_MASK;
  if (mode == ORD_POS) {
    pos = ~(uint32_t)(key >> 32);
    sk = (uint32_t)key;
  } else if (mode == ORD_BEST) {
    sk = (uint32_t)(key >> 32);
    pos = ~(uint32_t)key;
  } else {
    sk = ~(uint32_t)(key >> 32) & KM;
    pos = ~(uint32_t)key;
  }
}

// Slot of this lane among the lanes with `take`, from one LDS atomic per
// wave on *cnt (called by whole waves; the result is defined where take).
RK_DEV uint32_t wave_slot(uint32_t* cnt, bool take) {
  const uint64_t m = __ballot(take);
  if (m == 0) return 0;
  const int fl = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane_id() == fl) base = atomicAdd(cnt, (uint32_t)__popcll(m));
  return (uint32_t)__builtin_amdgcn_readlane((int)base, fl) + mbcnt(m);
}

// LDS a sort borrows from its kernel: 2048 bucket counters and reduction
// scratch.
struct SortLds {
  uint32_t* bins;
  uint64_t* s64;  // >= 3 * SEL_NT / 64 words
  uint32_t* s32;  // >= SEL_NT / 64 words
  uint32_t* ctab; // 256 words (the bucket sort's coarse bins)
  uint32_t cap;   // keys the sort buffer holds
};

// Sorts the n (64 < n <= BUCKET_SORT_MAX) distinct keys sbuf[0, n)
// descending by buckets and ranks, with sbuf[BUCKET_SORT_MAX, 2
// BUCKET_SORT_MAX) as the grouping buffer (the select's LDS holds at least
// 8192 keys); false (keys untouched) when they do not spread over the
// buckets.
RK_DEV bool bucket_sort_all(uint64_t* sbuf, uint32_t n, const SortLds& sl) {
  if (n > BUCKET_SORT_MAX || sl.cap < 2 * BUCKET_SORT_MAX) return false;
  return bucket_sort_desc<SEL_NT>(sbuf, sbuf + BUCKET_SORT_MAX, n, sl.bins, sl.ctab, sl.s64,
                                  sl.s32);
}

// Sorts the distinct keys sbuf[0, m) descending: by buckets and ranks when
// they spread (bucket_sort_desc), else by a bitonic network (zeros pad it to
// a power of two).
RK_DEV void sort_desc(uint64_t* sbuf, uint32_t m, const SortLds& sl) {
  __syncthreads();  // sbuf complete
  if (m > 64 && m <= BUCKET_SORT_MAX && bucket_sort_all(sbuf, m, sl)) return;
  const uint32_t P2 = next_pow2(m);
  for (uint32_t i = m + threadIdx.x; i < P2; i += SEL_NT) sbuf[i] = 0;
  __syncthreads();
  if (P2 > 1) block_bitonic_desc<SEL_NT>(sbuf, P2);
}

// Writes the row's output from sbuf[0, m), keys already in output order
// (re-keyed), plus its tail fill.
template <int DT, bool LARGEST>
RK_DEV void write_row(const KParams& p, const RowDesc& d, const uint64_t* sbuf, uint32_t m,
                      int mode) {
  using U = typename Traits<DT>::U;
  const U* rb = static_cast<const U*>(d.base);
  U* ov = static_cast<U*>(d.out_v);
  for (uint32_t i = threadIdx.x; i < m; i += SEL_NT) {
    uint32_t sk, pos;
    order_decode<DT>(mode, sbuf[i], sk, pos);
    const uint32_t bits = sk ? key_to_bits<DT, LARGEST>(sk) : (uint32_t)ldg(rb + pos);
    ov[i] = (U)bits;
    put_idx(p, d, i, caller_idx(p, d, pos));
  }
  for (int64_t i = (int64_t)m + threadIdx.x; i < d.slots; i += SEL_NT) {
    ov[i] = (U)0;
    put_idx(p, d, i, -1);
  }
}

// Sort sbuf[0, m) (re-keyed) and write the row's output plus its tail fill.
template <int DT, bool LARGEST>
RK_DEV void sort_and_write(const KParams& p, const RowDesc& d, uint64_t* sbuf, uint32_t m, int mode,
                           const SortLds& sl, int r = 0) {
  sort_desc(sbuf, m, sl);
  RK_STAMP(p, r, 13);
  (void)r;
  write_row<DT, LARGEST>(p, d, sbuf, m, mode);
}

// Rows of the ordered emit, after the executor wrote them (the last step of
// the fallback round's select): position order needs only the caller's
// indices and the tail fill (any k, no LDS); the other orders sort the k
// keys the emit wrote.  Rows are taken by workgroup index.
template <int DT, bool LARGEST>
RK_DEV void finalize_ordered_rows(const KParams& p, uint64_t* sbuf, const SortLds& sl, int mode) {
  using U = typename Traits<DT>::U;
  for (int r = blockIdx.x; r < p.R; r += gridDim.x) {
    if (p.st[r].phase != PH_ORDERED) continue;
    const RowDesc d = p.desc[r];
    const uint32_t m = (uint32_t)d.k;
    if (mode == ORD_POS) {
      U* ow = static_cast<U*>(d.out_v);
      if (d.idx_in)
        for (uint32_t i = threadIdx.x; i < m; i += SEL_NT)
          put_idx(p, d, i, caller_idx(p, d, (uint32_t)get_idx(p, d, i)));
      for (int64_t i = (int64_t)m + threadIdx.x; i < d.slots; i += SEL_NT) {
        ow[i] = (U)0;
        put_idx(p, d, i, -1);
      }
    } else {
      const U* ov = static_cast<const U*>(d.out_v);
      for (uint32_t i = threadIdx.x; i < m; i += SEL_NT)
        sbuf[i] = order_key<DT>(mode, sel_key<DT, LARGEST>((uint32_t)ov[i]),
                                (uint32_t)get_idx(p, d, i));
      __syncthreads();
      sort_and_write<DT, LARGEST>(p, d, sbuf, m, mode, sl);
    }
    __syncthreads();
    if (threadIdx.x == 0) p.st[r].phase = PH_DONE;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// select_finalize: per row, one 1024-thread workgroup with a 16384-key LDS
// buffer.  COPYALL rows are sorted whole; SMALL rows are loaded straight from
// the input; FILTER rows read their candidates (LDS when they fit, global
// otherwise), verify that at least k of them are >= T, radix-select the k
// largest composite keys, sort and write.  A sampled row that fails the
// verification goes to the exact passes.
// ---------------------------------------------------------------------------
// Two 512-thread workgroups per CU (4 waves per SIMD, up to 128 VGPRs): the
// select's loops hold 8 loads in flight and the 4096-key bucket sort 8 keys
// per thread in registers, which a 96- or 80-register budget spills to
// scratch memory (gfx950, 96: 72-156 bytes per lane).
// FIN: the instance launched after the exact-pass executor, which also
// finalizes the rows it emitted in order (a separate instance, so that the
// sampled round's select keeps its own, smaller code).
template <int DT, bool LARGEST, bool FIN = false>
__global__ __launch_bounds__(SEL_NT) __attribute__((amdgpu_waves_per_eu(4)))
void k_select_finalize(KParams p) {
  if (p.round && p.hdr->fallback_rows == 0) return;
  using U = typename Traits<DT>::U;
  extern __shared__ uint64_t sbuf[];
  __shared__ uint32_t hist[HIST_MAX_BINS];
  __shared__ uint32_t scratch[16];
  __shared__ uint64_t scratch64[3 * SEL_NT / 64];
  __shared__ uint32_t outv[4];
  __shared__ uint32_t slot_cnt;
  __shared__ uint32_t next_sh;
  __shared__ uint32_t ctab[256];
  const SortLds sl{hist, scratch64, scratch, ctab, (uint32_t)p.lds_keys};
  const int mode = order_mode(p.flags);
  for (;;) {
    if (threadIdx.x == 0) next_sh = atomicAdd(&p.hdr->q_select[p.round], 1u);
    __syncthreads();
    // wave-uniform (readfirstlane tells the compiler): the row's metadata
    // then comes in scalar loads and stays in scalar registers
    const uint32_t qi = __builtin_amdgcn_readfirstlane(next_sh);
    __syncthreads();
    if (qi >= (uint32_t)p.R) break;
    const int r = __builtin_amdgcn_readfirstlane((int)p.order[qi]);
    const uint32_t ph = p.st[r].phase;
    if (ph != PH_COPYALL && ph != PH_SMALL && ph != PH_FILTER) continue;
    uint32_t next_phase = PH_DONE;
    RK_STAMP(p, r, 8);
    const RowDesc d = p.desc[r];
    const U* rb = static_cast<const U*>(d.base);
    const int64_t len = d.len;
    const uint32_t k = (uint32_t)d.k;
    if (ph == PH_COPYALL && mode == ORD_POS) {
      // position order: the row itself, then the tail fill (any length)
      const uint32_t m = (uint32_t)(len < (int64_t)k ? len : (int64_t)k);
      U* ov = static_cast<U*>(d.out_v);
      for (uint32_t i = threadIdx.x; i < m; i += SEL_NT) {
        ov[i] = ldg(rb + i);
        put_idx(p, d, i, caller_idx(p, d, i));
      }
      for (int64_t i = (int64_t)m + threadIdx.x; i < d.slots; i += SEL_NT) {
        ov[i] = (U)0;
        put_idx(p, d, i, -1);
      }
    } else if (ph == PH_COPYALL) {
      const uint32_t m = (uint32_t)(len < (int64_t)k ? len : (int64_t)k);
      load_to_lds<SEL_NT, 4>(sbuf, rb, m, [&](U x, uint32_t i) {
        return order_key<DT>(mode, sel_key<DT, LARGEST>((uint32_t)x), i);
      });
      __syncthreads();
      sort_and_write<DT, LARGEST>(p, d, sbuf, m, mode, sl);
    } else {
      const uint64_t* keys;
      uint32_t n;
      bool in_lds;
      if (ph == PH_SMALL) {
        n = (uint32_t)len;
        load_to_lds<SEL_NT, 4>(sbuf, rb, n, [](U x, uint32_t i) {
          return ((uint64_t)sel_key<DT, LARGEST>((uint32_t)x) << 32) | (uint32_t)(~i);
        });
        keys = sbuf;
        in_lds = true;
      } else {
        const RowState st = p.st[r];
        n = p.cand_cnt[r];
        const uint64_t* src = p.cand + p.cand_base[r];
        bool ok = n <= st.cap && n >= k;
        uint32_t why = n > st.cap ? FB_OVERFLOW : FB_TOO_FEW;
        if (ok) {
          in_lds = n <= (uint32_t)p.lds_keys;
          uint32_t ge = 0;
          if (in_lds) {
            load_to_lds<SEL_NT, 4>(sbuf, src, n, [](uint64_t x, uint32_t) { return x; });
            keys = sbuf;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n; i += SEL_NT)
              ge += ((uint32_t)(ld_lds(keys + i) >> 32) >= st.T) ? 1u : 0u;
          } else {
            keys = src;
            constexpr int UN = 8;  // loads in flight per thread
            for (uint32_t i0 = threadIdx.x; i0 < n; i0 += SEL_NT * UN) {
              uint64_t x[UN];
#pragma unroll
              for (int u = 0; u < UN; ++u) {
                const uint32_t i = i0 + (uint32_t)u * SEL_NT;
                x[u] = i < n ? ldg(keys + i) : 0ull;
              }
#pragma unroll
              for (int u = 0; u < UN; ++u) {
                const uint32_t i = i0 + (uint32_t)u * SEL_NT;
                ge += (i < n && (uint32_t)(x[u] >> 32) >= st.T) ? 1u : 0u;
              }
            }
          }
          ge = block_sum<SEL_NT>(ge, scratch);
          ok = ge >= k;
        }
        if (!ok) {  // uniform
          exact_start_row<SEL_NT>(p, r, why);
          if (threadIdx.x == 0) {
            if (st.flags & RF_EXACT_T) atomicOr(&p.hdr->status, (uint32_t)ST_INTERNAL);
            atomicAdd(&p.hdr->fallback_rows, 1u);
          }
          __syncthreads();
          continue;
        }
      }
      __syncthreads();
      RK_STAMP(p, r, 9);
      // Candidates that one bucket sort takes (all of them >= T, at least k)
      // are sorted whole: the first k are the selection, best first, so the
      // radix select, the compaction and a second sort are skipped.
      if (in_lds && n > 64 && n <= BUCKET_SORT_MAX && bucket_sort_all(sbuf, n, sl)) {
        RK_STAMP(p, r, 10);
        RK_STAMP(p, r, 11);
        if (mode == ORD_BEST) {
          RK_STAMP(p, r, 13);
          write_row<DT, LARGEST>(p, d, sbuf, k, mode);
        } else {
          for (uint32_t i = threadIdx.x; i < k; i += SEL_NT) {
            const uint64_t x = sbuf[i];
            sbuf[i] = order_key<DT>(mode, (uint32_t)(x >> 32), ~(uint32_t)x);
          }
          sort_and_write<DT, LARGEST>(p, d, sbuf, k, mode, sl, r);
        }
        RK_STAMP(p, r, 12);
        __syncthreads();
        if (threadIdx.x == 0) p.st[r].phase = PH_DONE;
        __syncthreads();
        continue;
      }
      const uint64_t kth =
          in_lds ? block_radix_select<SEL_NT, 11, uint64_t, false>(keys, n, k, hist, scratch, scratch64, outv)
                 : block_radix_select<SEL_NT, 11, uint64_t, true>(keys, n, k, hist, scratch, scratch64, outv);
      RK_STAMP(p, r, 10);
      // A large selection (k > MERGE_CHUNK) of a sampled row, when the merge
      // kernels follow this launch: its k keys are compacted into the row's
      // candidate slots (cap >= 4k) and sorted there by several workgroups.
      const bool to_merge = p.merge && ph == PH_FILTER && k > (uint32_t)MERGE_CHUNK;
      uint64_t* mk = p.cand + p.cand_base[r];
      if (threadIdx.x == 0) slot_cnt = 0;
      if (in_lds) {
        // In place, one chunk of CH keys per thread at a time: every slot
        // handed out is below the number of keys read so far, so a write
        // never lands on a key that is still unread.
        constexpr int CH = 4;
        for (uint32_t c0 = 0; c0 < n; c0 += SEL_NT * CH) {
          uint64_t reg[CH];
#pragma unroll
          for (int j = 0; j < CH; ++j) {
            const uint32_t i = c0 + threadIdx.x + j * SEL_NT;
            reg[j] = i < n ? sbuf[i] : 0ull;
          }
          __syncthreads();
#pragma unroll
          for (int j = 0; j < CH; ++j) {
            const uint32_t i = c0 + threadIdx.x + j * SEL_NT;
            const bool take = i < n && reg[j] >= kth;
            const uint32_t slot = wave_slot(&slot_cnt, take);
            if (take && slot < k)
              sbuf[slot] = order_key<DT>(mode, (uint32_t)(reg[j] >> 32), ~(uint32_t)reg[j]);
          }
        }
      } else {
        __syncthreads();
        constexpr int UN = 8;  // loads in flight per thread
        for (uint32_t c0 = 0; c0 < n; c0 += SEL_NT * UN) {  // uniform trip count
          uint64_t x[UN];
#pragma unroll
          for (int u = 0; u < UN; ++u) {
            const uint32_t i = c0 + threadIdx.x + (uint32_t)u * SEL_NT;
            x[u] = i < n ? ldg(keys + i) : 0ull;
          }
          // In place (keys == mk): every slot handed out is below the number
          // of keys read so far; the barrier makes all of this step's reads
          // precede its writes.
          if (to_merge) __syncthreads();
#pragma unroll
          for (int u = 0; u < UN; ++u) {
            const uint32_t i = c0 + threadIdx.x + (uint32_t)u * SEL_NT;
            const bool take = i < n && x[u] >= kth;
            const uint32_t slot = wave_slot(&slot_cnt, take);
            if (take && slot < k) {
              const uint64_t o = order_key<DT>(mode, (uint32_t)(x[u] >> 32), ~(uint32_t)x[u]);
              if (to_merge) mk[slot] = o;
              else sbuf[slot] = o;
            }
          }
          if (to_merge) __syncthreads();
        }
      }
      __syncthreads();
      if (threadIdx.x == 0 && slot_cnt != k) atomicOr(&p.hdr->status, (uint32_t)ST_INTERNAL);
      RK_STAMP(p, r, 11);
      if (to_merge) {
        if (in_lds)
          for (uint32_t i = threadIdx.x; i < k; i += SEL_NT) mk[i] = sbuf[i];
        next_phase = PH_MERGE;
      } else {
        sort_and_write<DT, LARGEST>(p, d, sbuf, k, mode, sl, r);
      }
      RK_STAMP(p, r, 12);
    }
    __syncthreads();
    if (threadIdx.x == 0) p.st[r].phase = next_phase;
    __syncthreads();
  }
  // the rows the executor emitted in order (none before it has run)
  if (FIN && p.hdr->ordered_rows != 0) finalize_ordered_rows<DT, LARGEST>(p, sbuf, sl, mode);
  if (p.publish && threadIdx.x == 0) {
    // the last workgroup to finish copies the header (every other
    // workgroup's header updates precede its increment of sel_done)
    __threadfence();
    if (atomicAdd(&p.hdr->sel_done, 1u) == gridDim.x - 1) {
      __threadfence();
      const uint32_t* h = reinterpret_cast<const uint32_t*>(p.hdr);
      for (int i = 0; i < 8; ++i)
        p.publish[i] = __hip_atomic_load(h + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence_system();
    }
  }
}

// ---------------------------------------------------------------------------
// Large selections, k > MERGE_CHUNK (launched only when max_k exceeds it).
// The select leaves such a row's k order-keyed composite keys in its
// candidate slots (PH_MERGE, a final phase no later kernel acts on).
// k_merge_sort sorts every chunk of MERGE_CHUNK keys in LDS, one workgroup
// per chunk, the chunks of all rows at once; k_merge_rank gives each key its
// output position: its index in its own sorted chunk plus, for every other
// chunk of the row, the number of that chunk's keys ordered before it (a
// binary search in LDS; the keys are unique), and writes value and index
// there.  A 16384-key selection is then four 4096-key sorts and four rank
// passes on four CUs instead of one 16384-key sort on one.  (The reference
// sorts K > 4096 with a host loop of per-row Thrust sorts,
// topk_radixselect.h:285-307.)
// ---------------------------------------------------------------------------
// Chunks per row of the merge kernels: rows with k above the LDS select
// (lds_keys) never reach them (they take the exact passes and ordered emit).
RK_HOST_DEV int merge_chunks(const KParams& p) {
  const int k = p.max_k < p.lds_keys ? p.max_k : p.lds_keys;
  return (k + MERGE_CHUNK - 1) / MERGE_CHUNK;
}

// Threads of the chunk sort: 4 keys per thread, 4 waves per SIMD (measured
// on MI355X for 4096 keys: 31 us at 1024 threads, 33 at 512, 53 at 256; the
// bitonic network's 64-bit compare-exchanges are VALU-bound and one wave per
// SIMD leaves their latency exposed).
constexpr int MS_NT = 1024;
constexpr int MS_E = MERGE_CHUNK / MS_NT;   // keys per thread (registers)

template <int DT>
__global__ __launch_bounds__(MS_NT) void k_merge_sort(KParams p) {
  __shared__ uint64_t xs[MERGE_CHUNK];
  __shared__ uint64_t xt[MERGE_CHUNK];  // the bucket sort's grouping buffer
  __shared__ uint32_t bins[2048];
  __shared__ uint32_t ctab[256];
  __shared__ uint64_t s64[3 * MS_NT / 64];
  __shared__ uint32_t s32[MS_NT / 64];
  const int cpr = merge_chunks(p);
  const int64_t items = (int64_t)p.R * cpr;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int r = (int)(it / cpr);
    const uint32_t c0 = (uint32_t)(it % cpr) * MERGE_CHUNK;
    if (p.st[r].phase != PH_MERGE) continue;  // uniform
    const uint32_t k = (uint32_t)p.desc[r].k;
    if (c0 >= k) continue;
    const uint32_t n = min(k - c0, (uint32_t)MERGE_CHUNK);
    uint64_t* keys = p.cand + p.cand_base[r] + c0;
    [[maybe_unused]] const int sb = 16 + 4 * (int)(it % cpr);  // phase-timing slots
    RK_STAMP(p, r, sb);
    load_to_lds<MS_NT, MS_E>(xs, keys, n, [](uint64_t x, uint32_t) { return x; });
    __syncthreads();
    RK_STAMP(p, r, sb + 1);
    // Buckets and ranks when the keys spread over their top varying bits,
    // else the bitonic network in registers (zeros pad it: they sort last).
    if (n > 64 && bucket_sort_desc<MS_NT>(xs, xt, n, bins, ctab, s64, s32)) {
      for (uint32_t i = threadIdx.x; i < n; i += MS_NT) keys[i] = xs[i];
    } else {
      uint64_t v[MS_E];
#pragma unroll
      for (int e = 0; e < MS_E; ++e) {
        const uint32_t i = threadIdx.x * MS_E + e;
        v[e] = i < n ? xs[i] : 0ull;
      }
      __syncthreads();  // xs is the network's exchange buffer from here on
      bitonic_sort_regs<MS_NT, MS_E>(v, xs);
#pragma unroll
      for (int e = 0; e < MS_E; ++e) {
        const uint32_t i = threadIdx.x * MS_E + e;
        if (i < n) keys[i] = v[e];
      }
    }
    RK_STAMP(p, r, sb + 2);
    __syncthreads();
    RK_STAMP(p, r, sb + 3);
  }
}

template <int DT, bool LARGEST>
__global__ __launch_bounds__(SEL_NT) void k_merge_rank(KParams p) {
  using U = typename Traits<DT>::U;
  constexpr int PER = MERGE_CHUNK / SEL_NT;  // own keys per thread
  extern __shared__ uint64_t sbuf[];         // all k keys of the row
  const int mode = order_mode(p.flags);
  const int cpr = merge_chunks(p);
  const int64_t items = (int64_t)p.R * cpr;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int r = (int)(it / cpr);
    const int c = (int)(it % cpr);
    const uint32_t c0 = (uint32_t)c * MERGE_CHUNK;
    if (p.st[r].phase != PH_MERGE) continue;  // uniform
    const RowDesc d = p.desc[r];
    const uint32_t k = (uint32_t)d.k;
    if (c0 >= k) continue;
    const uint32_t n = min(k - c0, (uint32_t)MERGE_CHUNK);
    [[maybe_unused]] const int sb = 32 + 4 * c;  // phase-timing slots
    RK_STAMP(p, r, sb);
    load_to_lds<SEL_NT, 8>(sbuf, p.cand + p.cand_base[r], k, [](uint64_t x, uint32_t) { return x; });
    __syncthreads();
    RK_STAMP(p, r, sb + 1);
    uint64_t x[PER];
    uint32_t rank[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t i = threadIdx.x + (uint32_t)j * SEL_NT;
      x[j] = i < n ? sbuf[c0 + i] : 0ull;
      rank[j] = i;
    }
    for (int cc = 0; cc < cpr; ++cc) {
      const uint32_t d0 = (uint32_t)cc * MERGE_CHUNK;
      if (cc == c || d0 >= k) continue;  // uniform
      const uint32_t m = min(k - d0, (uint32_t)MERGE_CHUNK);
      const uint64_t* other = sbuf + d0;
      // keys of `other` before x: a prefix (sorted descending), found by a
      // branch-free binary search, the PER searches of a thread interleaved
      uint32_t cnt[PER];
#pragma unroll
      for (int j = 0; j < PER; ++j) cnt[j] = 0;
      for (uint32_t step = next_pow2(m + 1) >> 1; step > 0; step >>= 1) {
        uint64_t o[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) o[j] = other[min(cnt[j] + step, m) - 1];  // unconditional
#pragma unroll
        for (int j = 0; j < PER; ++j) {
          const uint32_t q = cnt[j] + step;
          if (q <= m && o[j] > x[j]) cnt[j] = q;
        }
      }
#pragma unroll
      for (int j = 0; j < PER; ++j) rank[j] += cnt[j];
    }
    RK_STAMP(p, r, sb + 2);
    const U* rb = static_cast<const U*>(d.base);
    U* ov = static_cast<U*>(d.out_v);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t i = threadIdx.x + (uint32_t)j * SEL_NT;
      if (i >= n) continue;
      uint32_t sk, pos;
      order_decode<DT>(mode, x[j], sk, pos);
      const uint32_t bits = sk ? key_to_bits<DT, LARGEST>(sk) : (uint32_t)ldg(rb + pos);
      ov[rank[j]] = (U)bits;
      put_idx(p, d, rank[j], caller_idx(p, d, pos));
    }
    if (c == 0) {
      for (int64_t i = (int64_t)k + threadIdx.x; i < d.slots; i += SEL_NT) {
        ov[i] = (U)0;
        put_idx(p, d, i, -1);
      }
    }
    __syncthreads();
    RK_STAMP(p, r, sb + 3);
  }
}

// ---------------------------------------------------------------------------
// Exact passes (RadiK's countBin / selectBin, radixselect_l.cuh:14-126,
// 173-267, on the selection key): per pass, a histogram of the next digit of
// the keys that share the resolved prefix, then the bin holding the k-th
// key.  A row leaves the passes as soon as the keys >= the selected bin's
// lower bound fit its candidate slots (threshold known exactly: one filter
// read, then the select), or, with every bit resolved, as PH_ORDERED (the
// ordered emit's counting and writing reads).
//
// All of it runs in one persistent launch, k_exact, over a device-side work
// list: the rows handed to it (xrows) advance through their passes on their
// own, each pass split into chunks of x_chunk_tiles tiles that any
// workgroup may claim; the workgroup that completes a pass's last chunk
// selects the bin and opens the row's next pass.  So an idle call costs one
// launch (no row fell back), and a row that falls back streams on the whole
// grid.  No workgroup waits on another one's progress except for chunks that
// running workgroups hold, so the launch needs no co-residency.
// ---------------------------------------------------------------------------
// Exact calls (RK_EXACT) only: in sampled calls the sample and select
// kernels start the rows they send to the exact passes (exact_start_row).
__global__ __launch_bounds__(256) void k_exact_init(KParams p) {
  if (p.hdr->fallback_rows == 0) return;
  for (int r = blockIdx.x; r < p.R; r += gridDim.x) {
    const RowState s = p.st[r];
    if (s.phase == PH_HIST && (s.flags & RF_NEED_INIT)) exact_start_row<256>(p, r, FB_NONE);
  }
}

// Histogram LDS of one streaming workgroup: one sub-histogram per wave.
struct HistLds {
  uint32_t sub[STREAM_NT / 64][HIST_MAX_BINS];
};

// Digit histogram of tiles [ta, tb) of PH_HIST row r added to the row's
// global histogram: per-wave LDS sub-histograms (one atomic per wave when
// all its counted lanes hit one bin, as on skewed rows), merged with one
// global atomic per non-empty bin.  The row's vector grid is streamed with
// buffer loads, software pipelined as in the filter; the loads keep the
// lines cached, since the row's next pass reads them again.  Called by a
// whole workgroup.
template <int DT, bool LARGEST, int U>
RK_DEV void hist_segment(const KParams& p, HistLds& lds, const RowState& st, const RowDesc& d,
                         int r, uint32_t ta, uint32_t tb) {
  using UT = typename Traits<DT>::U;
  constexpr int P = Traits<DT>::PACK;
  constexpr int KB = Traits<DT>::KEY_BITS;
  constexpr int NW = STREAM_NT / 64;
#ifdef RK_PHASE_TIMING
  const uint64_t hs0 = wall_clock64();
#endif
  for (int i = threadIdx.x; i < NW * HIST_MAX_BINS; i += STREAM_NT) (&lds.sub[0][0])[i] = 0;
  __syncthreads();
#ifdef RK_PHASE_TIMING
  const uint64_t hs1 = wall_clock64();
#endif
  const int lane = lane_id();
  const int wave_base = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63);
  uint32_t* h = lds.sub[wave_base >> 6];
  const int HB = p.tune.hist_bits;
  const uint32_t b = st.resolved;
  const uint32_t P0 = st.prefix;
  const int w = (KB - (int)b) < HB ? (KB - (int)b) : HB;
  const int shift = KB - (int)b - w;
  const uint32_t nb = 1u << w;
  const UT* rb = static_cast<const UT*>(d.base);
  const SegGrid g = seg_grid<DT>(d.base, d.len, p.tune.tile, ta, tb);

  // One key: counted when it carries the resolved prefix.  Within the lanes
  // that count it, the first one adds for all of them when they share a
  // bin (readfirstlane / ballot, no lane exchange through LDS).
  uint32_t kor = 0u, kand = 0xffffffffu;  // of the keys this thread counts
  auto count = [&](uint32_t u, bool valid) {
    const uint32_t sk = sel_key<DT, LARGEST>(u);
    const bool ok = valid && (b == 0 || (uint32_t)((uint64_t)sk >> (KB - b)) == P0);
    if (ok) {
      kor |= sk;
      kand &= sk;
      const uint32_t dg = (sk >> shift) & (nb - 1);
      const uint32_t d0 = __builtin_amdgcn_readfirstlane(dg);
      const uint64_t act = __ballot(true);
      const uint64_t same = __ballot(dg == d0);
      if (same == act) {
        if (mbcnt(act) == 0) atomicAdd(&h[d0], (uint32_t)__popcll(act));
      } else {
        atomicAdd(&h[dg], 1u);
      }
    }
  };

  // the row's head and tail elements outside the vector grid, one per lane
  if ((wave_base == 0 && g.first_seg) || (wave_base == STREAM_NT - 64 && g.last_seg)) {
    int64_t pos = -1;
    if (wave_base == 0 && g.first_seg && lane < g.a0) pos = lane;
    if (wave_base != 0 && lane < P && g.tail0 + lane < d.len) pos = g.tail0 + lane;
    count(pos >= 0 ? (uint32_t)rb[pos] : 0u, pos >= 0);
  }
  const uint4* vp = reinterpret_cast<const uint4*>(rb + g.al);
  constexpr uint32_t STEP = (uint32_t)STREAM_NT * U;
  constexpr int64_t CHUNK = int64_t(1) << 26;  // vectors per descriptor (32-bit offsets)
  for (int64_t c0 = 0; c0 < g.nv; c0 += CHUNK) {
    const uint32_t nvc = (uint32_t)((g.nv - c0) < CHUNK ? (g.nv - c0) : CHUNK);
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(vp + c0, nvc * 16u);
    auto load_set = [&](uint4 (&x)[U], uint32_t jb) {
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = bload16(rs, (jb + lane + u * STREAM_NT) * 16u);
    };
    uint4 x[U], y[U];
    uint32_t jb = (uint32_t)wave_base;
    if (jb < nvc) load_set(x, jb);
    for (; jb < nvc; jb += STEP) {  // wave-uniform
      load_set(y, jb + STEP);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool valid = jb + lane + u * STREAM_NT < nvc;
        const Vec<DT> v = unpack<DT>(x[u]);
#pragma unroll
        for (int c = 0; c < P; ++c) count(v.e[c], valid);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = y[u];
    }
  }
  {
    // the workgroup's OR and AND into the row's (its barriers also complete
    // the LDS histograms); OR of nothing is 0, AND of nothing all ones
    __shared__ uint64_t sor[NW];
    const uint64_t both = block_or64<STREAM_NT>(((uint64_t)(~kand) << 32) | kor, sor);
    if (threadIdx.x == 0 && both != 0) {
      atomicOr(&p.xst[r].kor, (uint32_t)both);
      atomicAnd(&p.xst[r].kand, ~(uint32_t)(both >> 32));
    }
  }
#ifdef RK_PHASE_TIMING
  const uint64_t hs2 = wall_clock64();
  uint32_t nz = 0;
#endif
  uint32_t* gh = p.ghist + (size_t)r * HIST_MAX_BINS;
  for (uint32_t i = threadIdx.x; i < nb; i += STREAM_NT) {
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) s += lds.sub[q][i];
    if (s) atomicAdd(&gh[i], s);
#ifdef RK_PHASE_TIMING
    nz += s ? 1u : 0u;
#endif
  }
#ifdef RK_PHASE_TIMING
  wait_vmem();
  __syncthreads();
  const uint64_t hs3 = wall_clock64();
  nz = block_sum<STREAM_NT>(nz, reinterpret_cast<uint32_t*>(&lds.sub[0][0]));
  if (threadIdx.x == 0) {
    uint32_t* xt = p.hdr->pad;
    atomicAdd(&xt[40], (uint32_t)(hs1 - hs0));
    atomicAdd(&xt[41], (uint32_t)(hs2 - hs1));
    atomicAdd(&xt[42], (uint32_t)(hs3 - hs2));
    atomicAdd(&xt[43], nz);
    atomicAdd(&xt[44], 1u);
  }
  __syncthreads();
#endif
}

// Ordered emit for rows whose ties at the k-th key V exceed the candidate
// slots (e.g. an all-equal row), and unsorted rows with k above the LDS
// select.  Counting pass: per chunk, the keys > V and == V; the pass's last
// workgroup turns the chunk counts into prefix sums.  Writing pass: each
// chunk writes, in row-position order, its keys > V and those == V whose
// rank among equals is below rem, at output position gt_before +
// min(rem, eq_before).  A chunk is a segment of the row's vector grid (its
// head and tail elements in the first and last chunk), the same elements in
// both passes.
template <int DT, bool LARGEST, int U>
RK_DEV void emit_count_chunk(const KParams& p, const RowState& st, const RowDesc& d, int r,
                             uint32_t c, uint32_t ta, uint32_t tb, uint32_t* scratch) {
  using UT = typename Traits<DT>::U;
  constexpr int P = Traits<DT>::PACK;
  const uint32_t V = st.V;
  const UT* rb = static_cast<const UT*>(d.base);
  const SegGrid g = seg_grid<DT>(d.base, d.len, p.tune.tile, ta, tb);
  const int lane = lane_id();
  const int wave_base = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63);
  uint32_t gt = 0, eq = 0;
  auto tally = [&](uint32_t u, bool valid) {
    const uint32_t sk = sel_key<DT, LARGEST>(u);
    gt += (valid && sk > V) ? 1u : 0u;
    eq += (valid && sk == V) ? 1u : 0u;
  };
  if ((wave_base == 0 && g.first_seg) || (wave_base == STREAM_NT - 64 && g.last_seg)) {
    int64_t pos = -1;
    if (wave_base == 0 && g.first_seg && lane < g.a0) pos = lane;
    if (wave_base != 0 && lane < P && g.tail0 + lane < d.len) pos = g.tail0 + lane;
    tally(pos >= 0 ? (uint32_t)rb[pos] : 0u, pos >= 0);
  }
  const uint4* vp = reinterpret_cast<const uint4*>(rb + g.al);
  constexpr uint32_t STEP = (uint32_t)STREAM_NT * U;
  constexpr int64_t CHUNK = int64_t(1) << 26;
  for (int64_t c0 = 0; c0 < g.nv; c0 += CHUNK) {
    const uint32_t nvc = (uint32_t)((g.nv - c0) < CHUNK ? (g.nv - c0) : CHUNK);
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(vp + c0, nvc * 16u);
    auto load_set = [&](uint4 (&x)[U], uint32_t jb) {
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = bload16(rs, (jb + lane + u * STREAM_NT) * 16u);
    };
    uint4 x[U], y[U];
    uint32_t jb = (uint32_t)wave_base;
    if (jb < nvc) load_set(x, jb);
    for (; jb < nvc; jb += STEP) {
      load_set(y, jb + STEP);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool valid = jb + lane + u * STREAM_NT < nvc;
        const Vec<DT> v = unpack<DT>(x[u]);
#pragma unroll
        for (int e = 0; e < P; ++e) tally(v.e[e], valid);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = y[u];
    }
  }
  gt = block_sum<STREAM_NT>(gt, scratch);
  eq = block_sum<STREAM_NT>(eq, scratch);
  if (threadIdx.x == 0)
    st_dev(reinterpret_cast<uint64_t*>(p.tile_cnt + p.tile_start[r] + c),
           ((uint64_t)eq << 32) | gt);
}

template <int DT, bool LARGEST, int U>
RK_DEV void emit_write_chunk(const KParams& p, const RowState& st, const RowDesc& d, int r,
                             uint32_t c, uint32_t ta, uint32_t tb, uint32_t* scratch) {
  using UT = typename Traits<DT>::U;
  constexpr int P = Traits<DT>::PACK;
  const uint32_t V = st.V;
  const uint32_t rem = st.rem;
  const uint32_t kk = (uint32_t)d.k;
  const UT* rb = static_cast<const UT*>(d.base);
  UT* ov = static_cast<UT*>(d.out_v);
  const SegGrid g = seg_grid<DT>(d.base, d.len, p.tune.tile, ta, tb);
  // keys > V and == V in the row's chunks before this one
  const uint64_t pre = ld_dev(reinterpret_cast<const uint64_t*>(p.tile_cnt + p.tile_start[r] + c));
  uint32_t run_g = (uint32_t)pre, run_e = (uint32_t)(pre >> 32);
  // One block step: thread t holds n elements at row positions first..,
  // consecutive in the row across threads (t, then e); keys > V and the
  // first rem keys == V are written at their ranks.
  // (n <= NMAX; the element loops are unrolled over NMAX, so the array
  // stays in registers)
  auto step = [&](auto nmax_tag, const uint32_t* u, uint32_t n, int64_t first) {
    constexpr int NMAX = decltype(nmax_tag)::value;
    uint32_t gq = 0;
#pragma unroll
    for (int e = 0; e < NMAX; ++e) {
      const uint32_t sk = sel_key<DT, LARGEST>(u[e]);
      gq += (uint32_t)e >= n ? 0u : (sk > V ? 1u : (sk == V ? 0x10000u : 0u));
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<STREAM_NT>(gq, scratch, tot);
    uint32_t gb = run_g + (ex & 0xffffu);
    uint32_t qb = run_e + (ex >> 16);
#pragma unroll
    for (int e = 0; e < NMAX; ++e) {
      const uint32_t sk = (uint32_t)e < n ? sel_key<DT, LARGEST>(u[e]) : 0u;
      const uint32_t pos = (uint32_t)(first + e);
      if ((uint32_t)e >= n) {
        // past this thread's elements
      } else if (sk > V) {
        const uint32_t o = gb + (qb < rem ? qb : rem);
        if (o < kk) {
          ov[o] = (UT)u[e];
          put_idx(p, d, o, pos);  // a position: mapped by finalize_ordered_rows
        }
        ++gb;
      } else if (sk == V) {
        if (qb < rem) {
          const uint32_t o = gb + qb;
          if (o < kk) {
            ov[o] = (UT)u[e];
            put_idx(p, d, o, pos);
          }
        }
        ++qb;
      }
    }
    run_g += tot & 0xffffu;
    run_e += tot >> 16;
  };
  // head elements (first chunk): one per thread, before every vector
  if (g.first_seg) {
    const int64_t pos = threadIdx.x;
    uint32_t u = 0;
    const uint32_t n = pos < g.a0 ? 1u : 0u;
    if (n) u = (uint32_t)rb[pos];
    step(std::integral_constant<int, 1>{}, &u, n, pos);
  }
  // body: UW consecutive vectors per thread (position order is
  // thread-major within a block step; at most 4, so that the step's
  // elements stay in registers); the next step's loads are issued before
  // this one is written
  constexpr int UW = U < 4 ? U : 4;
  const uint4* vp = reinterpret_cast<const uint4*>(rb + g.al);
  constexpr uint32_t BLK = (uint32_t)STREAM_NT * UW;
  constexpr int64_t CHUNK = int64_t(1) << 26;
  for (int64_t c0 = 0; c0 < g.nv; c0 += CHUNK) {
    const uint32_t nvc = (uint32_t)((g.nv - c0) < CHUNK ? (g.nv - c0) : CHUNK);
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(vp + c0, nvc * 16u);
    auto load_set = [&](uint4 (&x)[UW], uint32_t jb) {
#pragma unroll
      for (int u = 0; u < UW; ++u) x[u] = bload16(rs, (jb + threadIdx.x * UW + u) * 16u);
    };
    uint4 x[UW], y[UW];
    uint32_t jb = 0;
    load_set(x, jb);
    for (; jb < nvc; jb += BLK) {  // workgroup-uniform
      load_set(y, jb + BLK);
      const uint32_t j0 = jb + threadIdx.x * UW;
      const uint32_t nvec = j0 >= nvc ? 0u : (nvc - j0 < (uint32_t)UW ? nvc - j0 : (uint32_t)UW);
      uint32_t u[UW * P];
#pragma unroll
      for (int q = 0; q < UW; ++q) {
        const Vec<DT> v = unpack<DT>(x[q]);
#pragma unroll
        for (int e = 0; e < P; ++e) u[q * P + e] = v.e[e];
      }
      step(std::integral_constant<int, UW * P>{}, u, nvec * P, g.al + (int64_t)(c0 + j0) * P);
#pragma unroll
      for (int q = 0; q < UW; ++q) x[q] = y[q];
    }
  }
  // tail elements (last chunk): after every vector
  if (g.last_seg) {
    const int64_t pos = g.tail0 + threadIdx.x;
    uint32_t u = 0;
    const uint32_t n = pos < d.len ? 1u : 0u;
    if (n) u = (uint32_t)rb[pos];
    step(std::integral_constant<int, 1>{}, &u, n, pos);
  }
}

// The bracket pass over tiles [ta, tb) of a PH_BRACKET row (bracket [T, V]
// in selection keys): the keys above V counted (stored per chunk, the
// ordered emit's counts start from them) and the keys in [T, V] appended as
// composite keys, with device-coherent stores: the pass's finisher reads
// them in this launch.  Called by a whole workgroup.
template <int DT, bool LARGEST, int U>
RK_DEV void bracket_segment(const KParams& p, FilterLds& lds, const RowState& st,
                            const RowDesc& d, int r, uint32_t c, uint32_t ta, uint32_t tb,
                            uint32_t* scratch) {
  using UT = typename Traits<DT>::U;
  constexpr int P = Traits<DT>::PACK;
  const uint32_t lo = st.T, hi = st.V;
  const UT* rb = static_cast<const UT*>(d.base);
  const SegGrid g = seg_grid<DT>(d.base, d.len, p.tune.tile, ta, tb);
  const int lane = lane_id();
  const int wave_base = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63);
  AppendCtx ac;
  ac.stage = lds.stage[wave_base >> 6];
  ac.cand = p.cand + p.cand_base[r];
  ac.cand_cnt = &p.cand_cnt[r];
  ac.cap = st.cap;
  ac.coh = true;
  uint32_t staged = 0;
  bool full = false;
  uint32_t gt = 0;
  // the keys of one 16-byte vector (or one element), counted and appended
  auto take = [&](const uint4& x, uint32_t valid_mask, int64_t first) {
    const Vec<DT> v = unpack<DT>(x);
    uint32_t m = 0;
#pragma unroll
    for (int e = 0; e < P; ++e) {
      const uint32_t sk = sel_key<DT, LARGEST>(v.e[e]);
      const bool ok = (valid_mask >> e) & 1u;
      gt += (ok && sk > hi) ? 1u : 0u;
      m |= (ok && sk >= lo && sk <= hi ? 1u : 0u) << e;
    }
    if (__ballot(m != 0u)) append_vec<DT, LARGEST>(ac, staged, full, x, m, first);
  };
  if ((wave_base == 0 && g.first_seg) || (wave_base == STREAM_NT - 64 && g.last_seg)) {
    int64_t pos = -1;
    if (wave_base == 0 && g.first_seg && lane < g.a0) pos = lane;
    if (wave_base != 0 && lane < P && g.tail0 + lane < d.len) pos = g.tail0 + lane;
    uint4 x = make_uint4(0u, 0u, 0u, 0u);
    if (pos >= 0) x.x = (uint32_t)rb[pos];  // element 0 of the vector
    take(x, pos >= 0 ? 1u : 0u, pos);
  }
  const uint4* vp = reinterpret_cast<const uint4*>(rb + g.al);
  constexpr uint32_t STEP = (uint32_t)STREAM_NT * U;
  constexpr int64_t CHUNK = int64_t(1) << 26;
  for (int64_t c0 = 0; c0 < g.nv; c0 += CHUNK) {
    const uint32_t nvc = (uint32_t)((g.nv - c0) < CHUNK ? (g.nv - c0) : CHUNK);
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(vp + c0, nvc * 16u);
    auto load_set = [&](uint4 (&x)[U], uint32_t jb) {
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = bload16(rs, (jb + lane + u * STREAM_NT) * 16u);
    };
    uint4 x[U], y[U];
    uint32_t jb = (uint32_t)wave_base;
    if (jb < nvc) load_set(x, jb);
    for (; jb < nvc; jb += STEP) {
      load_set(y, jb + STEP);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t j = jb + lane + u * STREAM_NT;
        take(x[u], j < nvc ? (1u << P) - 1u : 0u, g.al + (int64_t)(c0 + j) * P);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = y[u];
    }
  }
  // staged keys: one reservation per workgroup, then coherent stores
  if (lane == 0) lds.flush_n[wave_base >> 6] = staged;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int w = 0; w < STREAM_NT / 64; ++w) tot += lds.flush_n[w];
    uint32_t b = tot ? atomicAdd(ac.cand_cnt, tot) : 0u;
    for (int w = 0; w < STREAM_NT / 64; ++w) {
      lds.flush_base[w] = b;
      b += lds.flush_n[w];
    }
  }
  __syncthreads();
  const uint32_t fb = lds.flush_base[wave_base >> 6];
  for (uint32_t i = lane; i < staged; i += 64)
    if (fb + i < ac.cap) st_dev(ac.cand + fb + i, ac.stage[i]);
  gt = block_sum<STREAM_NT>(gt, scratch);
  if (threadIdx.x == 0)
    st_dev(reinterpret_cast<uint64_t*>(p.tile_cnt + p.tile_start[r] + c), (uint64_t)gt);
}

// The bracket pass of row r is complete.  With G keys above the bracket and
// n inside it: if G < k <= G + n (and the bracket fit its slots), the k-th
// key V is the (k - G)-th largest composite key inside (a radix select over
// them, in global memory), each chunk's keys above V and equal to V are its
// count above the bracket plus its bracket keys above V, and its bracket keys
// equal to V (counted by row position in LDS), and the row goes on to the
// ordered emit's writing pass with their prefix sums.  Otherwise it takes
// the digit passes from the start.  Returns whether the row leaves.
template <int DT>
RK_DEV bool finish_bracket(const KParams& p, int r, const RowState& s, uint32_t nch, uint32_t ct,
                           HistLds& lh, uint32_t* scratch, uint64_t* scratch64, uint32_t* outv) {
  constexpr int KB = Traits<DT>::KEY_BITS;
  const RowDesc d = p.desc[r];
  const uint32_t k = (uint32_t)d.k;
  const uint32_t n = ld_dev(p.cand_cnt + r);
  uint64_t* tc = reinterpret_cast<uint64_t*>(p.tile_cnt + p.tile_start[r]);
  uint32_t G = 0;
  for (uint32_t c = threadIdx.x; c < nch; c += STREAM_NT) G += (uint32_t)ld_dev(tc + c);
  G = block_sum<STREAM_NT>(G, scratch);
  uint32_t* lgt = &lh.sub[0][0];  // per chunk: bracket keys > V, then == V
  uint32_t* leq = lgt + 4096;
  const bool ok = n <= s.cap && G < k && k <= G + n && nch <= 4096u;  // uniform
  if (!ok) {
    if (threadIdx.x == 0) {
      // the digit passes from the start (histogram of the row is still zero)
      RowState& w = p.st[r];
      st_dev(&w.prefix, 0u);
      st_dev(&w.resolved, 0u);
      st_dev(&w.rem, k);
      st_dev(&w.why, (uint32_t)FB_BRACKET);
      st_dev(&w.cap, 0u);
      st_dev(&w.phase, (uint32_t)PH_HIST);
    }
    return false;
  }
  const uint64_t* cand = p.cand + p.cand_base[r];
  const uint64_t kth = block_radix_select<STREAM_NT, 11, uint64_t, true, true>(
      cand, n, k - G, lh.sub[1], scratch, scratch64, outv);
  const uint32_t V = (uint32_t)(kth >> 32);
  for (uint32_t c = threadIdx.x; c < nch; c += STREAM_NT) {
    lgt[c] = 0u;
    leq[c] = 0u;
  }
  __syncthreads();
  // the row's vector grid: chunk c holds elements [a0 + c*ct*tile, ...),
  // the head elements (before a0) are in chunk 0, the tail in the last
  const SegGrid g = seg_grid<DT>(d.base, d.len, p.tune.tile, 0u, 1u);
  const int64_t span = (int64_t)ct * p.tune.tile;
  for (uint32_t i = threadIdx.x; i < n; i += STREAM_NT) {
    const uint64_t x = ld_dev(cand + i);
    const uint32_t key = (uint32_t)(x >> 32);
    if (key < V) continue;
    const int64_t pos = (int64_t)(~(uint32_t)x);
    int64_t c = pos < g.a0 ? 0 : (pos - g.a0) / span;
    if (c > (int64_t)nch - 1) c = (int64_t)nch - 1;
    atomicAdd(key > V ? &lgt[c] : &leq[c], 1u);
  }
  __syncthreads();
  // chunk prefix sums of (keys > V, keys == V), as the counting pass leaves them
  uint32_t carry_g = 0, carry_e = 0;
  for (uint32_t b0 = 0; b0 < nch; b0 += STREAM_NT) {
    const uint32_t c = b0 + threadIdx.x;
    const uint32_t gv = c < nch ? (uint32_t)ld_dev(tc + c) + lgt[c] : 0u;
    const uint32_t ev = c < nch ? leq[c] : 0u;
    uint32_t tg, te;
    const uint32_t eg = block_excl_scan<STREAM_NT>(gv, scratch, tg);
    const uint32_t ee = block_excl_scan<STREAM_NT>(ev, scratch, te);
    if (c < nch) st_dev(tc + c, ((uint64_t)(carry_e + ee) << 32) | (carry_g + eg));
    carry_g += tg;
    carry_e += te;
  }
  if (threadIdx.x == 0) {
    RowState& w = p.st[r];
    st_dev(&w.V, V);
    st_dev(&w.rem, k - carry_g);  // keys == V to take (carry_g: all keys > V)
    st_dev(&w.flags, (uint32_t)RF_EMIT_WRITE);
    st_dev(&w.phase, (uint32_t)PH_ORDERED);
    atomicAdd(&p.hdr->ordered_rows, 1u);
  }
  (void)KB;
  return false;
}

// A histogram pass of row r is complete: the bin of the k-th key (the
// global histogram is cleared as it is read, for the next pass), then the
// row's next stage.  Returns (to every thread) whether the row leaves the
// executor: to the select (threshold known) or on an internal error.
template <int KB>
RK_DEV bool finish_hist(const KParams& p, int r, const RowState& s, uint32_t* hist,
                        uint32_t* scratch, uint32_t* outv) {
  const int HB = p.tune.hist_bits;
  const int b = (int)s.resolved;
  const int w = (KB - b) < HB ? (KB - b) : HB;
  const uint32_t nb = 1u << w;
  uint32_t* g = p.ghist + (size_t)r * HIST_MAX_BINS;
  // all of a thread's bins loaded before any is used: one round trip
  constexpr int PER = HIST_MAX_BINS / STREAM_NT;
  uint32_t hv[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t i = threadIdx.x + (uint32_t)j * STREAM_NT;
    hv[j] = i < nb ? ld_dev(g + i) : 0u;
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t i = threadIdx.x + (uint32_t)j * STREAM_NT;
    if (i < nb) {
      hist[i] = hv[j];
      st_dev(g + i, 0u);
    }
  }
  __syncthreads();
  block_find_crossing<STREAM_NT>(hist, (int)nb, s.rem, scratch, outv, &outv[3]);
  if (threadIdx.x == 0) {
    RowState& st = p.st[r];
    const uint32_t bin = outv[0];
    const uint32_t counted = outv[3];
    uint32_t leave = 0;
    if (bin == 0xffffffffu) {
      atomicOr(&p.hdr->status, (uint32_t)ST_INTERNAL);
      st_dev(&st.phase, (uint32_t)PH_DONE);
      leave = 1;
    } else {
      uint32_t prefix = (s.prefix << w) | bin;
      uint32_t resolved = (uint32_t)(b + w);
      if (outv[2] == counted && resolved < (uint32_t)KB) {
        // Every key the pass counted is in this bin: they share the bits
        // down to their common prefix (OR == AND there), which the next
        // window skips, as RadiK's adaptive scaling spreads skewed rows
        // (utils.cuh:255-324), but exactly, in key space.
        const uint32_t kor = ld_dev(&p.xst[r].kor), kand = ld_dev(&p.xst[r].kand);
        const uint32_t diff = kor ^ kand;
        const uint32_t cp = diff == 0u ? (uint32_t)KB : (uint32_t)(__clz(diff) - (32 - KB));
        if (cp > resolved) {
          resolved = cp;
          prefix = resolved >= 32u ? kand : (kand >> (KB - (int)resolved));
        }
      }
      st_dev(&p.xst[r].kor, 0u);
      st_dev(&p.xst[r].kand, 0xffffffffu);
      const uint32_t rem = s.rem - outv[1];
      const uint32_t k = (uint32_t)p.desc[r].k;
      const uint64_t need = (uint64_t)(k - rem) + outv[2];
      if (need <= s.cap) {
        // the filter pass follows here, then the row leaves for the select
        st_dev(&st.T, prefix << (KB - (int)resolved));
        st_dev(&st.phase, (uint32_t)PH_FILTER);
        st_dev(&st.flags, (uint32_t)RF_EXACT_T);
        st_dev(p.cand_cnt + r, 0u);
      } else if ((int)resolved == KB) {
        st_dev(&st.V, prefix);
        st_dev(&st.rem, rem);
        st_dev(&st.phase, (uint32_t)PH_ORDERED);
        st_dev(&st.flags, 0u);
        atomicAdd(&p.hdr->ordered_rows, 1u);
      } else {
        st_dev(&st.prefix, prefix);
        st_dev(&st.resolved, resolved);
        st_dev(&st.rem, rem);
      }
    }
    outv[3] = leave;
  }
  __syncthreads();
  return outv[3] != 0;
}

// The counting pass of an ordered row is complete: chunk counts -> their
// exclusive prefix sums, in place.
RK_DEV void finish_count(const KParams& p, int r, uint32_t nch, uint32_t* scratch) {
  uint64_t* tc = reinterpret_cast<uint64_t*>(p.tile_cnt + p.tile_start[r]);
  uint32_t carry_g = 0, carry_e = 0;
  for (uint32_t b0 = 0; b0 < nch; b0 += STREAM_NT) {
    const uint32_t i = b0 + threadIdx.x;
    const uint64_t v = i < nch ? ld_dev(tc + i) : 0ull;
    uint32_t tg, te;
    const uint32_t eg = block_excl_scan<STREAM_NT>((uint32_t)v, scratch, tg);
    const uint32_t ee = block_excl_scan<STREAM_NT>((uint32_t)(v >> 32), scratch, te);
    if (i < nch) st_dev(tc + i, ((uint64_t)(carry_e + ee) << 32) | (carry_g + eg));
    carry_g += tg;
    carry_e += te;
  }
  if (threadIdx.x == 0) st_dev(&p.st[r].flags, (uint32_t)RF_EMIT_WRITE);
}

// Longest a workgroup waits for claimable work before it reports an
// internal error and leaves (wall clock, 100 MHz): no legitimate wait comes
// near it (a chunk's pass plus a bin selection take microseconds).
constexpr uint64_t X_MAX_WAIT = 100000000ull;  // 1 s

// The executor runs 1 or 2 workgroups per CU by default (4 at most from the
// autotuner's candidates): registers are bounded for 2 waves per SIMD, which
// keeps every variant free of scratch spills.
template <int DT, bool LARGEST, int U>
__global__ __launch_bounds__(STREAM_NT) __attribute__((amdgpu_waves_per_eu(U >= 8 ? 1 : 2)))
void k_exact(KParams p) {
  const uint32_t nrows = min(p.hdr->x_rows, (uint32_t)p.R);  // set by earlier launches
  if (nrows == 0) return;
  constexpr int KB = Traits<DT>::KEY_BITS;
  __shared__ union {
    FilterLds f;
    HistLds h;
  } lds;
  __shared__ uint32_t scratch[16];
  __shared__ uint64_t scratch64[3 * STREAM_NT / 64];
  __shared__ uint32_t outv[4];
  __shared__ int r_sh;
  __shared__ uint32_t c_sh, seq_sh, last_sh, cont_sh;
  const int tid = threadIdx.x;
  uint32_t cursor = blockIdx.x % nrows;
  uint64_t wait0 = 0;
  uint32_t idle_naps = 1;
  RowState st;  // pass state of row st_r in pass st_seq
  int st_r = -1;
  uint32_t st_seq = 0;
#ifdef RK_PHASE_TIMING
  // tuning builds: the header's spare words hold workgroup 0's entry and
  // exit and, for the first listed row, each stage's first claim and end
  uint32_t* const xt = p.hdr->pad;
  const int r_t = (int)p.xrows[0];
  if (blockIdx.x == 0 && tid == 0) xt[0] = (uint32_t)wall_clock64();
  // per-workgroup totals (ticks of 10 ns): claiming, idle waits, chunk work,
  // chunk completion, pass finishing; chunks done, failed claims
  uint32_t a_claim = 0, a_idle = 0, a_work = 0, a_done = 0, a_fin = 0, a_n = 0, a_fail = 0;
  uint64_t t_mark = wall_clock64();
  auto lap = [&](uint32_t& acc) {
    const uint64_t t = wall_clock64();
    acc += (uint32_t)(t - t_mark);
    t_mark = t;
  };
#define RK_LAP(acc) lap(acc)
#else
#define RK_LAP(acc) do {} while (0)
#endif
  bool cont = false;  // this workgroup opened a pass and took its first chunk
  for (;;) {
    // Claim a chunk: wave 0 looks at 64 rows of the list at a time, from
    // this workgroup's cursor, and claims in the first row with chunks left.
    if (!cont && tid < 64) {
      int got = -1;
      uint32_t gc = 0, gseq = 0;
      for (uint32_t base = 0; base < nrows && got < 0; base += 64) {
        const uint32_t i = base + (uint32_t)tid;
        int r = -1;
        bool avail = false;
        if (i < nrows) {
          r = (int)p.xrows[(cursor + i) % nrows];
          const uint64_t w = __hip_atomic_load(&p.xst[r].claim, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
          avail = (uint32_t)(w >> 32) != XSEQ_DONE && (uint32_t)w < p.xst[r].nch;
        }
        uint64_t m = __ballot(avail);
        while (m) {
          const int l = __ffsll((unsigned long long)m) - 1;
          m &= m - 1;
          uint64_t old = 0;
          if (tid == l)
            old = __hip_atomic_fetch_add(&p.xst[r].claim, 1ull, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
          old = __shfl(old, l, 64);
          const int rl = __shfl(r, l, 64);
#ifdef RK_PHASE_TIMING
          if (!((uint32_t)(old >> 32) != XSEQ_DONE && (uint32_t)old < p.xst[rl].nch)) ++a_fail;
#endif
          if ((uint32_t)(old >> 32) != XSEQ_DONE && (uint32_t)old < p.xst[rl].nch) {
            got = rl;
            gc = (uint32_t)old;
            gseq = (uint32_t)(old >> 32);
            cursor = (cursor + base + (uint32_t)l) % nrows;
            break;
          }
        }
      }
      if (tid == 0) {
        r_sh = got;
        c_sh = gc;
        seq_sh = gseq;
      }
    }
    __syncthreads();
    const int r = r_sh;
    const uint32_t c = c_sh, seq = seq_sh;
    __syncthreads();
    if (r < 0) {
      // Nothing claimable: done when every row has left, else wait for the
      // workgroups that hold the last chunks of a pass to open the next one.
      const uint32_t left = __hip_atomic_load(&p.hdr->x_left, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
      if (left == 0) {
#ifdef RK_PHASE_TIMING
        RK_LAP(a_idle);
        if (blockIdx.x == 0 && tid == 0) xt[1] = (uint32_t)wall_clock64();
        if (tid == 0) {
          atomicAdd(&xt[30], a_claim);
          atomicAdd(&xt[31], a_idle);
          atomicAdd(&xt[32], a_work);
          atomicAdd(&xt[33], a_done);
          atomicAdd(&xt[34], a_fin);
          atomicAdd(&xt[35], a_n);
          atomicAdd(&xt[36], a_fail);
          atomicAdd(&xt[37], 1u);
        }
#endif
        return;
      }
      const uint64_t now = wall_clock64();
      if (wait0 == 0) wait0 = now;
      if (now - wait0 > X_MAX_WAIT) {
        if (tid == 0) atomicOr(&p.hdr->status, (uint32_t)ST_INTERNAL);
        return;
      }
      // back off while idle (up to ~1.7 us between polls), so that waiting
      // workgroups do not crowd the claim words' memory channel
      for (uint32_t i = 0; i < idle_naps; ++i) __builtin_amdgcn_s_sleep(2);
      idle_naps = idle_naps < 32 ? idle_naps * 2 : 32;
      RK_LAP(a_idle);
      continue;
    }
    RK_LAP(a_claim);
    wait0 = 0;
    idle_naps = 1;
#ifdef RK_PHASE_TIMING
    if (r == r_t && c == 0 && tid == 0 && seq < 12) xt[4 + 2 * seq] = (uint32_t)wall_clock64();
#endif
    // The row's pass state, as the workgroup that opened this pass left it
    // (read after the claim that returned the pass's sequence number; it
    // does not change within a pass, so a workgroup keeps it for the next
    // chunk of the same pass).
    if (r != st_r || seq != st_seq) {
      const uint32_t* src = reinterpret_cast<const uint32_t*>(p.st + r);
      uint32_t* dst = reinterpret_cast<uint32_t*>(&st);
#pragma unroll
      for (int i = 0; i < (int)(sizeof(RowState) / 4); ++i) dst[i] = ld_dev(src + i);
      st_r = r;
      st_seq = seq;
    }
    const RowDesc d = p.desc[r];  // constant in this launch: plain loads
    const uint32_t nch = p.xst[r].nch, ct = p.xst[r].ct;
    const uint32_t tiles = p.tile_start[r + 1] - p.tile_start[r];
    const uint32_t ta = c * ct;
    const uint32_t tb = tiles - ta < ct ? tiles : ta + ct;
    if (st.phase == PH_HIST) {
      hist_segment<DT, LARGEST, U>(p, lds.h, st, d, r, ta, tb);
    } else if (st.phase == PH_BRACKET) {
      bracket_segment<DT, LARGEST, U>(p, lds.f, st, d, r, c, ta, tb, scratch);
    } else if (st.phase == PH_FILTER) {
      filter_segment<DT, LARGEST, U>(p, lds.f, st, d, p.cand_base[r], r, ta, tb);
    } else if (st.phase == PH_ORDERED && !(st.flags & RF_EMIT_WRITE)) {
      emit_count_chunk<DT, LARGEST, U>(p, st, d, r, c, ta, tb, scratch);
    } else if (st.phase == PH_ORDERED) {
      emit_write_chunk<DT, LARGEST, U>(p, st, d, r, c, ta, tb, scratch);
    } else if (tid == 0) {
      atomicOr(&p.hdr->status, (uint32_t)ST_INTERNAL);  // a claimed chunk in no pass
    }
#ifdef RK_PHASE_TIMING
    ++a_n;
#endif
    RK_LAP(a_work);
    // Chunk done: this workgroup's histogram atomics and chunk counts
    // (device-coherent) are performed before the count that may let another
    // workgroup finish the pass.  (Candidates and outputs are read by later
    // launches only.)
    wait_vmem();
    __syncthreads();
    if (tid == 0) {
      const uint32_t prev = __hip_atomic_fetch_add(&p.xst[r].done, 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
      last_sh = prev + 1 == nch ? 1u : 0u;
    }
    __syncthreads();
    RK_LAP(a_done);
    cont = false;
    if (!last_sh) continue;
    // This workgroup completed the pass: the row's next stage.
#ifdef RK_PHASE_TIMING
    if (r == r_t && tid == 0 && seq < 12) {
      xt[5 + 2 * seq] = (uint32_t)wall_clock64();
      xt[2] = nch;
      xt[3] = seq + 1;
    }
#endif
    bool leave;
    if (st.phase == PH_HIST) {
      leave = finish_hist<KB>(p, r, st, lds.h.sub[0], scratch, outv);
    } else if (st.phase == PH_BRACKET) {
      leave = finish_bracket<DT>(p, r, st, nch, ct, lds.h, scratch, scratch64, outv);
    } else if (st.phase == PH_ORDERED && !(st.flags & RF_EMIT_WRITE)) {
      finish_count(p, r, nch, scratch);
      leave = false;
    } else {
      leave = true;  // filter pass done (the select follows) / ordered rows written
    }
    // The next pass's state (device-coherent stores) is performed before the
    // claim word that opens it.
    wait_vmem();
    __syncthreads();
    if (tid == 0) {
      if (leave) {
        st_dev(&p.xst[r].claim, (uint64_t)XSEQ_DONE << 32);
        wait_vmem();
        __hip_atomic_fetch_add(&p.hdr->x_left, 0xffffffffu, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      } else {
        // the next pass opens with its first chunk taken by this workgroup,
        // which goes on to it without a claim round trip
        st_dev(&p.xst[r].done, 0u);
        wait_vmem();
        st_dev(&p.xst[r].claim, ((uint64_t)(seq + 1) << 32) | (uint64_t)1);
        c_sh = 0u;
        seq_sh = seq + 1;
      }
      cont_sh = leave ? 0u : 1u;
    }
    __syncthreads();
    cont = cont_sh != 0;
    RK_LAP(a_fin);
  }
#undef RK_LAP
}


// ---------------------------------------------------------------------------
// Host side: the launch sequence.
// ---------------------------------------------------------------------------
struct Launch {
  hipStream_t stream;
  int row_grid;   // grid of per-row kernels
  int grid;       // grid of streaming kernels
  int part;       // PART_ALL, or one half of the sequence (PART_MAIN / PART_FALLBACK)
  uint32_t* publish;  // PART_MAIN: header words to mapped host memory (or null)
};

template <int DT, bool LARGEST>
static void set_lds_attrs(size_t dyn) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sample<DT, LARGEST>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sample<DT, LARGEST, true>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_select_finalize<DT, LARGEST>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_select_finalize<DT, LARGEST, true>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_merge_rank<DT, LARGEST>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
}

// Optional timing of the streaming filter kernel (the dominant kernel):
// a pair of HIP events recorded around its first launch in each call.
struct TimingState {
  std::mutex mu;
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
};
static TimingState& timing() {
  static TimingState t;
  return t;
}

static bool timing_begin(hipStream_t s, std::pair<hipEvent_t, hipEvent_t>& ev) {
  TimingState& t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  if (!t.on) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
  if (!t.pool.empty()) {
    ev = t.pool.back();
    t.pool.pop_back();
  } else {
    if (hipEventCreate(&ev.first) != hipSuccess) return false;
    if (hipEventCreate(&ev.second) != hipSuccess) return false;
  }
  return hipEventRecord(ev.first, s) == hipSuccess;
}

static void timing_end(hipStream_t s, const std::pair<hipEvent_t, hipEvent_t>& ev) {
  TimingState& t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  (void)hipEventRecord(ev.second, s);
  t.pending.push_back(ev);
}

// Workgroups of a persistent kernel that can be resident at once on the
// device: the static partition needs every workgroup running from the start.
template <typename K>
static int resident_grid(K kernel, int threads, size_t dyn = 0) {
  static std::mutex mu;
  static std::vector<std::pair<std::pair<const void*, size_t>, int>> memo;
  const std::pair<const void*, size_t> key{reinterpret_cast<const void*>(kernel), dyn};
  std::lock_guard<std::mutex> lk(mu);
  for (auto& m : memo)
    if (m.first == key) return m.second;
  int dev = 0, cus = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, dyn) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int g = cus * per_cu;
  memo.emplace_back(key, g);
  return g;
}

// Compute units of the current device (memoized per device).
static int device_cus() {
  static std::mutex mu;
  static std::vector<int> memo;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  if ((int)memo.size() <= dev) memo.resize(dev + 1, 0);
  if (memo[dev] <= 0) {
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    memo[dev] = cus > 0 ? cus : 1;
  }
  return memo[dev];
}

template <int DT, bool LARGEST, int U>
static hipError_t run_typed_u(const KParams& p, const Launch& L0) {
  const size_t dyn = (size_t)p.lds_keys * sizeof(uint64_t);
  const size_t dyn_sample = (size_t)p.sample_keys * sizeof(uint32_t);
  const int rows_blocks = (p.R + 255) / 256;
  hipStream_t s = L0.stream;
  // Persistent streaming kernels: every workgroup resident from the start.
  const int fgrid = std::min(L0.grid, resident_grid(k_filter<DT, LARGEST, U>, STREAM_NT));
  // The exact-pass executor: the call's grid on the exact path, else its
  // own default (the grid of a sampled call is the filter's).
  const int xgrid = std::min((p.flags & FL_EXACT) ? L0.grid
                                                  : device_cus() * exact_grid_per_cu(p.total_len),
                             resident_grid(k_exact<DT, LARGEST, U>, STREAM_NT));
  // Per-row kernels loop over rows: one resident wave of workgroups suffices.
  auto rows = [&](int resident) { return std::max(1, std::min(L0.row_grid, resident)); };
  const int g_sample = rows(resident_grid(k_sample<DT, LARGEST>, SEL_NT, dyn_sample));
  const int g_select = rows(resident_grid(k_select_finalize<DT, LARGEST>, SEL_NT, dyn));
  const int g_init = rows(resident_grid(k_exact_init, 256));
  KParams p1 = p;  // fallback round: kernels return at once when no row fell back
  p1.round = 1;
  // In a sampled call the fallback round is almost always idle, and an idle
  // launch costs about as much as dispatching its workgroups: its per-row
  // kernels run one workgroup per CU.  The exact-pass executor keeps its
  // full grid, so that rows that do fall back stream at full bandwidth.
  const bool exact_call = (p.flags & FL_EXACT) != 0;
  // Rows with k > MERGE_CHUNK: the select compacts their k keys and these
  // two kernels sort them over many workgroups (4096-key chunks, then every
  // key placed by its rank), instead of one workgroup's sort of all k.
  auto launch_merge = [&](const KParams& pm) {
    const int cpr = merge_chunks(pm);
    const int64_t items = (int64_t)pm.R * cpr;
    const size_t drank = (size_t)cpr * MERGE_CHUNK * sizeof(uint64_t);
    const int g_ms = (int)std::min<int64_t>(items, resident_grid(k_merge_sort<DT>, MS_NT));
    const int g_mr = (int)std::min<int64_t>(
        items, resident_grid(k_merge_rank<DT, LARGEST>, SEL_NT, drank));
    k_merge_sort<DT><<<g_ms, MS_NT, 0, s>>>(pm);
    k_merge_rank<DT, LARGEST><<<g_mr, SEL_NT, drank, s>>>(pm);
  };
  int fb_sel = g_select;
  if (!exact_call) {
    const int cus = device_cus();
    fb_sel = std::min(fb_sel, cus);
  }
  // A sampled call of one row: its sample workgroup plans the call first
  // (one launch fewer on the latency-bound single-row path).
  const bool plan_in_sample = !exact_call && p.R == 1 && L0.part != PART_FALLBACK;
  if (L0.part != PART_FALLBACK && !plan_in_sample) {
    if (p.R <= 1024) {
      if (p.R <= 64) k_plan_scan<DT, true, 64><<<1, 64, 0, s>>>(p);
      else k_plan_scan<DT, true><<<1, 1024, 0, s>>>(p);
    } else {
      k_plan_rows<DT><<<rows_blocks, 256, 0, s>>>(p);
      k_plan_scan<DT, false><<<1, 1024, 0, s>>>(p);
    }
  }
  if (!(p.flags & FL_EXACT) && L0.part != PART_FALLBACK) {
    if (plan_in_sample) k_sample<DT, LARGEST, true><<<1, SEL_NT, dyn_sample, s>>>(p);
    else k_sample<DT, LARGEST><<<g_sample, SEL_NT, dyn_sample, s>>>(p);
    std::pair<hipEvent_t, hipEvent_t> ev;
    const bool timed = timing_begin(s, ev);
    k_filter<DT, LARGEST, U><<<fgrid, STREAM_NT, 0, s>>>(p);
    if (timed) timing_end(s, ev);
    KParams pm = p;
    pm.publish = L0.part == PART_MAIN ? L0.publish : nullptr;
    pm.merge = merge_chunks(p) > 1;
    k_select_finalize<DT, LARGEST><<<g_select, SEL_NT, dyn, s>>>(pm);
    if (pm.merge) launch_merge(pm);
  }
  if (L0.part == PART_MAIN) return hipGetLastError();
  // rows entering the exact passes are initialised by the kernel that sends
  // them there; only an exact call starts them all here
  if (exact_call) k_exact_init<<<g_init, 256, 0, s>>>(p1);
  k_exact<DT, LARGEST, U><<<xgrid, STREAM_NT, 0, s>>>(p1);
  // COPYALL / SMALL rows of an exact-path call are finalized here too.  An
  // exact call's large selections are sorted by the merge kernels, as in a
  // sampled call's first round (a sampled call's rare fallback rows with
  // k > MERGE_CHUNK sort in their select's workgroup, so that its fallback
  // round stays two launches).
  KParams pf = exact_call ? p : p1;
  pf.merge = exact_call && merge_chunks(p) > 1;
  k_select_finalize<DT, LARGEST, true><<<fb_sel, SEL_NT, dyn, s>>>(pf);
  if (pf.merge) launch_merge(pf);
  return hipGetLastError();
}

template <int DT, bool LARGEST>
static hipError_t run_typed(const KParams& p, const Launch& L) {
  static std::once_flag once;
  std::call_once(once, [&] { set_lds_attrs<DT, LARGEST>((size_t)LDS_KEYS_MAX * sizeof(uint64_t)); });
  switch (p.tune.unroll) {
    case 2: return run_typed_u<DT, LARGEST, 2>(p, L);
    case 8: return run_typed_u<DT, LARGEST, 8>(p, L);
    default: return run_typed_u<DT, LARGEST, 4>(p, L);
  }
}

// Entry used by the C-ABI layer.
hipError_t run_pipeline(int dtype, const KParams& p, hipStream_t stream, int row_grid, int grid,
                        int part, uint32_t* publish) {
  if (part != PART_ALL && (p.flags & FL_EXACT)) return hipErrorInvalidValue;
  Launch L{stream, row_grid, grid, part, publish};
  const bool largest = p.flags & FL_LARGEST;
  switch (dtype) {
    case F32: return largest ? run_typed<F32, true>(p, L) : run_typed<F32, false>(p, L);
    case F16: return largest ? run_typed<F16, true>(p, L) : run_typed<F16, false>(p, L);
    case BF16: return largest ? run_typed<BF16, true>(p, L) : run_typed<BF16, false>(p, L);
    default: return hipErrorInvalidValue;
  }
}

int lds_keys_max() { return LDS_KEYS_MAX; }

// Copies the first 32 bytes of a workspace header (status, totals, fallback
// and ordered row counts) to `dst`, which may be mapped host memory: a
// stream-ordered way for a host loop to learn a call's outcome from the
// completion of this kernel alone, without a copy engine.
__global__ __launch_bounds__(64) void k_publish(const uint32_t* hdr, uint32_t* dst) {
  if (threadIdx.x < 8) dst[threadIdx.x] = hdr[threadIdx.x];
  __threadfence_system();
}

hipError_t publish_header(const void* workspace, uint32_t* dst, hipStream_t stream) {
  k_publish<<<1, 64, 0, stream>>>(static_cast<const uint32_t*>(workspace), dst);
  return hipGetLastError();
}

int timing_enable(int on) {
  TimingState& t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  t.on = on != 0;
  return 0;
}

// Waits for the recorded events; returns the summed milliseconds and count.
int timing_read(double* ms, int32_t* launches) {
  TimingState& t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  double total = 0.0;
  int32_t n = 0;
  for (auto& ev : t.pending) {
    if (hipEventSynchronize(ev.second) != hipSuccess) return -1;
    float f = 0.0f;
    if (hipEventElapsedTime(&f, ev.first, ev.second) != hipSuccess) return -1;
    total += f;
    ++n;
    t.pool.push_back(ev);
  }
  t.pending.clear();
  if (ms) *ms = total;
  if (launches) *launches = n;
  return 0;
}

}  // namespace rk
