// Shared host/device layout of the top-k pipeline: row descriptors, row
// states, the workspace header and the kernel parameter block.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace rk {

// Row phases.  Each kernel acts only on rows in its phase, so the launch
// sequence is fixed and needs no host round trip.
enum Phase : uint32_t {
  PH_DONE = 0,     // output written
  PH_COPYALL = 1,  // len <= k: the whole row is the answer
  PH_SMALL = 2,    // row fits in LDS: select directly from the input
  PH_SAMPLE = 3,   // long row: estimate a threshold from a sample
  PH_FILTER = 4,   // threshold T known: compact candidates >= T, then select
  PH_HIST = 5,     // exact radix passes over the row (fallback / plain path)
  PH_ORDERED = 6,  // k-th key V known but its ties overflow the candidate
                   // buffer: ordered emit in row-position order
  PH_MERGE = 7,    // k > MERGE_CHUNK selected keys (order-keyed) in the row's
                   // candidate slots: sorted in chunks, then ranked and written
  PH_BRACKET = 8,  // k above the LDS select (unsorted): the sample brackets the
                   // k-th key in [T, V]; one read counts the keys above V per
                   // chunk and compacts those in [T, V], the k-th key is
                   // selected among them, then the ordered emit writes
};

// Keys per chunk of the multi-workgroup sort of large selections.
constexpr int MERGE_CHUNK = 4096;

enum RowFlag : uint32_t {
  RF_NEED_INIT = 1u,  // PH_HIST row whose histogram/state must be reset
  RF_EXACT_T = 2u,    // T came from exact histograms: the select cannot fail
  RF_EMIT_WRITE = 4u, // PH_ORDERED row past its counting pass: writing
};

enum Status : uint32_t {
  ST_BAD_ROW = 1u,       // negative length / k, k above max_k, output too small
  ST_TILES = 2u,         // tile list exceeds the workspace bound
  ST_CANDIDATES = 4u,    // candidate capacity exceeds the workspace bound
  ST_INTERNAL = 8u,      // invariant broken (should never be set)
};

struct RowDesc {
  const void* base;    // first element of the row
  const void* idx_in;  // optional caller indices, parallel to base (int32 or int64)
  void* out_v;         // first output value slot of the row
  void* out_i;         // first output index slot of the row (int32 or int64)
  int64_t len;
  int32_t k;
  int32_t slots;  // output slots reserved for the row (>= min(len, k))
};

struct RowState {
  uint32_t phase;
  uint32_t flags;
  uint32_t T;         // PH_FILTER threshold key
  uint32_t prefix;    // PH_HIST: resolved high key bits
  uint32_t resolved;  // PH_HIST: number of resolved bits
  uint32_t rem;       // PH_HIST / PH_ORDERED: elements still to take at the boundary
  uint32_t V;         // PH_ORDERED: the k-th key
  uint32_t cap;       // candidate capacity of the row
  uint32_t why;       // why the row left the sampled path (diagnostics, Fallback)
};

// Per-row work state of the exact-pass executor (k_exact).  A row's passes
// (digit histograms, then the filter or the ordered emit's two passes) are
// stages of nch chunks each; `claim` hands the chunks of the current stage
// out and carries the stage's sequence number, so a workgroup that claims a
// chunk knows which stage it belongs to.
struct XRow {
  uint64_t claim;  // (stage sequence << 32) | next chunk of the stage
  uint32_t done;   // chunks of the current stage completed
  uint32_t nch;    // chunks per stage
  uint32_t ct;     // tiles per chunk
  uint32_t kor;    // histogram pass: OR and AND of the keys it counted, so
  uint32_t kand;   //   that the next window can skip the bits they all share
  uint32_t pad;
};
constexpr uint32_t XSEQ_DONE = 0xffffffffu;  // claim's sequence once the row has left

enum Fallback : uint32_t {
  FB_NONE = 0,
  FB_SAMPLE_DEGENERATE = 1,  // threshold would admit every key (NaN-heavy / rank >= sample)
  FB_SAMPLE_TOO_MANY = 2,    // estimated candidates exceed the row's slots (heavy ties)
  FB_OVERFLOW = 3,           // more candidates than slots arrived
  FB_TOO_FEW = 4,            // fewer than k candidates were >= T (unrepresentative sample)
  FB_BRACKET = 5,            // the k-th key fell outside its sampled bracket (or the
                             // bracket overflowed its slots): digit passes instead
};

struct Header {
  uint32_t status;
  uint32_t total_tiles;
  uint64_t total_cand;
  uint32_t fallback_rows;  // rows that left the sampled path for exact passes
  uint32_t ordered_rows;   // rows that needed the ordered emit
  uint32_t q_sample;       // row queue heads of the per-row kernels
  uint32_t q_select[2];    //   (select: sampled round, fallback round)
  uint32_t sel_done;       // workgroups of the sampled-round select that finished
  uint32_t q_chunk[2];     // chunk queue heads of the filter (sampled, fallback round)
  uint32_t x_rows;         // rows handed to the exact-pass executor (xrows entries)
  uint32_t x_left;         // of those, rows whose passes are not finished
  uint32_t pad[50];
};
static_assert(sizeof(Header) == 256, "workspace header is 256 bytes");

// Tunables chosen by the autotuner per shape bucket.
struct Tune {
  int32_t tile;           // elements per tile (multiple of 64)
  int32_t grid;           // workgroups of the persistent streaming kernels
  int32_t sample_groups;  // groups of contiguous elements sampled per row
  int32_t group_elems;    // elements per sample group (multiple of 8)
  int32_t hist_bits;      // digit width of the exact passes (8..11)
  int32_t unroll;         // 16-byte loads in flight per lane (2, 4 or 8)
};

struct KParams {
  // inputs
  const void* vals;
  const int64_t* row_off;
  const int32_t* row_k;
  const int64_t* out_off;
  const void* idx_in;           // int32, or int64 when idx64
  const void* const* row_ptrs;  // pointer form (serving)
  const int64_t* row_len;       // pointer form (serving)
  void* const* out_val_ptrs;    // pointer form: per-row output values
  int32_t* const* out_idx_ptrs; // pointer form: per-row output indices
  void* out_vals;
  void* out_idx;  // int32, or int64 when idx64
  int32_t idx64;  // caller indices and output indices are int64 (IdxType = int64_t)
  int32_t K;      // k for every row when row_k is null; slot stride when out_off is null
  int32_t R;      // rows
  int32_t max_k;
  uint32_t flags;
  int64_t total_len;  // bound on sum of lengths (tile / candidate sizing)
  // workspace
  Header* hdr;
  RowDesc* desc;
  RowState* st;
  uint32_t* tile_start;  // [R + 1]
  uint64_t* cand_base;   // [R + 1]
  uint32_t* cand_cnt;    // [R]
  uint32_t* order;       // [2R] rows by descending cost class: largest k first
                         // (select), then longest first (sample)
  uint32_t* ghist;       // [R][2048]
  uint2* tile_cnt;       // [tile_capacity] ordered emit: keys > V, == V per tile
  uint32_t* chunk_row;   // [tile_capacity] row of each filter chunk, or ~0u
  uint64_t* cand;        // [cand_capacity]
  uint32_t* xrows;       // [R] rows of the exact-pass executor, in arrival order
  XRow* xst;             // [R] their work state
  uint64_t tile_capacity;
  uint64_t cand_capacity;
  int32_t lds_keys;     // LDS key capacity of the select kernels (this launch)
  int32_t small_max;    // rows up to this length select straight from the input
  int32_t sample_keys;  // LDS key capacity of the sample kernel (this launch)
  int32_t round;        // 0: sampled round; 1: fallback round (skips when idle)
  int32_t chunk_tiles;  // tiles per chunk of the interleaved filter partition
  int32_t merge;        // select: sort rows with k > MERGE_CHUNK in chunks
                        // (the merge kernels follow this select launch)
  int32_t x_chunk_tiles;  // most tiles per chunk of the exact-pass executor
  Tune tune;
  // When set (sampled-round select only): the last select workgroup to finish
  // copies the header's first 8 words here (mapped host memory).
  uint32_t* publish;
};

// Public flag bits (mirrors of the C-ABI values).
enum : uint32_t {
  FL_LARGEST = 1u,
  FL_SORTED = 2u,
  FL_ASCEND = 4u,
  FL_EXACT = 8u,
};

// Workgroups per CU of the exact-pass executor by default: 1, or 2 for calls
// of 2^26 keys and more (measured on MI355X: fewer workgroups contend less
// for the device-scope atomics every chunk pays; more stream faster once the
// work is large).
inline int exact_grid_per_cu(int64_t total_len) { return total_len >= (int64_t(1) << 26) ? 2 : 1; }

constexpr int HIST_MAX_BINS = 2048;
constexpr int TILE_MIN = 1024;

// Which part of the launch sequence a call issues.  PART_MAIN is the sampled
// round (plan, sample, filter, select_finalize); PART_FALLBACK the exact
// passes and ordered emit that only do work when rows fell back.  A caller
// that reads the header between the two (the serving loop) skips the idle
// second half; PART_ALL issues both back to back.
enum Part : int { PART_ALL = 0, PART_MAIN = 1, PART_FALLBACK = 2 };

}  // namespace rk
