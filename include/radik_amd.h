/*
 * radik_amd — C ABI of the MI355X (gfx950) radix top-k.
 *
 * Every entry point takes plain pointers and sizes; device pointers are HIP
 * device memory; `stream` is a hipStream_t (NULL = the null stream).  No
 * call allocates device memory, synchronises the host, or throws: errors are
 * returned as RK_E_* codes.  All row metadata of the batched entry points is
 * read on the device, so a call can be captured into a hipGraph and replayed.
 *
 * Semantics (identical for every entry point):
 *   - a row's selection key is the order-preserving flip of its bit pattern
 *     (radik/RadixSelect/utils.cuh:329-334); NaN ranks below every number in
 *     both directions;
 *   - the min(len, k) best elements are selected, ties at the k-th key broken
 *     by the lower row position (the reference's filter admits ties by an
 *     atomic race, radik/RadixSelect/radixselect_l.cuh:417-421);
 *   - RK_SORTED: ordered by key in the RK_ASCEND / descending direction,
 *     equal keys by position ascending; otherwise in position order;
 *   - indices are row-local positions, or idx_in[row_start + position];
 *   - output slots beyond min(len, k) are filled with value 0 / index -1
 *     (radik/RadixSelect/topk_radixselect.h:312-320).
 */
#ifndef RADIK_AMD_H_
#define RADIK_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types */
enum { RK_F32 = 0, RK_F16 = 1, RK_BF16 = 2 };

/* flags: RK_LARGEST mirrors the LARGEST template flag, RK_SORTED the
 * needSorting argument and RK_ASCEND the ASCEND template flag of
 * topKRadixSelectL (radik/RadixSelect/topk_radixselect.h:40-70).
 * RK_EXACT forces RadiK's multi-pass histogram path (no sampled threshold):
 * the "plain" kernel the batched path is measured against. */
enum { RK_LARGEST = 1u, RK_SORTED = 2u, RK_ASCEND = 4u, RK_EXACT = 8u };

/* return codes */
enum {
  RK_OK = 0,
  RK_E_ARG = -1,         /* bad argument (null pointer, negative size, ...) */
  RK_E_WORKSPACE = -2,   /* workspace missing or smaller than required */
  RK_E_UNSUPPORTED = -3, /* sorted output with k above rk_max_k(), or dtype unknown */
  RK_E_ALIGN = -4,       /* value pointer not aligned to its element size */
  RK_E_HIP = -5,         /* a HIP runtime call failed */
  RK_E_DEVICE = -6,      /* device-side validation failed (rk_topk_status) */
  RK_E_IO = -7,          /* batch / result file missing, unreadable or malformed */
  RK_E_SCRATCH = -8      /* call too large for rk_topk_batched_noalloc's scratch */
};

/* Largest k for sorted output on the current device (16384 on MI355X: the
 * LDS sort).  Unsorted calls accept any k: rows whose k exceeds this limit
 * (e.g. a median, k = len / 2) are written by an ordered emit in position
 * order, after a sampled bracket of the k-th key (two reads of the row) or,
 * on the exact path and for rows the bracket misses, the digit passes. */
int32_t rk_max_k(void);

/* Workspace bytes for a batched call over at most `num_rows` rows with at
 * most `total_len` elements in total and every k <= max_k.
 * Replaces getRadixSelectLWorkSpaceSize (topk_radixselect.h:22-33), which
 * sizes by num_rows * longest_row; this sizes by total length and bounded
 * candidate slots, so uneven batches need no num_rows * max_len buffer. */
int rk_topk_workspace_size(int dtype, int32_t num_rows, int64_t total_len,
                           int32_t max_k, uint32_t flags, size_t* bytes);

/* Batched top-k over uneven rows; all metadata on the device.
 * Replaces topKRadixSelectL (topk_radixselect.h:59-323) for batches:
 *   vals      device, concatenated rows (valIn)
 *   row_off   device int64[num_rows + 1]: row r is vals[row_off[r], row_off[r+1])
 *             (the reference builds int offsets from a host taskLen[], :90-98)
 *   row_k     device int32[num_rows]: per-row k (the reference has one K, :106-112)
 *   total_len bound on row_off[num_rows] - row_off[0]; max_k bound on every k
 *   idx_in    optional device int32, parallel to vals (idxIn / WITHIDXIN)
 *   out_vals, out_idx  device; row r writes [out_off[r], out_off[r+1])
 *   out_off   device int64[num_rows + 1]; each row needs >= min(len, k) slots
 *   workspace device, >= rk_topk_workspace_size(...) bytes
 * Stream ordered, asynchronous, graph-capturable. */
int rk_topk_batched(int dtype, const void* vals, const int64_t* row_off,
                    const int32_t* row_k, int32_t num_rows, int64_t total_len,
                    int32_t max_k, uint32_t flags, const int32_t* idx_in,
                    void* out_vals, int32_t* out_idx, const int64_t* out_off,
                    void* workspace, size_t ws_bytes, void* stream);

/* rk_topk_batched with 64-bit indices: idx_in (optional) and out_idx are
 * int64_t, the reference's IdxType = int64_t instantiation of
 * topKRadixSelectL<IdxType, ...>(valIn, idxIn, valOut, idxOut, ...)
 * (topk_radixselect.h:59-70).  Row lengths stay below 2^31 (positions
 * written without idx_in are row-local); the tail fill index is -1.  Same
 * workspace as rk_topk_batched. */
int rk_topk_batched_i64(int dtype, const void* vals, const int64_t* row_off,
                        const int32_t* row_k, int32_t num_rows, int64_t total_len,
                        int32_t max_k, uint32_t flags, const int64_t* idx_in,
                        void* out_vals, int64_t* out_idx, const int64_t* out_off,
                        void* workspace, size_t ws_bytes, void* stream);

/* Allocation-free form of rk_topk_batched.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory or hand in
 * a workspace (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and results as rk_topk_batched,
 * without the workspace: the call's scratch is a device array of
 * RK_NOALLOC_SCRATCH_BYTES built into the library, one per GPU, held from the
 * library's load on that GPU until the process ends.  A call whose
 * rk_topk_workspace_size exceeds it returns RK_E_SCRATCH before any device
 * work.  The call allocates nothing and is graph-capturable.  All calls on
 * one GPU share that array: calls that overlap on one GPU (two streams or two
 * host threads without ordering between them) race, so serialize them.  The
 * device status of these calls is not exposed; a row the device rejects
 * (k above max_k, bad offsets) is left unwritten. */
#define RK_NOALLOC_SCRATCH_BYTES (64u << 20)
int rk_topk_batched_noalloc(int dtype, const void* vals, const int64_t* row_off,
                            const int32_t* row_k, int32_t num_rows, int64_t total_len,
                            int32_t max_k, uint32_t flags, const int32_t* idx_in,
                            void* out_vals, int32_t* out_idx, const int64_t* out_off,
                            void* stream);

/* Same, with each row and its output at their own device addresses
 * (row_ptrs, row_len, row_k, out_val_ptrs and out_idx_ptrs are device
 * arrays; row r writes row_k[r] slots at out_val_ptrs[r] / out_idx_ptrs[r]):
 * the form a serving loop uses for requests resident in their own buffers.
 * No counterpart in the reference (it takes one contiguous valIn). */
int rk_topk_rows(int dtype, const void* const* row_ptrs, const int64_t* row_len,
                 const int32_t* row_k, int32_t num_rows, int64_t total_len,
                 int32_t max_k, uint32_t flags, void* const* out_val_ptrs,
                 int32_t* const* out_idx_ptrs, void* workspace, size_t ws_bytes,
                 void* stream);

/* Workspace bytes for rk_topk_compat: same arguments as
 * getRadixSelectLWorkSpaceSize<ValType>(K, n, taskNum, &bytes)
 * (topk_radixselect.h:22-33), n = the longest row. */
int rk_compat_workspace_size(int dtype, int32_t K, int32_t n, int32_t taskNum,
                             uint32_t flags, size_t* bytes);

/* The reference's surface: host taskLen[], one K, output [taskNum][K] with
 * the tail fill; mirrors topKRadixSelectL(valIn, idxIn, valOut, idxOut,
 * workSpace, taskLen, taskNum, K, stream, needSorting)
 * (topk_radixselect.h:59-70).  Copies taskNum+1 offsets to the device from
 * host memory, so unlike rk_topk_batched it is not graph-capturable. */
int rk_topk_compat(int dtype, const void* valIn, const int32_t* idxIn,
                   void* valOut, int32_t* idxOut, void* workspace,
                   size_t ws_bytes, const int32_t* taskLen, int32_t taskNum,
                   int32_t K, uint32_t flags, void* stream);

/* rk_topk_compat with IdxType = int64_t: idxIn (optional) and idxOut are
 * int64_t (topk_radixselect.h:59-70 instantiated with int64_t indices). */
int rk_topk_compat_i64(int dtype, const void* valIn, const int64_t* idxIn,
                       void* valOut, int64_t* idxOut, void* workspace,
                       size_t ws_bytes, const int32_t* taskLen, int32_t taskNum,
                       int32_t K, uint32_t flags, void* stream);

/* Waits for `stream`, then reads the device status of the last call that
 * used `workspace`: 0 when every row was valid.  Optional outputs: rows
 * finished by the exact-pass executor (rows that left the sampled path, and
 * rows with k above rk_max_k()), and rows that needed the ordered emit.  No
 * counterpart in the reference (it throws, error_check.hpp:14-58). */
int rk_topk_status(const void* workspace, void* stream, uint32_t* status,
                   uint32_t* fallback_rows, uint32_t* ordered_rows);

/* Human-readable text of a return code. */
const char* rk_strerror(int code);

/* Tunables of the batched path: tile size (keys per streaming work unit,
 * default 16384, smaller for small calls), persistent grid of the streaming
 * kernels (default 4 workgroups per CU for the sampled path's filter; 1, or
 * 2 from 2^26 keys, for an RK_EXACT call's executor), sample runs and keys
 * per run (1024 x 16), digit width of the exact passes (8..11 bits, default
 * 11: 3 passes over 32-bit keys) and 16-byte loads in flight per thread (2,
 * 4 or 8; default 2, and 4 for RK_EXACT calls).  A field of 0 selects its
 * default. */
typedef struct rk_tune {
  int32_t tile;
  int32_t grid;
  int32_t sample_groups;
  int32_t group_elems;
  int32_t hist_bits;
  int32_t unroll;
} rk_tune;

/* Reads / overrides the tunables used by subsequent calls in this process
 * whose shape bucket has no autotuned entry.  rk_set_tune leaves fields < 0
 * unchanged; rk_get_tune reports 0 for a field at its default. */
int rk_get_tune(rk_tune* out);
int rk_set_tune(const rk_tune* in);

/* Profiling hook (no counterpart in the reference): while enabled, every
 * batched call records a pair of HIP events around its streaming filter
 * kernel on the call's stream.  rk_timing_read waits for them, returns the
 * summed kernel milliseconds and the number of launches timed, and clears. */
int rk_timing_enable(int on);
int rk_timing_read(double* filter_ms, int32_t* launches);

/* Autotuner (no counterpart in the reference, whose grid constants are fixed
 * for one GPU, topk_radixselect.h:17): times the batched call on the given
 * buffers over a grid of tunables and caches the fastest for the shape bucket
 * (dtype, RK_EXACT, ceil log2 of total_len, num_rows and max_k); later batched
 * calls in that bucket use it.  Returns the winner and its median time. */
int rk_autotune_batched(int dtype, const void* vals, const int64_t* row_off,
                        const int32_t* row_k, int32_t num_rows, int64_t total_len,
                        int32_t max_k, uint32_t flags, void* out_vals,
                        int32_t* out_idx, const int64_t* out_off, void* workspace,
                        size_t ws_bytes, void* stream, int32_t reps,
                        rk_tune* best, float* best_ms);
int rk_tune_cache_save(const char* path);
int rk_tune_cache_load(const char* path);
int rk_tune_cache_clear(void);
int rk_tune_cache_size(void);

/* hipGraph capture of any sequence of radik_amd calls issued on `stream`
 * between rk_graph_begin and rk_graph_end (the batched entry points make no
 * host round trip and no memset, so they replay correctly); replaying reads
 * the buffers' current contents. */
typedef struct rk_graph rk_graph;
int rk_graph_begin(void* stream);
int rk_graph_end(void* stream, rk_graph** out);
int rk_graph_launch(rk_graph* graph, void* stream);
int rk_graph_destroy(rk_graph* graph);

/* ---------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference, whose harness launches one
 * batch per call, radik/test/test_radik.cu:135-160).  A server owns `streams`
 * slots, each a stream with its own workspace, sized at creation for
 * max_batch_rows rows of at most max_batch_keys keys in total; creation is
 * the only call that allocates device memory.  A dispatcher thread admits
 * pending requests in arrival order while the bytes of the requests in
 * flight (input + output) plus the server's own buffers stay within
 * mem_budget, coalesces them into one rows-form call per free slot, and
 * completes them when the slot's event fires.  A request whose bytes alone
 * exceed the budget is rejected (RK_E_WORKSPACE from rk_server_wait).
 * plain = 1 is the baseline: one request per RK_EXACT call, one stream.
 * Results are identical to rk_topk_rows on the request alone.
 * ------------------------------------------------------------------------- */
typedef struct rk_server rk_server;

typedef struct rk_server_config {
  int64_t max_batch_keys; /* keys per batch; also the longest request */
  int64_t mem_budget;     /* device bytes: server buffers + requests in flight */
  int32_t dtype;
  uint32_t flags;         /* RK_LARGEST / RK_SORTED / RK_ASCEND of every request */
  int32_t max_k;          /* largest k of any request */
  int32_t max_batch_rows; /* requests per batch */
  int32_t streams;        /* batches in flight at once */
  int32_t plain;
} rk_server_config;

typedef struct rk_server_stats {
  int64_t completed, rejected, failed;
  int64_t batches, fallback_batches; /* fallback: batches that ran exact passes */
  int64_t server_bytes, peak_inflight_bytes;
  double mean_batch_rows;
  double p50_us, p99_us, mean_us, max_us; /* latency: arrival to observed completion */
  double issue_us;                        /* mean host time to issue one batch part */
  int32_t max_batch_rows;
  int32_t pad_;
  int64_t retained; /* requests held: queued, in flight, or done and not yet waited for */
} rk_server_stats;

int rk_server_create(const rk_server_config* cfg, rk_server** out);
/* Queues one request: a device row of `len` elements, k <= cfg.max_k; writes
 * min(len, k) results and the tail fill up to k slots into out_vals /
 * out_idx.  The row and outputs must stay valid until the request completes. */
int rk_server_submit(rk_server* s, const void* vals, int64_t len, int32_t k, void* out_vals,
                     int32_t* out_idx, uint64_t* ticket);
/* Blocks until the request completes; returns its status and latency.  A
 * request's outcome is reported once: the server then releases it (memory
 * stays bounded by what is queued, in flight or not yet waited for), and a
 * second wait on the ticket returns RK_E_ARG.  rk_server_replay reports and
 * releases its own requests. */
int rk_server_wait(rk_server* s, uint64_t ticket, double* latency_us);
/* Blocks until nothing is queued or in flight. */
int rk_server_drain(rk_server* s);
/* Open-loop load: submits request i at arrival_s[i] seconds after the call
 * starts (from the calling thread), waits for all, and returns per-request
 * latency from the scheduled arrival (-1 if not completed) and the makespan.
 * Returns RK_E_DEVICE if a request failed device-side validation; rejected
 * requests only show as -1. */
int rk_server_replay(rk_server* s, int32_t n, const void* const* vals, const int64_t* lens,
                     const int32_t* ks, void* const* out_vals, int32_t* const* out_idx,
                     const double* arrival_s, double* latency_us, double* makespan_s);
/* Statistics over every request since creation (nearest-rank percentiles). */
int rk_server_get_stats(rk_server* s, rk_server_stats* out);
/* Drains, stops the dispatcher and frees the server's buffers. */
int rk_server_destroy(rk_server* s);

/* ---- Fused k-nearest-neighbour select -----------------------------------
 * No counterpart in the reference, which selects from values already in
 * memory: this is its top-k fused into the producer of the values, a
 * distance GEMM (radik/RadixSelect/topk_radixselect.h:59-70 is the select it
 * extends).  Scores of nq queries against n database vectors are computed
 * with MFMA (bf16 or f16 inputs, fp32 accumulation) and filtered against a
 * per-query threshold inside the same kernel, so the nq x n score matrix is
 * never written; the candidates are selected exactly (ties: lower database
 * index).
 *   queries [nq][d], base [n][d]: row-major, 16-byte aligned, dtype RK_BF16
 *   or RK_F16, d a multiple of 16 in [16, 256], 1 <= k <= min(n, 16384).
 *   metric RK_KNN_IP: inner product, largest first; RK_KNN_L2: squared L2
 *   distance |q|^2 - (2 q.x - |x|^2), smallest first.
 *   flags: RK_SORTED (best first, ties by index) or 0 (database order).
 *   out_vals [nq][k] fp32 scores / distances, out_idx [nq][k] int32 indices.
 * rk_topk_status(workspace) afterwards reports RK_E_DEVICE if a query's
 * candidates at the safe threshold outgrew its slots: the safe threshold is
 * the k-th best score of S evenly spaced database vectors (every floor(n/S)-th),
 * so this happens when very many scores tie at it, or when the database's
 * layout repeats with that stride so that the sampled vectors score far
 * below the query's true k-th best.  Such a query's outputs are filled with
 * value NaN and index -1.  The fallback_rows output counts the queries that
 * needed the second, safe-threshold pass. */
enum { RK_KNN_IP = 0, RK_KNN_L2 = 1 };
int rk_knn_workspace_size(int dtype, int32_t nq, int64_t n, int32_t d, int32_t k, size_t* bytes);
int rk_knn_topk(int dtype, int metric, const void* queries, const void* base, int32_t nq, int64_t n,
                int32_t d, int32_t k, uint32_t flags, float* out_vals, int32_t* out_idx,
                void* workspace, size_t ws_bytes, void* stream);
/* The scores themselves, [nq][n] fp32, with the same arithmetic as
 * rk_knn_topk (the unfused path: scores to memory, then rk_topk_batched). */
int rk_knn_scores(int dtype, int metric, const void* queries, const void* base, int32_t nq,
                  int64_t n, int32_t d, float* out, void* stream);

/* ---- CSR batch and result files -----------------------------------------
 * No counterpart in the reference, whose harness builds its rows in memory
 * (radik/test/test_radik.cu:88-127): a little-endian file form of the
 * batched entry's CSR layout (rk_topk_batched), for capturing a batch of
 * requests and replaying it, and for storing its results.
 *   batch file:  64-byte header (magic "RKBATCH1", dtype, flags, rows,
 *                total), row_off int64[rows+1] (row_off[0] = 0, nondecreasing,
 *                row_off[rows] = total), row_k int32[rows] padded to 8
 *                bytes, values[total] (4 bytes for RK_F32, 2 for RK_F16 /
 *                RK_BF16, as bits).
 *   result file: 64-byte header (magic "RKRESLT1", dtype, 0, rows, outputs),
 *                out_off int64[rows+1] (same rules, ending at outputs),
 *                out_vals[outputs], out_idx int32[outputs].
 * Buffers are host memory; no device work.  The *_info calls read the
 * header (dtype; total = values of a batch, outputs of a result) so the
 * caller can size the buffers of the matching *_read call.  That call takes
 * the dtype and sizes back (max_rows, max_total: what the buffers hold) and
 * rejects a file that has since changed its element type or grown past them
 * (RK_E_IO) before writing anything; it validates the offsets and the
 * file's exact length.  Errors: RK_E_ARG (arguments), RK_E_UNSUPPORTED
 * (dtype), RK_E_IO (the file). */
typedef struct {
  int32_t dtype;
  uint32_t flags;  /* batch: the RK_* flags the batch was captured with */
  int64_t rows;
  int64_t total;
} rk_file_info;
int rk_batch_write(const char* path, int dtype, uint32_t flags, int32_t rows,
                   const int64_t* row_off, const int32_t* row_k, const void* values);
int rk_batch_info(const char* path, rk_file_info* info);
int rk_batch_read(const char* path, int dtype, int64_t max_rows, int64_t max_total,
                  int64_t* row_off, int32_t* row_k, void* values);
int rk_result_write(const char* path, int dtype, int32_t rows, const int64_t* out_off,
                    const void* out_vals, const int32_t* out_idx);
int rk_result_info(const char* path, rk_file_info* info);
int rk_result_read(const char* path, int dtype, int64_t max_rows, int64_t max_total,
                   int64_t* out_off, void* out_vals, int32_t* out_idx);

#ifdef __cplusplus
}
#endif

#endif /* RADIK_AMD_H_ */
