// Fused k-nearest-neighbour select on MI355X: the top-k of every query's
// scores against a database, with the scores computed by MFMA (bf16 or f16
// inputs, fp32 accumulation) and filtered against a per-query threshold in
// the same kernel, so that the [queries x database] score matrix never goes
// to HBM.  The surviving candidates are selected exactly with the rules of
// the batched top-k (composite key = (selection key << 32) | ~index, so ties
// go to the lower database index; sorted output best first).
//
// Launch sequence of one call (device-side decisions only, no host sync, no
// memset: graph-capturable):
//   1. k_knn<SCORES>   scores against S evenly spaced database vectors
//   2. k_knn_threshold per query, over its S sample keys: T_safe = the k-th
//                      best (at least k database scores reach it, since the
//                      sample is a subset) and T = the r-th best with
//                      r = ceil(mu + 4 sqrt(mu) + 2), mu = S k / n (about r n / S
//                      candidates instead of k n / S)
//   (L2 only: k_knn_xnorm, the squared norm of every database vector)
//   3. k_knn<FILTER>   all n scores; keys >= T appended as candidates
//   4. k_knn_select    exact top-k of the candidates; a query with fewer than
//                      k candidates (or more than its slots) is queued for a
//                      retry with T_safe
//   5. k_knn<FILTER>   the queued queries only (returns at once if none)
//   6. k_knn_select    the queued queries; still too many candidates (only
//                      with massive ties at T_safe) sets the status word
//
// Scores: inner product q.x (largest first), or squared L2 distance
// |q|^2 - (2 q.x - |x|^2) (smallest first), in fp32 with a fixed operation
// order shared by the fused and the materialising (rk_knn_scores) paths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "../../include/radik_amd.h"
#include "rk_device.hpp"
#include "rk_pipeline.hpp"

namespace rk {
namespace {

constexpr int KNN_NT = 256;       // 4 waves; each owns 32 queries (or 64)
constexpr int KNN_DMAX = 256;     // query fragments are held in registers
constexpr int KNN_SAMPLE = 16384; // sampled database vectors (at most)
constexpr int KNN_SEL_NT = 512;   // threads of the per-query kernels
constexpr int KNN_STAGE = 256;    // staged candidates per wave (filter)
constexpr int KNN_K_MAX = 16384;  // k bound (LDS sort of the selection)

enum KnnMode : int { MODE_SCORES = 0, MODE_FILTER = 1 };

typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int DT>
struct Mfma;
template <>
struct Mfma<BF16> {
  typedef __bf16 v8 __attribute__((ext_vector_type(8)));
  static RK_DEV f32x16 run(const uint4& a, const uint4& b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8, a),
                                                   __builtin_bit_cast(v8, b), c, 0, 0, 0);
  }
};
template <>
struct Mfma<F16> {
  typedef _Float16 v8 __attribute__((ext_vector_type(8)));
  static RK_DEV f32x16 run(const uint4& a, const uint4& b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8, a),
                                                  __builtin_bit_cast(v8, b), c, 0, 0, 0);
  }
};

// Sum of squares of the 8 elements of a 16-byte fragment, in element order.
template <int DT>
RK_DEV float sq8(const uint4& x, float acc) {
  const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const float f = bits_to_float<DT>((w[i] >> (16 * h)) & 0xffffu);
      acc = __fadd_rn(acc, __fmul_rn(f, f));
    }
  }
  return acc;
}

struct KnnArgs {
  const void* q;         // queries [nq][d]
  const void* x;         // database vectors, vector j at x + j * ldx elements
  int32_t nq;            // queries (list entries when `list` is set)
  int32_t d;
  int64_t nv;            // database vectors of this launch
  int64_t ldx;           // elements between consecutive vectors
  int32_t metric;        // RK_KNN_IP / RK_KNN_L2
  int32_t chunk_v;       // vectors per workgroup (multiple of 32)
  int32_t qblocks;       // ceil(nq / queries per workgroup)
  int32_t chunks;        // ceil(nv / chunk_v)
  // scores mode
  float* out;            // [nq][ldo]
  int64_t ldo;
  // filter mode
  const uint32_t* list;  // query of list entry e (null: e itself)
  const uint32_t* nlist; // device count of list entries (null: nq)
  const uint32_t* T;     // per-query threshold key
  uint32_t* cnt;         // per-query candidate counters
  uint64_t* cand;        // per-query candidate slots (cap each)
  uint32_t cap;
  const float* xnorm;    // squared norms of the launch's vectors (L2 filter
                         // mode; scores mode computes them from the B fragments)
};

// Score of one accumulator element: inner product, or the squared L2
// distance |q|^2 - (2 q.x - |x|^2) (the same operations in every path).
// 2 q.x - |x|^2 is one fused multiply-add: 2 q.x is exact in fp32 (short of
// overflow), so it equals the product-then-difference rounded once.
RK_DEV float knn_value(int metric, float dot, float qn, float xn) {
  if (metric == RK_KNN_IP) return dot;
  return __fsub_rn(qn, __fmaf_rn(2.0f, dot, -xn));
}

// MFMA tiles: each wave holds MT tiles of 32 queries and multiplies them by
// 32 database vectors per step (v_mfma_f32_32x32x16, K = 16 per
// instruction), so one B fragment feeds MT instructions.  Lane l holds A row
// l % 32 and B column l % 32 at k = 16 t + 8 (l / 32) .. +7, and accumulator
// i of row 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32 (the layout was
// checked on MI355X).
//
// Filter mode tests every score with one float compare against its query's
// threshold, held in registers: !(v < tf) for largest-first (!(v > tf) for
// smallest-first) is true for every score whose key reaches T (tf is the
// float of T, NaN when T = 0, so that everything passes), for NaN scores, and
// for +-0 next to a zero threshold.  The compares go in groups of four
// accumulators with one wave-uniform branch per group: only when some lane
// passes one of them is the exact key compared and the candidate staged;
// with thresholds near the k-th best score that is a small fraction of the
// groups.
template <int DT, int KS, int MT, int MODE, int METRIC>
__global__ __launch_bounds__(KNN_NT) __attribute__((amdgpu_waves_per_eu(MODE == MODE_FILTER && KS <= 8 ? 4 : 1))) void k_knn(KnnArgs a) {
  using U = typename Traits<DT>::U;
  constexpr int NW = KNN_NT / 64;
  constexpr int QW = 32 * MT;  // queries per wave
  constexpr bool LARGEST = (METRIC == RK_KNN_IP);
  __shared__ __attribute__((aligned(16))) float sh_qn[NW][QW];
  __shared__ uint32_t sh_t[NW][QW];
  __shared__ uint32_t sh_q[NW][QW];
  __shared__ uint64_t st_key[MODE == MODE_FILTER ? NW : 1][KNN_STAGE];
  __shared__ uint32_t st_q[MODE == MODE_FILTER ? NW : 1][KNN_STAGE];
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  // Workgroup -> (query block, chunk): the query blocks of one chunk run on
  // one XCD (workgroup i goes to XCD i % 8), so its vectors are read from
  // HBM once into that XCD's L2 and shared.
  const int id = blockIdx.x;
  const int xcd = id & 7;
  const int local = id >> 3;
  const int qb = local % a.qblocks;
  const int chunk = (local / a.qblocks) * 8 + xcd;
  if (chunk >= a.chunks) return;
  const int32_t nlist = a.nlist ? (int32_t)*a.nlist : a.nq;
  if (qb * QW * NW >= nlist) return;  // uniform over the workgroup
  const int32_t e0 = (qb * NW + w) * QW;  // first list entry of this wave
  const int nsteps = a.d / 16;
  const U* Q = static_cast<const U*>(a.q);
  const U* X = static_cast<const U*>(a.x);

  // This lane's A rows and their fragments.
  uint4 af[MT][KS];
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    const int32_t ea = e0 + m * 32 + (lane & 31);
    const bool va = ea < nlist;
    const uint32_t qa = va ? (a.list ? a.list[ea] : (uint32_t)ea) : 0u;
    float qpart = 0.0f;
#pragma unroll
    for (int t = 0; t < KS; ++t) {
      af[m][t] = make_uint4(0u, 0u, 0u, 0u);
      if (t < nsteps && va)
        af[m][t] = ldg16<false>(Q + (size_t)qa * a.d + t * 16 + (lane >> 5) * 8);
      qpart = sq8<DT>(af[m][t], qpart);
    }
    const float qn = __fadd_rn(qpart, __shfl_xor(qpart, 32, 64));
    if (lane < 32) {
      sh_qn[w][m * 32 + lane] = qn;
      sh_q[w][m * 32 + lane] = va ? qa : 0xffffffffu;
      sh_t[w][m * 32 + lane] = (MODE == MODE_FILTER && va) ? a.T[qa] : 0u;
    }
  }
  __syncthreads();
  // accumulator i of this lane belongs to row rowof(i) of a 32-query tile
  auto rowof = [&](int i) { return 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3); };
  // per-accumulator query norms (L2; the lean filter reads them from LDS
  // per compare group, like its thresholds)
  float qr[MT][16];
  if constexpr (METRIC == RK_KNN_L2 && MODE != MODE_FILTER) {
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int i = 0; i < 16; ++i) qr[m][i] = sh_qn[w][m * 32 + rowof(i)];
  }
  // Per-row float thresholds and query norms (filter mode) stay in LDS: the
  // four rows of a compare group are consecutive, so a group reads each with
  // one 16-byte load.  With 32-vector tiles and one prefetch set this keeps
  // the filter within 128 registers, 4 waves per SIMD at d <= 128 (measured
  // 2 % faster than 3 for inner product, 8 % for L2).
  constexpr bool LEAN = (MODE == MODE_FILTER);
  __shared__ __attribute__((aligned(16))) float sh_tf[LEAN ? NW : 1][LEAN ? QW : 4];
  float tf[MT][16];
  if constexpr (MODE == MODE_FILTER) {
    auto tf_of = [&](int row) {
      const uint32_t T = sh_t[w][row];
      const float inf = __builtin_inff();
      return sh_q[w][row] == 0xffffffffu ? (LARGEST ? inf : -inf)
           : T == 0u ? __builtin_nanf("")
           : __uint_as_float(key_to_bits<F32, LARGEST>(T));
    };
    if constexpr (LEAN) {
      if (lane < QW) sh_tf[w][lane] = tf_of(lane);
      __syncthreads();
    } else {
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int i = 0; i < 16; ++i) tf[m][i] = tf_of(m * 32 + rowof(i));
    }
  }
  const int64_t v0 = (int64_t)chunk * a.chunk_v;
  const int64_t v1 = std::min<int64_t>(a.nv, v0 + a.chunk_v);
  // B tiles (TV vectors x d) go through LDS once per workgroup, read from
  // this chunk's range with buffer loads (vectors past v1 and chunks past d
  // read zeros without a branch; the host keeps a chunk's bytes below 2^32);
  // every wave reads its B fragments from there.
  //
  // Scores mode: the 256 threads load a tile's 16-byte chunks into
  // registers and store them to one of two LDS buffers.  Rows are padded by
  // 16 bytes, so the 16-byte fragment reads of a wave's lanes (one vector
  // each) fall in distinct banks.
  //
  // Filter mode: the chunks go straight to LDS (bload16_lds: no registers,
  // no LDS store instructions) into a ring of three buffers: step s reads
  // tile s and starts tile s+2 into the buffer step s-1 read.  (The compiler
  // waits for this wave's loads before the next step's LDS reads, so a tile
  // has about one step to arrive.)  Measured on MI355X against loads through
  // registers: filter launch -2.8 % (inner product), -6.6 % (L2).
  // One load instruction of a wave fills 64 consecutive slots, so rows
  // cannot be padded; instead chunk c of tile vector v sits at slot
  // v * CPV + (c ^ swz(v)), which puts the fragment reads of 16 consecutive
  // lanes (16 vectors, one chunk each) in 16 distinct 16-byte bank groups.
  constexpr int TV = 32;                                   // vectors per tile
  constexpr int CPV = KS * 2;                              // chunks per vector
  constexpr int CHUNKS = TV * CPV;                         // chunks per tile
  constexpr int LPT = (CHUNKS + KNN_NT - 1) / KNN_NT;      // chunks per thread
  constexpr int STRIDE = KS * 32 + 16;                     // LDS bytes per vector
  constexpr bool RING = (MODE == MODE_FILTER);
  __shared__ __attribute__((aligned(16)))
      unsigned char btile[RING ? 3 : 2][RING ? CHUNKS * 16 : TV * STRIDE];
  // L2 filter: the tile's TV squared norms travel with it (wave 0 loads them)
  constexpr bool XN_TILE = (MODE == MODE_FILTER && METRIC == RK_KNN_L2);
  __shared__ float xtile[XN_TILE ? 3 : 1][TV];
  const __amdgpu_buffer_rsrc_t rsx =
      make_rsrc(XN_TILE ? (const void*)(a.xnorm + v0) : (const void*)X,
                XN_TILE && v1 > v0 ? (uint32_t)((v1 - v0) * 4) : 0u);
  const __amdgpu_buffer_rsrc_t rs =
      make_rsrc(X + v0 * a.ldx, v1 > v0 ? (uint32_t)(((v1 - v0 - 1) * a.ldx + a.d) * sizeof(U)) : 0u);
  auto gload = [&](uint4 (&r)[LPT], float& xr, int64_t vt) {
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      const int c = (int)threadIdx.x + j * KNN_NT;
      const int v = c / CPV, pc = c % CPV;
      const int64_t vb = vt + v;
      const bool ok = c < CHUNKS && vb < v1 && pc < nsteps * 2;
      r[j] = bload16(rs, ok ? (uint32_t)((vb - v0) * a.ldx * (int64_t)sizeof(U)) + (uint32_t)pc * 16u
                            : RSRC_OOB);
    }
    if constexpr (XN_TILE) {
      const int64_t vb = vt + (int64_t)threadIdx.x;
      xr = bload4f(rsx, (threadIdx.x < TV && vb < v1) ? (uint32_t)((vb - v0) * 4) : RSRC_OOB);
    }
  };
  auto lstore = [&](const uint4 (&r)[LPT], float xr, int buf) {
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      const int c = (int)threadIdx.x + j * KNN_NT;
      if (c < CHUNKS)
        *reinterpret_cast<uint4*>(&btile[buf][(c / CPV) * STRIDE + (c % CPV) * 16]) = r[j];
    }
    if constexpr (XN_TILE)
      if (threadIdx.x < TV) xtile[buf][threadIdx.x] = xr;
  };
  auto lfrag = [&](uint4 (&bf)[KS], int buf) {
#pragma unroll
    for (int t = 0; t < KS; ++t)
      bf[t] = *reinterpret_cast<const uint4*>(&btile[buf][(lane & 31) * STRIDE + t * 32 + (lane >> 5) * 16]);
  };
  // Candidates of this wave wait in an LDS stage (written and read by this
  // wave only; LDS operations of one wave complete in order) and go to
  // their queries' slots in bulk, so the score loop issues no global atomic.
  uint32_t sn = 0;  // staged entries (wave-uniform)
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t j = (uint32_t)lane; j < sn; j += 64) {
      const uint32_t q = st_q[w][j];
      const uint32_t slot = atomicAdd(&a.cnt[q], 1u);
      if (slot < a.cap) a.cand[(size_t)q * a.cap + slot] = st_key[w][j];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    sn = 0;
  };
  // scores of one tile: MT accumulator sets
  auto mma = [&](const uint4 (&bf)[KS], f32x16 (&c)[MT]) {
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      c[m] = f32x16{};
#pragma unroll
      for (int t = 0; t < KS; ++t) c[m] = Mfma<DT>::run(af[m][t], bf[t], c[m]);
    }
  };
  // squared norm of this lane's vector of the tile (L2)
  auto xn_of = [&](const uint4 (&bf)[KS], int buf) -> float {
    float xn = 0.0f;
    if constexpr (METRIC == RK_KNN_L2) {
      if constexpr (XN_TILE) {
        xn = xtile[buf][lane & 31];
      } else {
        float xpart = 0.0f;
#pragma unroll
        for (int t = 0; t < KS; ++t) xpart = sq8<DT>(bf[t], xpart);
        xn = __fadd_rn(xpart, __shfl_xor(xpart, 32, 64));
      }
    }
    return xn;
  };
  // compares (filter) or stores (scores) of the tile at vt
  auto cmp = [&](const f32x16 (&c)[MT], float xn, int64_t vt) {
    const int64_t vb = vt + (lane & 31);
    const bool vv = vb < v1;
    if constexpr (MODE == MODE_SCORES) {
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int row = m * 32 + rowof(i);
          const float v = (METRIC == RK_KNN_IP) ? c[m][i] : knn_value(METRIC, c[m][i], qr[m][i], xn);
          const uint32_t q = sh_q[w][row];
          if (vv && q != 0xffffffffu) a.out[(size_t)q * a.ldo + vb] = v;
        }
    } else {
      // Compares in groups of four accumulators: one wave-uniform branch per
      // group; the exact check runs only in a group where some lane passed.
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int g = 0; g < 16; g += 4) {
          float v[4], tg[4], qg[4];
          bool pass[4];
          uint64_t any = 0;
          if constexpr (LEAN) {
            const float4 t4 = *reinterpret_cast<const float4*>(&sh_tf[w][m * 32 + rowof(g)]);
            tg[0] = t4.x; tg[1] = t4.y; tg[2] = t4.z; tg[3] = t4.w;
            if constexpr (METRIC == RK_KNN_L2) {
              const float4 q4 = *reinterpret_cast<const float4*>(&sh_qn[w][m * 32 + rowof(g)]);
              qg[0] = q4.x; qg[1] = q4.y; qg[2] = q4.z; qg[3] = q4.w;
            }
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              tg[e] = tf[m][g + e];
              qg[e] = qr[m][g + e];
            }
          }
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int i = g + e;
            v[e] = (METRIC == RK_KNN_IP) ? c[m][i] : knn_value(METRIC, c[m][i], qg[e], xn);
            pass[e] = LARGEST ? !(v[e] < tg[e]) : !(v[e] > tg[e]);
            any |= __ballot(pass[e]);
          }
          if (!any) continue;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            if (!__ballot(pass[e])) continue;
            const int row = m * 32 + rowof(g + e);
            const uint32_t q = sh_q[w][row];
            const uint32_t key = sel_key<F32, LARGEST>(__float_as_uint(v[e]));
            const bool take = pass[e] && vv && q != 0xffffffffu && key >= sh_t[w][row];
            const uint64_t tm = __ballot(take);
            if (tm) {
              const uint32_t nt = (uint32_t)__popcll(tm);
              if (sn + nt > KNN_STAGE) flush();
              if (take) {
                const uint32_t pos = sn + mbcnt(tm);
                st_key[w][pos] = ((uint64_t)key << 32) | (uint32_t)(~(uint32_t)vb);
                st_q[w][pos] = q;
              }
              sn += nt;
            }
          }
        }
    }
  };
  if constexpr (RING) {
    constexpr int LCPV = CPV == 4 ? 2 : CPV == 8 ? 3 : CPV == 16 ? 4 : 5;
    static_assert((1 << LCPV) == CPV, "chunks per vector: a power of two from 4 to 32");
    constexpr int CPW = CHUNKS / NW;              // slots per wave per tile
    constexpr int DI = CPW >= 64 ? CPW / 64 : 1;  // its load instructions
    constexpr int DL = CPW >= 64 ? 64 : CPW;      // active lanes per instruction
    static_assert(CPW % DL == 0 && DI * DL == CPW, "whole load instructions per wave");
    auto swz = [](int v) { return (v ^ (v >> LCPV)) & (CPV - 1); };
    // the DI load instructions of this wave for the tile at vt
    auto dma = [&](int ring, int64_t vt) {
#pragma unroll
      for (int j = 0; j < DI; ++j) {
        const int slot = w * CPW + j * 64 + lane;
        const int v = slot >> LCPV;
        const int c = (slot & (CPV - 1)) ^ swz(v);
        const int64_t vb = vt + v;
        const bool ok = vb < v1 && c < nsteps * 2;
        if (DL == 64 || lane < DL)
          bload16_lds(rs,
                      ok ? (uint32_t)((vb - v0) * a.ldx * (int64_t)sizeof(U)) + (uint32_t)c * 16u
                         : RSRC_OOB,
                      &btile[ring][(w * CPW + j * 64) * 16]);
      }
    };
    // B fragments: chunk 2t + h (h = lane / 32) of vector lane % 32, at
    // byte (v * CPV + ((2t ^ h) ^ swz(v))) * 16 = a0 ^ 32t
    auto lfrag_ring = [&](uint4 (&bf)[KS], int ring) {
      const int v = lane & 31;
      const uint32_t a0 = (uint32_t)(v * CPV + ((lane >> 5) ^ swz(v))) * 16u;
#pragma unroll
      for (int t = 0; t < KS; ++t)
        bf[t] = *reinterpret_cast<const uint4*>(&btile[ring][a0 ^ (uint32_t)(32 * t)]);
    };
    // L2: the tile's squared norms travel the same way (wave 0, one 4-byte
    // load per lane), so wave 0 issues DI + 1 loads per tile.
    auto xdma = [&](int ring, int64_t vt) {
      if constexpr (XN_TILE) {
        const int64_t vb = vt + lane;
        if (w == 0 && lane < TV)
          bload4_lds(rsx, vb < v1 ? (uint32_t)((vb - v0) * 4) : RSRC_OOB,
                     reinterpret_cast<unsigned char*>(&xtile[ring][0]));
      }
    };
    // End of a step: this wave's loads issued before this step's have
    // landed (vmcnt counts vector memory operations in issue order), its LDS
    // writes are done, and after the barrier every wave's are.  The
    // compiler's own waits cover only this wave's loads, so the order across
    // waves rests on this wait.
    auto ring_barrier = [&]() {
      if (XN_TILE && w == 0)
        asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(DI + 1) : "memory");
      else
        asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(DI) : "memory");
    };
    // The query fragments' loads complete here, once, not at their first
    // use inside the loop (s_waitcnt vmcnt, expcnt, lgkmcnt 0).
    __builtin_amdgcn_s_waitcnt(0);
    dma(0, v0);
    xdma(0, v0);
    dma(1, v0 + TV);
    xdma(1, v0 + TV);
    ring_barrier();
    // Step s reads ring slot r0 (tile s) and starts tile s+2 into r2 (read
    // last in step s-1).  Loads past v1 read zeros; every wave of the
    // workgroup runs the same trip count.
    int r0 = 0, r1 = 1, r2 = 2;
    for (int64_t vt = v0; vt < v1; vt += TV) {
      uint4 bf[KS];
      lfrag_ring(bf, r0);
      dma(r2, vt + 2 * TV);
      xdma(r2, vt + 2 * TV);
      f32x16 c[MT];
      mma(bf, c);
      cmp(c, xn_of(bf, r0), vt);
      ring_barrier();
      const int rt = r0;
      r0 = r1;
      r1 = r2;
      r2 = rt;
    }
    // the last loads (zeros past v1) land before the workgroup's LDS is freed
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  } else {
    // Per step: fragments of tile s from one LDS buffer, tile s+1 (in
    // registers) to the other, loads of tile s+3 issued, then the MFMAs and
    // stores of tile s; one barrier per step.  With two register sets each
    // tile's loads have two steps to arrive (a chunk's first reader on an
    // XCD waits for HBM).  Loads past v1 read zeros, so every wave runs the
    // same trip count (rounded up to an even number of steps).
    constexpr int NSET = 2;
    auto body = [&](uint4 (&R)[LPT], float& xr, int buf, int64_t vt) {
      uint4 bf[KS];
      lfrag(bf, buf);
      lstore(R, xr, buf ^ 1);
      gload(R, xr, vt + (NSET + 1) * TV);
      f32x16 c[MT];
      mma(bf, c);
      cmp(c, xn_of(bf, buf), vt);
      __syncthreads();
    };
    uint4 rr[NSET][LPT];
    float xx[NSET] = {};
    gload(rr[0], xx[0], v0);
    lstore(rr[0], xx[0], 0);
#pragma unroll
    for (int j = 0; j < NSET; ++j) gload(rr[j], xx[j], v0 + (j + 1) * TV);
    __syncthreads();
    for (int64_t vt = v0; vt < v1; vt += 2 * TV) {
      body(rr[0], xx[0], 0, vt);
      body(rr[NSET - 1], xx[NSET - 1], 1, vt + TV);
    }
  }
  if constexpr (MODE == MODE_FILTER)
    if (sn) flush();
}

// Squared norm of every database vector, with the operation order of the
// score kernel's own (two halves of each 16-element step summed separately
// in element order, then added), so that scores built from either are
// bit-identical.
template <int DT, int KS>
__global__ __launch_bounds__(256) void k_knn_xnorm(const void* x, int64_t n, int32_t d, float* out) {
  using U = typename Traits<DT>::U;
  // 128 vectors per workgroup.  The block's 16-byte chunks are read with
  // consecutive lanes on consecutive chunks (whole cache lines per load)
  // into LDS rows padded by two chunks, so that the reads below, thread
  // (vector v, half h) taking chunk 2t + h of row v, fall in 16 distinct
  // bank groups.  Each thread then sums its half in the same order as before.
  // Measured on MI355X (2^20 x 128 bf16): 54.7 -> 44.0 us against two
  // threads per vector loading their half's chunks straight from memory.
  constexpr int VPB = 128, CPV = 2 * KS, STR = CPV + 2, PER = VPB * CPV / 256;
  __shared__ uint4 rows[VPB * STR];
  const int tid = threadIdx.x;
  const int64_t v0 = (int64_t)blockIdx.x * VPB;
  const U* X = static_cast<const U*>(x);
  const int nch = d / 8;  // 16-byte chunks per vector (d is a multiple of 16)
  uint4 c[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int ci = tid + u * 256, v = ci / CPV, k = ci % CPV;
    const int64_t j = v0 + v;
    c[u] = (j < n && k < nch) ? ldg16<false>(X + j * d + k * 8) : make_uint4(0u, 0u, 0u, 0u);
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int ci = tid + u * 256;
    rows[(ci / CPV) * STR + ci % CPV] = c[u];
  }
  __syncthreads();
  const int v = tid >> 1, h = tid & 1;
  float acc = 0.0f;
#pragma unroll
  for (int t = 0; t < KS; ++t) acc = sq8<DT>(rows[v * STR + 2 * t + h], acc);
  const float other = __shfl_xor(acc, 1, 64);
  const int64_t j = v0 + v;
  if (j < n && h == 0) out[j] = __fadd_rn(acc, other);
}

template <int DT>
void launch_xnorm(const void* x, int64_t n, int32_t d, float* out, hipStream_t s) {
  const unsigned nb = (unsigned)((n + 127) / 128);
  if (d <= 32) k_knn_xnorm<DT, 2><<<nb, 256, 0, s>>>(x, n, d, out);
  else if (d <= 64) k_knn_xnorm<DT, 4><<<nb, 256, 0, s>>>(x, n, d, out);
  else if (d <= 128) k_knn_xnorm<DT, 8><<<nb, 256, 0, s>>>(x, n, d, out);
  else k_knn_xnorm<DT, 16><<<nb, 256, 0, s>>>(x, n, d, out);
}

// Per query: T_safe = the k-th best of the S sample keys, T = the r-th best.
// Also resets the query's candidate counter and (block 0) the header.
__global__ __launch_bounds__(KNN_SEL_NT) void k_knn_threshold(const float* sample, int32_t S,
                                                              int32_t nq, int64_t n, int32_t k,
                                                              int32_t metric, uint32_t* T,
                                                              uint32_t* T_safe, uint32_t* cnt,
                                                              Header* hdr) {
  constexpr uint32_t TOP_CAP = 1024;  // keys >= T_safe kept for the second select
  __shared__ __attribute__((aligned(16))) uint32_t keys[KNN_SAMPLE];
  __shared__ uint32_t top[TOP_CAP];
  __shared__ uint32_t hist[HIST_MAX_BINS];
  __shared__ uint32_t scratch[16];
  __shared__ uint64_t scratch64[16];
  __shared__ uint32_t outv[4];
  __shared__ uint32_t ntop;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hdr->status = 0;
    hdr->fallback_rows = 0;
    hdr->ordered_rows = 0;
    hdr->total_cand = 0;
  }
  const bool largest = (metric == RK_KNN_IP);
  const float mu = (float)S * (float)k / (float)n;
  const float rf = ceilf(mu + 4.0f * sqrtf(mu) + 2.0f);
  const uint32_t r = (S == n || rf >= (float)k) ? (uint32_t)k : (uint32_t)rf;
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    const uint32_t* row = reinterpret_cast<const uint32_t*>(sample) + (size_t)q * S;
    if (threadIdx.x == 0) ntop = 0;
    // The keys go to LDS with the OR and AND of all of them, so the bits
    // every key shares are known.  One histogram of the next 11 bits finds
    // the bin b of the k-th best key; the keys in bins >= b (the k-1 better
    // ones and the bin's own) are gathered, and both thresholds are selected
    // among them: three passes over the sample instead of a radix select's
    // four plus the gather for T.  When the gathered keys would exceed
    // TOP_CAP the selects run over the whole sample.  (Measured on MI355X,
    // 1024 queries x 16384 sample keys, k = 100: 46.6 -> 40.0 us with the
    // 16-byte loads below.)
    uint32_t kor = 0u, kand = ~0u;
    auto keyof = [&](uint32_t bits) {
      const uint32_t key = largest ? sel_key<F32, true>(bits) : sel_key<F32, false>(bits);
      kor |= key;
      kand &= key;
      return key;
    };
    if ((S & 3) == 0) {
      // 16-byte loads, all of a thread's in flight at once (8 at S = 16384;
      // rows start 64 KB apart in the 256-byte aligned workspace)
      const uint4* row4 = reinterpret_cast<const uint4*>(row);
      const uint32_t n4 = (uint32_t)S / 4;
      for (uint32_t i0 = threadIdx.x; i0 < n4; i0 += KNN_SEL_NT * 8) {
        uint4 x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const uint32_t i = i0 + (uint32_t)u * KNN_SEL_NT;
          x[u] = i < n4 ? ldg16<false>(row4 + i) : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const uint32_t i = i0 + (uint32_t)u * KNN_SEL_NT;
          if (i < n4)
            *reinterpret_cast<uint4*>(&keys[4 * i]) =
                make_uint4(keyof(x[u].x), keyof(x[u].y), keyof(x[u].z), keyof(x[u].w));
        }
      }
    } else {
      load_to_lds<KNN_SEL_NT, 8>(keys, row, (uint32_t)S,
                                 [&](uint32_t bits, uint32_t) { return keyof(bits); });
    }
    {
      const uint64_t both = block_or64<KNN_SEL_NT>(((uint64_t)(~kand) << 32) | kor, scratch64);
      kand = ~(uint32_t)(both >> 32);
      kor = (uint32_t)both;
    }  // its barriers also complete the keys
    const uint32_t diff = kor ^ kand;
    uint32_t ts = kand, tp = kand;  // all keys equal: every rank is that key
    if (diff != 0u) {
      const int vb = 32 - __builtin_clz(diff);  // low bits that vary
      const int shift = vb > 11 ? vb - 11 : 0;
      const uint32_t nb = 1u << (vb - shift);
      for (uint32_t i = threadIdx.x; i < nb; i += KNN_SEL_NT) hist[i] = 0;
      __syncthreads();
      for (int i = threadIdx.x; i < S; i += KNN_SEL_NT)
        atomicAdd(&hist[(keys[i] >> shift) & (nb - 1u)], 1u);
      __syncthreads();
      block_find_crossing<KNN_SEL_NT>(hist, (int)nb, (uint32_t)k, scratch, outv);
      // S >= k, so the k-th key's bin exists; keys >= lo: the bins >= b
      const uint32_t high = vb >= 32 ? 0u : (kand >> vb) << vb;
      const uint32_t lo = high | (outv[0] << shift);
      const uint32_t ngt = outv[1] + outv[2];
      __syncthreads();
      if (ngt <= TOP_CAP) {
        for (int i = threadIdx.x; i < S; i += KNN_SEL_NT) {
          const uint32_t x = keys[i];
          if (x >= lo) top[atomicAdd(&ntop, 1u)] = x;
        }
        __syncthreads();
        ts = block_radix_select<KNN_SEL_NT, 11, uint32_t, false>(top, ngt, k, hist, scratch,
                                                                 scratch64, outv);
        tp = r < (uint32_t)k ? block_radix_select<KNN_SEL_NT, 11, uint32_t, false>(
                                   top, ngt, r, hist, scratch, scratch64, outv)
                             : ts;
      } else {
        ts = block_radix_select_from<KNN_SEL_NT, 11, uint32_t, false>(keys, S, k, diff, kand, hist,
                                                                      scratch, outv);
        tp = r < (uint32_t)k ? block_radix_select_from<KNN_SEL_NT, 11, uint32_t, false>(
                                   keys, S, r, diff, kand, hist, scratch, outv)
                             : ts;
      }
    }
    if (threadIdx.x == 0) {
      T_safe[q] = ts;
      T[q] = tp;
      cnt[q] = 0;
    }
    __syncthreads();
  }
}

// Per query: the exact top-k of its candidates (keys >= T), written best
// first (sorted) or in database order.  Pass 0 queues a query whose
// candidates cannot hold its top-k for a retry with T_safe; pass 1 reports it.
__global__ __launch_bounds__(KNN_SEL_NT) void k_knn_select(int pass, const uint32_t* list,
                                                           const uint32_t* nlist_dev, int32_t nq,
                                                           int32_t k, int32_t metric, int sorted,
                                                           uint32_t* T, const uint32_t* T_safe,
                                                           uint32_t* cnt, const uint64_t* cand,
                                                           uint32_t cap, int32_t lds_keys,
                                                           uint32_t* faillist, Header* hdr,
                                                           float* out_v, int32_t* out_i) {
  extern __shared__ uint64_t sbuf[];
  __shared__ uint32_t hist[HIST_MAX_BINS];
  __shared__ uint32_t ctab[256];
  __shared__ uint32_t scratch[16];
  __shared__ uint64_t scratch64[3 * KNN_SEL_NT / 64];
  __shared__ uint32_t outv[4];
  __shared__ uint32_t slot_cnt;
  const int32_t nl = nlist_dev ? (int32_t)*nlist_dev : nq;
  const bool largest = (metric == RK_KNN_IP);
  for (int e = blockIdx.x; e < nl; e += gridDim.x) {
    const uint32_t q = list ? list[e] : (uint32_t)e;
    const uint32_t n = cnt[q];
    if (n < (uint32_t)k || n > cap) {
      if (threadIdx.x == 0) {
        if (pass == 0) {
          faillist[atomicAdd(&hdr->fallback_rows, 1u)] = q;
          T[q] = T_safe[q];
          cnt[q] = 0;
        } else {
          atomicOr(&hdr->status, (uint32_t)ST_CANDIDATES);
        }
      }
      if (pass != 0) {
        // no exact result: the row reads as failed (value NaN, index -1)
        for (uint32_t i = threadIdx.x; i < (uint32_t)k; i += KNN_SEL_NT) {
          out_v[(size_t)q * k + i] = __builtin_nanf("");
          out_i[(size_t)q * k + i] = -1;
        }
      }
      __syncthreads();
      continue;
    }
    const uint64_t* src = cand + (size_t)q * cap;
    const bool in_lds = n <= (uint32_t)lds_keys;
    if (in_lds) load_to_lds<KNN_SEL_NT, 8>(sbuf, src, n, [](uint64_t x, uint32_t) { return x; });
    __syncthreads();
    // Up to half the LDS buffer of candidates: one bucket sort of all of them
    // (the other half groups them; composite keys are distinct) leaves the
    // selection best first in sbuf[0, k), so the radix select and the
    // compaction are skipped (measured on MI355X, 1024 queries with ~600
    // candidates, k = 100: 28.2 -> 18.5 us).  Keys that do not spread over the
    // buckets are left untouched, and the radix select runs.
    const bool whole = in_lds && n > 64 && n <= (uint32_t)lds_keys / 2 &&
                       bucket_sort_desc<KNN_SEL_NT>(sbuf, sbuf + lds_keys / 2, n, hist, ctab,
                                                    scratch64, scratch);
    if (whole) {
      if (!sorted) {  // database order: the k selected keys re-keyed, sorted below
        for (uint32_t i = threadIdx.x; i < (uint32_t)k; i += KNN_SEL_NT) {
          const uint64_t x = sbuf[i];
          sbuf[i] = ((uint64_t)(uint32_t)x << 32) | (x >> 32);
        }
      }
    } else {
      const uint64_t kth =
          in_lds ? block_radix_select<KNN_SEL_NT, 11, uint64_t, false>(sbuf, n, k, hist, scratch, scratch64, outv)
                 : block_radix_select<KNN_SEL_NT, 11, uint64_t, true>(src, n, k, hist, scratch, scratch64, outv);
      if (threadIdx.x == 0) slot_cnt = 0;
      __syncthreads();
      // Keep the k keys >= kth (composite keys are distinct), re-keyed so that
      // one descending sort gives the output order: best first, or database
      // order ((~index << 32) | key).
      constexpr int CH = 4;
      for (uint32_t c0 = 0; c0 < n; c0 += KNN_SEL_NT * CH) {
        uint64_t reg[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          const uint32_t i = c0 + threadIdx.x + j * KNN_SEL_NT;
          reg[j] = i < n ? (in_lds ? sbuf[i] : src[i]) : 0ull;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          const uint32_t i = c0 + threadIdx.x + j * KNN_SEL_NT;
          const bool take = i < n && reg[j] >= kth;
          const uint64_t m = __ballot(take);
          if (m) {
            const int fl = __ffsll((unsigned long long)m) - 1;
            uint32_t base = 0;
            if (lane_id() == fl) base = atomicAdd(&slot_cnt, (uint32_t)__popcll(m));
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, fl) + mbcnt(m);
            if (take && base < (uint32_t)k)
              sbuf[base] = sorted ? reg[j]
                                  : (((uint64_t)(uint32_t)reg[j] << 32) | (reg[j] >> 32));
          }
        }
      }
    }
    __syncthreads();
    if (!(whole && sorted)) {
      const uint32_t P2 = next_pow2((uint32_t)k);
      for (uint32_t i = (uint32_t)k + threadIdx.x; i < P2; i += KNN_SEL_NT) sbuf[i] = 0;
      __syncthreads();
      if (P2 > 1) block_bitonic_desc<KNN_SEL_NT>(sbuf, P2);
    }
    for (uint32_t i = threadIdx.x; i < (uint32_t)k; i += KNN_SEL_NT) {
      const uint64_t v = sbuf[i];
      const uint32_t sk = sorted ? (uint32_t)(v >> 32) : (uint32_t)v;
      const uint32_t idx = ~(sorted ? (uint32_t)v : (uint32_t)(v >> 32));
      const uint32_t bits = largest ? key_to_bits<F32, true>(sk) : key_to_bits<F32, false>(sk);
      out_v[(size_t)q * k + i] = __uint_as_float(bits);
      out_i[(size_t)q * k + i] = (int32_t)idx;
    }
    __syncthreads();
  }
}

struct KnnLayout {
  size_t hdr, T, T_safe, cnt, fail, xnorm, big, total;
  int32_t S;
  uint32_t cap;
};

size_t a256(size_t x) { return (x + 255) & ~size_t(255); }

uint32_t next_pow2_host(uint32_t x) {
  uint32_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

KnnLayout knn_layout(int32_t nq, int64_t n, int32_t k) {
  KnnLayout L{};
  L.S = (int32_t)std::min<int64_t>(n, KNN_SAMPLE);
  // candidate slots per query: twice the expected count at T_safe, plus
  // 4096, at most n
  const double exp_safe = (double)k * (double)n / (double)L.S;
  L.cap = (uint32_t)std::min<double>((double)n, 2.0 * exp_safe + 4096.0);
  const size_t Q = (size_t)std::max(nq, 1);
  size_t o = 0;
  L.hdr = o; o = a256(o + sizeof(Header));
  L.T = o; o = a256(o + Q * 4);
  L.T_safe = o; o = a256(o + Q * 4);
  L.cnt = o; o = a256(o + Q * 4);
  L.fail = o; o = a256(o + Q * 4);
  L.xnorm = o; o = a256(o + (size_t)n * 4);
  // the sample scores are dead once the thresholds exist: the candidate
  // slots reuse their bytes
  L.big = o;
  o = a256(o + std::max(Q * (size_t)L.S * 4, Q * (size_t)L.cap * 8));
  L.total = o;
  return L;
}

// Query tiles per wave.  Scores mode: two 32-query tiles share each B
// fragment when their A fragments fit the register budget (d <= 128).
// Filter mode: one, since B comes from LDS there and the registers freed
// (A fragments, thresholds) buy a third wave per SIMD, which measured ~2 %
// faster than two tiles at two waves.
template <int KS, int MODE>
constexpr int knn_mt() { return (MODE == MODE_SCORES && KS <= 8) ? 2 : 1; }

template <int DT, int KS, int MODE, int METRIC>
hipError_t launch_knn(const KnnArgs& a0, hipStream_t s) {
  constexpr int MT = knn_mt<KS, MODE>();
  constexpr int BQ = 32 * MT * (KNN_NT / 64);  // queries per workgroup
  KnnArgs a = a0;
  const int64_t tiles = (a.nv + 31) / 32;
  const int64_t qblocks = (a.nq + BQ - 1) / BQ;
  // about 2048 workgroups: chunk = tiles * qblocks / 2048 tiles, in [4, 256]
  // (the 16384-vector sample launch of 1024 queries gets 512 workgroups;
  // measured on MI355X: 27.8 us with a floor of 8 tiles (256 workgroups),
  // 21.6 us with 4, 25.0 us with 2)
  int64_t ct = (tiles * qblocks + 2047) / 2048;
  ct = std::max<int64_t>(4, std::min<int64_t>(256, ct));
  ct = std::min<int64_t>(ct, std::max<int64_t>(tiles, 1));
  // a chunk's byte range (one buffer descriptor) stays below 2^32; 32
  // vectors always fit, since a vector stride is at most 2^31 / 2^14
  // elements of 2 bytes
  ct = std::min<int64_t>(ct, std::max<int64_t>(1, (int64_t)0xffffff00 / (a.ldx * 2) / 32));
  a.chunk_v = (int32_t)(ct * 32);
  a.qblocks = (int32_t)std::max<int64_t>(qblocks, 1);
  a.chunks = (int32_t)((a.nv + a.chunk_v - 1) / a.chunk_v);
  const int64_t per_xcd = (int64_t)a.qblocks * ((a.chunks + 7) / 8);
  const int64_t grid = per_xcd * 8;
  if (grid <= 0 || a.nq <= 0) return hipSuccess;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;  // beyond one launch
  k_knn<DT, KS, MT, MODE, METRIC><<<(unsigned)grid, KNN_NT, 0, s>>>(a);
  return hipGetLastError();
}

template <int DT, int MODE, int METRIC>
hipError_t launch_knn_d(const KnnArgs& a, hipStream_t s) {
  if (a.d <= 32) return launch_knn<DT, 2, MODE, METRIC>(a, s);
  if (a.d <= 64) return launch_knn<DT, 4, MODE, METRIC>(a, s);
  if (a.d <= 128) return launch_knn<DT, 8, MODE, METRIC>(a, s);
  return launch_knn<DT, 16, MODE, METRIC>(a, s);
}

template <int DT, int MODE>
hipError_t launch_knn_m(const KnnArgs& a, hipStream_t s) {
  return a.metric == RK_KNN_IP ? launch_knn_d<DT, MODE, RK_KNN_IP>(a, s)
                               : launch_knn_d<DT, MODE, RK_KNN_L2>(a, s);
}

template <int MODE>
hipError_t launch_knn_t(int dtype, const KnnArgs& a, hipStream_t s) {
  return dtype == RK_BF16 ? launch_knn_m<BF16, MODE>(a, s) : launch_knn_m<F16, MODE>(a, s);
}

int check_common(int dtype, int metric, const void* queries, const void* base, int32_t nq, int64_t n,
                 int32_t d) {
  if (dtype != RK_BF16 && dtype != RK_F16) return RK_E_UNSUPPORTED;
  if (metric != RK_KNN_IP && metric != RK_KNN_L2) return RK_E_ARG;
  if (nq < 0 || n < 1 || n > INT32_MAX || d < 16 || d > KNN_DMAX || d % 16) return RK_E_ARG;
  if ((nq && !queries) || !base) return RK_E_ARG;
  if (((uintptr_t)queries % 16) || ((uintptr_t)base % 16)) return RK_E_ALIGN;
  return RK_OK;
}

}  // namespace
}  // namespace rk

using namespace rk;

extern "C" int rk_knn_workspace_size(int dtype, int32_t nq, int64_t n, int32_t d, int32_t k,
                                     size_t* bytes) {
  if (!bytes) return RK_E_ARG;
  if (dtype != RK_BF16 && dtype != RK_F16) return RK_E_UNSUPPORTED;
  if (nq < 0 || n < 1 || n > INT32_MAX || d < 16 || d > KNN_DMAX || d % 16) return RK_E_ARG;
  if (k < 1 || k > KNN_K_MAX || k > n) return RK_E_ARG;
  *bytes = knn_layout(nq, n, k).total;
  return RK_OK;
}

extern "C" int rk_knn_scores(int dtype, int metric, const void* queries, const void* base,
                             int32_t nq, int64_t n, int32_t d, float* out, void* stream) {
  const int rc = check_common(dtype, metric, queries, base, nq, n, d);
  if (rc != RK_OK) return rc;
  if (nq && !out) return RK_E_ARG;
  if ((uintptr_t)out % 4) return RK_E_ALIGN;
  KnnArgs a{};
  a.q = queries;
  a.x = base;
  a.nq = nq;
  a.d = d;
  a.nv = n;
  a.ldx = d;
  a.metric = metric;
  a.out = out;
  a.ldo = n;
  return launch_knn_t<MODE_SCORES>(dtype, a, static_cast<hipStream_t>(stream)) == hipSuccess
             ? RK_OK
             : RK_E_HIP;
}

extern "C" int rk_knn_topk(int dtype, int metric, const void* queries, const void* base,
                           int32_t nq, int64_t n, int32_t d, int32_t k, uint32_t flags,
                           float* out_vals, int32_t* out_idx, void* workspace, size_t ws_bytes,
                           void* stream) {
  int rc = check_common(dtype, metric, queries, base, nq, n, d);
  if (rc != RK_OK) return rc;
  if (k < 1 || k > KNN_K_MAX || k > n) return RK_E_ARG;
  if (flags & ~(uint32_t)RK_SORTED) return RK_E_ARG;
  if (nq && (!out_vals || !out_idx)) return RK_E_ARG;
  if (((uintptr_t)out_vals % 4) || ((uintptr_t)out_idx % 4)) return RK_E_ALIGN;
  if (!workspace) return RK_E_WORKSPACE;
  if ((uintptr_t)workspace % 256) return RK_E_ALIGN;
  const KnnLayout L = knn_layout(nq, n, k);
  if (ws_bytes < L.total) return RK_E_WORKSPACE;
  if (nq == 0) return RK_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* ws = static_cast<char*>(workspace);
  Header* hdr = reinterpret_cast<Header*>(ws + L.hdr);
  uint32_t* T = reinterpret_cast<uint32_t*>(ws + L.T);
  uint32_t* T_safe = reinterpret_cast<uint32_t*>(ws + L.T_safe);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(ws + L.cnt);
  uint32_t* fail = reinterpret_cast<uint32_t*>(ws + L.fail);
  float* sample = reinterpret_cast<float*>(ws + L.big);
  uint64_t* cand = reinterpret_cast<uint64_t*>(ws + L.big);
  const uint32_t* nfail = &hdr->fallback_rows;
  const int lds_keys = std::max(4096, (int)next_pow2_host((uint32_t)k));
  const size_t dyn = (size_t)lds_keys * 8;
  static std::once_flag once;
  std::call_once(once, [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_knn_select),
                              hipFuncAttributeMaxDynamicSharedMemorySize, KNN_K_MAX * 8);
  });
  // 1. sample scores: S vectors at stride floor(n / S)
  KnnArgs a{};
  a.q = queries;
  a.d = d;
  a.nq = nq;
  a.metric = metric;
  const int64_t stride = n / L.S;
  a.x = base;
  a.nv = L.S;
  a.ldx = stride * d;
  a.out = sample;
  a.ldo = L.S;
  if (launch_knn_t<MODE_SCORES>(dtype, a, s) != hipSuccess) return RK_E_HIP;
  // 2. thresholds
  const int tgrid = std::min(nq, 4096);
  k_knn_threshold<<<tgrid, KNN_SEL_NT, 0, s>>>(sample, L.S, nq, n, k, metric, T, T_safe, cnt, hdr);
  if (hipGetLastError() != hipSuccess) return RK_E_HIP;
  // squared database norms (L2), once per call
  float* xnorm = reinterpret_cast<float*>(ws + L.xnorm);
  if (metric == RK_KNN_L2) {
    if (dtype == RK_BF16) launch_xnorm<BF16>(base, n, d, xnorm, s);
    else launch_xnorm<F16>(base, n, d, xnorm, s);
    if (hipGetLastError() != hipSuccess) return RK_E_HIP;
  }
  // 3./5. fused filter; 4./6. select
  const int sgrid = std::min(nq, 8192);
  for (int pass = 0; pass < 2; ++pass) {
    KnnArgs f{};
    f.q = queries;
    f.x = base;
    f.nq = nq;
    f.d = d;
    f.nv = n;
    f.ldx = d;
    f.metric = metric;
    f.list = pass ? fail : nullptr;
    f.nlist = pass ? nfail : nullptr;
    f.T = T;
    f.cnt = cnt;
    f.cand = cand;
    f.cap = L.cap;
    f.xnorm = metric == RK_KNN_L2 ? xnorm : nullptr;
    if (launch_knn_t<MODE_FILTER>(dtype, f, s) != hipSuccess) return RK_E_HIP;
    k_knn_select<<<sgrid, KNN_SEL_NT, dyn, s>>>(pass, pass ? fail : nullptr, pass ? nfail : nullptr,
                                                nq, k, metric, (flags & RK_SORTED) ? 1 : 0, T, T_safe,
                                                cnt, cand, L.cap, lds_keys, fail, hdr, out_vals,
                                                out_idx);
    if (hipGetLastError() != hipSuccess) return RK_E_HIP;
  }
  return RK_OK;
}
