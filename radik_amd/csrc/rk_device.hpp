// Device-side building blocks for the gfx950 radix top-k: element traits,
// the order-preserving key flip, wave64 / workgroup scans and reductions,
// the workgroup radix select over keys, and the LDS bitonic sort.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RK_DEV __device__ __forceinline__
#define RK_HOST_DEV __host__ __device__ __forceinline__

namespace rk {

enum Dtype : int { F32 = 0, F16 = 1, BF16 = 2 };

template <int DT>
struct Traits;
template <>
struct Traits<F32> {
  using U = uint32_t;                 // storage word of one element
  static constexpr int KEY_BITS = 32; // key width
  static constexpr int PACK = 4;      // elements per 16-byte load
  static constexpr uint32_t KEY_MASK = 0xffffffffu;
};
template <>
struct Traits<F16> {
  using U = uint16_t;
  static constexpr int KEY_BITS = 16;
  static constexpr int PACK = 8;
  static constexpr uint32_t KEY_MASK = 0xffffu;
};
template <>
struct Traits<BF16> {
  using U = uint16_t;
  static constexpr int KEY_BITS = 16;
  static constexpr int PACK = 8;
  static constexpr uint32_t KEY_MASK = 0xffffu;
};

// ---------------------------------------------------------------------------
// Keys.  The flip follows radik/RadixSelect/utils.cuh:329-334 (mask = sign ?
// all-ones : sign-bit; key = bits ^ mask), applied at the element's native
// width (32 bits for f32, 16 for f16/bf16).  The selection key makes "larger
// is better" for both directions (bit-inverted for smallest-k) and sends NaN
// to 0, the worst key, so NaN is chosen only when fewer than k non-NaN exist.
// ---------------------------------------------------------------------------
template <int DT, bool LARGEST>
RK_DEV uint32_t sel_key(uint32_t u) {
  if constexpr (DT == F32) {
    uint32_t f = u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
    if (!LARGEST) f = ~f;
    return ((u & 0x7fffffffu) > 0x7f800000u) ? 0u : f;
  } else {
    constexpr uint32_t NAN_ABS = (DT == F16) ? 0x7c00u : 0x7f80u;
    uint32_t f = (u ^ ((u & 0x8000u) ? 0xffffu : 0x8000u)) & 0xffffu;
    if (!LARGEST) f ^= 0xffffu;
    return ((u & 0x7fffu) > NAN_ABS) ? 0u : f;
  }
}

// Inverse of sel_key for a non-zero key: the element's bit pattern.
template <int DT, bool LARGEST>
RK_DEV uint32_t key_to_bits(uint32_t sk) {
  if constexpr (DT == F32) {
    uint32_t f = LARGEST ? sk : ~sk;
    return (f & 0x80000000u) ? (f ^ 0x80000000u) : ~f;
  } else {
    uint32_t f = LARGEST ? sk : (sk ^ 0xffffu);
    return (f & 0x8000u) ? (f ^ 0x8000u) : (~f & 0xffffu);
  }
}

template <int DT>
RK_DEV float bits_to_float(uint32_t u) {
  if constexpr (DT == F32) {
    return __uint_as_float(u);
  } else if constexpr (DT == BF16) {
    return __uint_as_float(u << 16);
  } else {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)u);
  }
}

template <int DT>
RK_DEV bool bits_is_nan(uint32_t u) {
  if constexpr (DT == F32) return (u & 0x7fffffffu) > 0x7f800000u;
  else if constexpr (DT == F16) return (u & 0x7fffu) > 0x7c00u;
  else return (u & 0x7fffu) > 0x7f80u;
}

// Prefilter of the streaming filter.  For a key threshold T > 0 the XOR mask
// m below, chosen by the half of key space T lies in, makes (raw ^ m) as an
// unsigned integer equal the selection key of every element in that half;
// the elements of the other half all map below T (none can pass) or all
// above it (the exact test decides).  So a 16-byte vector can hold a key >= T
// only if max_u(raw_i ^ m) >= T: one xor and one unsigned max per word
// (packed halves for 16-bit types).  Integer only, so it does not depend on
// the float unit's NaN or denormal handling; NaN patterns may pass (their
// key is 0 < T) and the exact per-element test sel_key(u) >= T removes them.
// (Checked exhaustively over all 16-bit patterns: no element with a key >= T
// fails the prefilter.)
template <int DT, bool LARGEST>
RK_DEV uint32_t prefilter_mask(uint32_t T) {
  constexpr uint32_t KM = Traits<DT>::KEY_MASK, TOP = (KM >> 1) + 1;
  uint32_t m = LARGEST ? (T >= TOP ? TOP : KM) : (T >= TOP ? 0u : (KM ^ TOP));
  if constexpr (Traits<DT>::KEY_BITS == 16) m |= m << 16;
  return m;
}

typedef unsigned short rk_u16x2 __attribute__((ext_vector_type(2)));
template <int DT>
RK_DEV uint32_t vec_max_xm(const uint4& x, uint32_t m) {
  if constexpr (DT == F32) {
    return max(max(x.x ^ m, x.y ^ m), max(x.z ^ m, x.w ^ m));
  } else {
    auto h = [&](uint32_t w) { return __builtin_bit_cast(rk_u16x2, w ^ m); };
    const rk_u16x2 r = __builtin_elementwise_max(__builtin_elementwise_max(h(x.x), h(x.y)),
                                                 __builtin_elementwise_max(h(x.z), h(x.w)));
    return max((uint32_t)r.x, (uint32_t)r.y);
  }
}

// Scalar loads through explicit address spaces.  Generic pointers may alias
// LDS, which stops the compiler from moving a global load above an earlier
// LDS store or atomic; naming the address space lets loops keep several
// global loads in flight.
template <typename T>
RK_DEV T ldg(const T* p) {
  return *((const __attribute__((address_space(1))) T*)p);
}
template <typename T>
RK_DEV T ld_lds(const T* p) {
  return *((const __attribute__((address_space(3))) T*)p);
}

// 16-byte loads through an explicit global (address space 1) pointer, so the
// compiler emits global_load_dwordx4 rather than flat loads (row bases reach
// the kernels as generic pointers).  NT selects the non-temporal (streaming)
// form for data read exactly once.
typedef unsigned int rk_v4u __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) rk_v4u* rk_gv4;
template <bool NT>
RK_DEV uint4 ldg16(const void* p) {
  rk_v4u v;
  if constexpr (NT) v = __builtin_nontemporal_load((rk_gv4)p);
  else v = *((rk_gv4)p);
  return make_uint4(v.x, v.y, v.z, v.w);
}

// Buffer loads: one resource descriptor (base + byte range, in SGPRs) per
// streamed range and 32-bit per-lane byte offsets; the hardware returns 0 for
// offsets outside the range, so the loads need no guards.  Descriptor word 3
// of the gfx9 family (incl. gfx950): 0x00020000.
RK_DEV __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  // The callers pass workgroup-uniform values; readfirstlane makes that
  // explicit so the descriptor lives in SGPRs (no per-load waterfall loop).
  const uint64_t a = (uint64_t)base;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  const uint32_t n = __builtin_amdgcn_readfirstlane(bytes);
  void* b = (void*)(((uint64_t)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(b, (short)0, (int)n, 0x00020000);
}
// 16-byte buffer load with the non-temporal (streaming) hint: data read once.
RK_DEV uint4 bload16_nt(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  const rk_v4u v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)byte_off, 0, 2);
  return make_uint4(v.x, v.y, v.z, v.w);
}
// 16-byte buffer load, cached normally (data other workgroups read again).
RK_DEV uint4 bload16(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  const rk_v4u v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)byte_off, 0, 0);
  return make_uint4(v.x, v.y, v.z, v.w);
}
// 4-byte buffer load, cached normally.
RK_DEV float bload4f(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0));
}
// A byte offset beyond every range a descriptor here describes (loads at it
// return zeros).
constexpr uint32_t RSRC_OOB = 0xfffffff0u;

// 16-byte buffer load straight into LDS (gfx950): the wave's active lanes
// write the 16-byte slots dst + 16 * lane (dst wave-uniform); offsets
// outside the descriptor's range write zeros.  No VGPR is written; the load
// counts in vmcnt until its data is in LDS, and the compiler waits for it
// (s_waitcnt vmcnt) before LDS accesses it cannot tell apart from its slots.
RK_DEV void bload16_lds(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, unsigned char* dst) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)dst, 16,
                                           (int)byte_off, 0, 0, 0);
}
// The same with 4 bytes per lane (slots dst + 4 * lane).
RK_DEV void bload4_lds(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, unsigned char* dst) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)dst, 4,
                                           (int)byte_off, 0, 0, 0);
}

// Unpacked 16-byte vector of PACK elements (bit patterns, low element first).
template <int DT>
struct Vec {
  uint32_t e[Traits<DT>::PACK];
};

template <int DT>
RK_DEV Vec<DT> unpack(const uint4& x) {
  Vec<DT> v;
  if constexpr (DT == F32) {
    v.e[0] = x.x; v.e[1] = x.y; v.e[2] = x.z; v.e[3] = x.w;
  } else {
    v.e[0] = x.x & 0xffffu; v.e[1] = x.x >> 16;
    v.e[2] = x.y & 0xffffu; v.e[3] = x.y >> 16;
    v.e[4] = x.z & 0xffffu; v.e[5] = x.z >> 16;
    v.e[6] = x.w & 0xffffu; v.e[7] = x.w >> 16;
  }
  return v;
}

// ---------------------------------------------------------------------------
// Device-coherent accesses between the workgroups of one running kernel.
// The XCDs' L2 caches are not coherent with each other for plain accesses;
// agent-scope atomic loads and stores (and read-modify-writes) are, without
// the L2 write-back / invalidation of an agent-scope fence.  A workgroup
// that signals another one with an atomic first waits for its own earlier
// accesses to complete (wait_vmem), so they are performed before the signal.
// ---------------------------------------------------------------------------
template <typename T>
RK_DEV T ld_dev(const T* a) {
  return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
RK_DEV void st_dev(T* a, T v) {
  __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// This wave's outstanding vector memory operations (loads, stores, atomics)
// completed.
RK_DEV void wait_vmem() { __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// ---------------------------------------------------------------------------
// Wave64 / workgroup primitives.
// ---------------------------------------------------------------------------
RK_DEV int lane_id() { return threadIdx.x & 63; }

// popcount(mask & lanes below this one)
RK_DEV uint32_t mbcnt(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave64 scans and reductions on DPP (data-parallel primitive) moves: four
// row_shr steps (distances 1, 2, 4, 8) scan or reduce each row of 16 lanes
// in VALU operations, and readlane of the rows' last lanes (15, 31, 47, 63)
// joins the four rows in scalar registers.  (Lane shuffles would go through
// the LDS crossbar, ds_bpermute, with its latency on every step.)  Called
// with the whole wave active.
//
// row_shr:n (DPP control 0x110 + n, gfx9 family): lane i of a row reads lane
// i - n of the same row; a lane whose source lies before the row keeps
// `id`, the operation's identity.
template <int N>
RK_DEV uint32_t row_shr(uint32_t v, uint32_t id) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x110 + N, 0xf, 0xf, false);
}
template <int N>
RK_DEV uint64_t row_shr64(uint64_t v, uint64_t id) {
  return ((uint64_t)row_shr<N>((uint32_t)(v >> 32), (uint32_t)(id >> 32)) << 32) |
         row_shr<N>((uint32_t)v, (uint32_t)id);
}
RK_DEV uint32_t readlane32(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
RK_DEV uint64_t readlane64(uint64_t v, int l) {
  return ((uint64_t)readlane32((uint32_t)(v >> 32), l) << 32) | readlane32((uint32_t)v, l);
}

// Inclusive scan (op) of each row of 16 lanes.
template <typename T, typename Op>
RK_DEV T row_scan(T x, T id, Op op) {
  if constexpr (sizeof(T) == 8) {
    x = op(x, row_shr64<1>(x, id));
    x = op(x, row_shr64<2>(x, id));
    x = op(x, row_shr64<4>(x, id));
    x = op(x, row_shr64<8>(x, id));
  } else {
    x = op(x, row_shr<1>(x, id));
    x = op(x, row_shr<2>(x, id));
    x = op(x, row_shr<4>(x, id));
    x = op(x, row_shr<8>(x, id));
  }
  return x;
}

// Reduction (op) over the wave, the same value in every lane.
template <typename T, typename Op>
RK_DEV T wave_reduce(T x, T id, Op op) {
  x = row_scan(x, id, op);
  if constexpr (sizeof(T) == 8)
    return op(op(readlane64(x, 15), readlane64(x, 31)), op(readlane64(x, 47), readlane64(x, 63)));
  else
    return op(op(readlane32(x, 15), readlane32(x, 31)), op(readlane32(x, 47), readlane32(x, 63)));
}

// Inclusive sum over the lanes up to this one.
template <typename T>
RK_DEV T wave_incl_scan_t(T x) {
  auto add = [](T a, T b) { return a + b; };
  x = row_scan(x, T(0), add);
  T r0, r1, r2;
  if constexpr (sizeof(T) == 8) {
    r0 = readlane64(x, 15);
    r1 = readlane64(x, 31);
    r2 = readlane64(x, 47);
  } else {
    r0 = readlane32(x, 15);
    r1 = readlane32(x, 31);
    r2 = readlane32(x, 47);
  }
  const int row = lane_id() >> 4;
  return x + (row >= 1 ? r0 : T(0)) + (row >= 2 ? r1 : T(0)) + (row >= 3 ? r2 : T(0));
}

RK_DEV uint32_t wave_incl_scan(uint32_t x) { return wave_incl_scan_t<uint32_t>(x); }
RK_DEV uint64_t wave_incl_scan64(uint64_t x) { return wave_incl_scan_t<uint64_t>(x); }

RK_DEV uint32_t wave_sum(uint32_t x) {
  return wave_reduce<uint32_t>(x, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}

RK_DEV uint64_t wave_or64(uint64_t x) {
  return wave_reduce<uint64_t>(x, 0ull, [](uint64_t a, uint64_t b) { return a | b; });
}

RK_DEV uint64_t wave_min64(uint64_t x) {
  return wave_reduce<uint64_t>(x, ~0ull, [](uint64_t a, uint64_t b) { return a < b ? a : b; });
}

RK_DEV uint64_t wave_max64(uint64_t x) {
  return wave_reduce<uint64_t>(x, 0ull, [](uint64_t a, uint64_t b) { return a > b ? a : b; });
}

// Exclusive scan over the workgroup in thread order.  scratch: >= NT/64 words.
template <int NT>
RK_DEV uint32_t block_excl_scan(uint32_t x, uint32_t* scratch, uint32_t& total) {
  constexpr int NW = NT / 64;
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(x);
  if (lane == 63) scratch[wid] = inc;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const uint32_t v = scratch[w];
    pre += (w < wid) ? v : 0u;
    tot += v;
  }
  __syncthreads();
  total = tot;
  return pre + inc - x;
}

template <int NT>
RK_DEV uint32_t block_sum(uint32_t x, uint32_t* scratch) {
  constexpr int NW = NT / 64;
  x = wave_sum(x);
  if (lane_id() == 0) scratch[threadIdx.x >> 6] = x;
  __syncthreads();
  uint32_t tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) tot += scratch[w];
  __syncthreads();
  return tot;
}

template <int NT>
RK_DEV uint64_t block_or64(uint64_t x, uint64_t* scratch) {
  constexpr int NW = NT / 64;
  x = wave_or64(x);
  if (lane_id() == 0) scratch[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t r = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) r |= scratch[w];
  __syncthreads();
  return r;
}

// Workgroup minimum and maximum of 64-bit values.  scratch: >= 2 NT/64 words.
template <int NT>
RK_DEV void block_minmax64(uint64_t& lo, uint64_t& hi, uint64_t* scratch) {
  constexpr int NW = NT / 64;
  lo = wave_min64(lo);
  hi = wave_max64(hi);
  if (lane_id() == 0) {
    scratch[threadIdx.x >> 6] = lo;
    scratch[NW + (threadIdx.x >> 6)] = hi;
  }
  __syncthreads();
  lo = ~0ull;
  hi = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    lo = scratch[w] < lo ? scratch[w] : lo;
    hi = scratch[NW + w] > hi ? scratch[NW + w] : hi;
  }
  __syncthreads();
}

// Exclusive 64-bit scan over the workgroup in thread order.  scratch: >=
// NT/64 words.
template <int NT>
RK_DEV uint64_t block_excl_scan64(uint64_t x, uint64_t* scratch, uint64_t& total) {
  constexpr int NW = NT / 64;
  const int wid = threadIdx.x >> 6;
  const uint64_t inc = wave_incl_scan64(x);
  if (lane_id() == 63) scratch[wid] = inc;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const uint64_t v = scratch[w];
    pre += (w < wid) ? v : 0ull;
    tot += v;
  }
  __syncthreads();
  total = tot;
  return pre + inc - x;
}

// Workgroup OR, minimum and maximum of 64-bit values in one exchange.
// scratch: >= 3 NT/64 words.
template <int NT>
RK_DEV void block_or_min_max64(uint64_t& o, uint64_t& lo, uint64_t& hi, uint64_t* scratch) {
  constexpr int NW = NT / 64;
  o = wave_or64(o);
  lo = wave_min64(lo);
  hi = wave_max64(hi);
  const int wid = threadIdx.x >> 6;
  if (lane_id() == 0) {
    scratch[wid] = o;
    scratch[NW + wid] = lo;
    scratch[2 * NW + wid] = hi;
  }
  __syncthreads();
  o = 0;
  lo = ~0ull;
  hi = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    o |= scratch[w];
    lo = scratch[NW + w] < lo ? scratch[NW + w] : lo;
    hi = scratch[2 * NW + w] > hi ? scratch[2 * NW + w] : hi;
  }
  __syncthreads();
}

// dst[i] = f(src[i], i) for i < n, over the workgroup, with UN global loads
// in flight per thread: all loads of a batch are issued before the first
// LDS store, so a copy of n elements costs ceil(n / (NT * UN)) memory round
// trips instead of ceil(n / NT).  (A plain load -> store loop waits for each
// load before its store.)
template <int NT, int UN, typename S, typename D, typename F>
RK_DEV void load_to_lds(D* dst, const S* src, uint32_t n, F&& f) {
  for (uint32_t i0 = threadIdx.x; i0 < n; i0 += NT * UN) {
    S x[UN];
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      const uint32_t i = i0 + (uint32_t)u * NT;
      x[u] = i < n ? ldg(src + i) : S(0);
    }
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      const uint32_t i = i0 + (uint32_t)u * NT;
      if (i < n) dst[i] = f(x[u], i);
    }
  }
}

// Finds, scanning from the top bin down, the bin b with
//   above(b) < k <= above(b) + hist[b]
// where above(b) = sum of hist over bins > b.  Result in out[0..2] =
// {b, above(b), hist[b]}; out[0] = 0xffffffff if sum(hist) < k; sum(hist)
// in *total_out if given.  Thread t owns a contiguous run of bins counted
// from the top, so an exclusive prefix scan over threads yields above().
template <int NT>
RK_DEV void block_find_crossing(const uint32_t* hist, int nbins, uint32_t k,
                                uint32_t* scratch, uint32_t* out,
                                uint32_t* total_out = nullptr) {
  if (threadIdx.x == 0) out[0] = 0xffffffffu;
  const int per = (nbins + NT - 1) / NT;
  const int hi = nbins - (int)threadIdx.x * per;
  uint32_t s = 0;
  for (int j = 0; j < per; ++j) {
    const int b = hi - 1 - j;
    if (b >= 0) s += hist[b];
  }
  uint32_t total;
  const uint32_t above = block_excl_scan<NT>(s, scratch, total);  // has barriers
  if (total_out && threadIdx.x == 0) *total_out = total;
  if (s != 0 && above < k && k <= above + s) {
    uint32_t a = above;
    for (int j = 0; j < per; ++j) {
      const int b = hi - 1 - j;
      if (b < 0) break;
      const uint32_t c = hist[b];
      if (k <= a + c) {
        out[0] = (uint32_t)b;
        out[1] = a;
        out[2] = c;
        break;
      }
      a += c;
    }
  }
  __syncthreads();
}

// Workgroup radix select: returns a threshold K such that
//   #{i : keys[i] >= K} >= k  and  #{i : keys[i] > K'} < k for K' >= K in
// the selected bin, i.e. keys >= K are exactly the k largest when the keys
// are distinct (composite keys), and at least the k largest otherwise;
// out[3] receives #{i : keys[i] >= K}.  `keys` points to global memory
// (GLOBAL) or LDS.  hist: LDS, 1 << HB words.  The scan starts below the
// common prefix of all keys (the key-space form of RadiK's adaptive
// scaling): bits every key shares are never histogrammed.  This form takes
// the bits that vary (diff) and one key (ref) from its caller.  Loads are
// issued UN at a time before their LDS atomics.  COH: global keys written
// in the same launch by other workgroups, read with device-coherent loads.
template <int NT, int HB, typename KeyT, bool GLOBAL, bool COH = false>
RK_DEV KeyT block_radix_select_from(const KeyT* keys, uint32_t n, uint32_t k, uint64_t diff,
                                    KeyT ref, uint32_t* hist, uint32_t* scratch,
                                    uint32_t* out) {
  constexpr int BITS = (int)sizeof(KeyT) * 8;
  constexpr int UN = 4;
  auto ld = [&](uint32_t i) -> KeyT {
    if constexpr (GLOBAL && COH) return ld_dev(keys + i);
    else if constexpr (GLOBAL) return ldg(keys + i);
    else return ld_lds(keys + i);
  };
  if (diff == 0) {  // all keys equal: every key is >= ref
    if (threadIdx.x == 0) out[3] = n;
    __syncthreads();
    return ref;
  }
  int lo = 64 - __builtin_clzll(diff);  // number of low bits still unresolved
  KeyT prefix = (lo >= BITS) ? (KeyT)0 : (KeyT)((ref >> lo) << lo);
  uint32_t rem = k, cnt = 0;
  while (lo > 0) {
    const int w = lo < HB ? lo : HB;
    const int shift = lo - w;
    const uint32_t nb = 1u << w;
    for (uint32_t i = threadIdx.x; i < nb; i += NT) hist[i] = 0;
    __syncthreads();
    for (uint32_t i0 = threadIdx.x; i0 < n; i0 += NT * UN) {
      KeyT x[UN];
#pragma unroll
      for (int u = 0; u < UN; ++u) {
        const uint32_t i = i0 + u * NT;
        x[u] = i < n ? ld(i) : (KeyT)0;
      }
#pragma unroll
      for (int u = 0; u < UN; ++u) {
        const uint32_t i = i0 + u * NT;
        const bool match = (lo >= BITS) || ((x[u] >> lo) == (prefix >> lo));
        if (i < n && match) atomicAdd(&hist[(uint32_t)(x[u] >> shift) & (nb - 1)], 1u);
      }
    }
    __syncthreads();
    block_find_crossing<NT>(hist, (int)nb, rem, scratch, out);
    const uint32_t b = out[0], above = out[1];
    cnt = out[2];
    __syncthreads();
    if (b == 0xffffffffu) {  // cannot happen when n >= k
      if (threadIdx.x == 0) out[3] = 0;
      __syncthreads();
      return prefix;
    }
    prefix |= (KeyT)b << shift;
    rem -= above;
    lo = shift;
    if (cnt == rem) break;  // the whole bin is taken
  }
  // keys above the final bin: k - rem; keys in it (all >= the threshold): cnt
  if (threadIdx.x == 0) out[3] = k - rem + cnt;
  __syncthreads();
  return prefix;
}

template <int NT, int HB, typename KeyT, bool GLOBAL, bool COH = false>
RK_DEV KeyT block_radix_select(const KeyT* keys, uint32_t n, uint32_t k, uint32_t* hist,
                               uint32_t* scratch, uint64_t* scratch64, uint32_t* out) {
  constexpr int UN = 4;
  auto ld = [&](uint32_t i) -> KeyT {
    if constexpr (GLOBAL && COH) return ld_dev(keys + i);
    else if constexpr (GLOBAL) return ldg(keys + i);
    else return ld_lds(keys + i);
  };
  const KeyT k0 = ld(0);
  uint64_t diff = 0;
  for (uint32_t i0 = threadIdx.x; i0 < n; i0 += NT * UN) {
    KeyT x[UN];
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      const uint32_t i = i0 + u * NT;
      x[u] = i < n ? ld(i) : k0;
    }
#pragma unroll
    for (int u = 0; u < UN; ++u) diff |= (uint64_t)(x[u] ^ k0);
  }
  diff = block_or64<NT>(diff, scratch64);
  return block_radix_select_from<NT, HB, KeyT, GLOBAL, COH>(keys, n, k, diff, k0, hist, scratch,
                                                            out);
}

// Descending bitonic sort of P2 (power of two) 64-bit keys held in LDS.
// The keys are spread over as few waves as possible: thread t owns E
// consecutive keys in registers (nthr = P2 / E threads, whole waves except
// when P2 < 64 E).  A compare-exchange at distance `stride` runs
//   in registers      when stride < E,
//   by lane shuffles  when the partner is in the same wave (stride / E < 64),
//   through LDS       otherwise (the only stages with a workgroup barrier).
// Waves holding no keys skip the register and shuffle stages, so a sort of
// up to 64 E keys is done by one wave with no barrier at all.
template <int NT, int E>
RK_DEV void bitonic_sort_e(uint64_t* a, uint32_t P2) {
  const uint32_t t = threadIdx.x;
  const uint32_t nthr = P2 / E;  // threads holding keys
  const bool wave_busy = (t & ~63u) < nthr;
  uint64_t v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = t < nthr ? a[t * E + e] : 0ull;
  for (uint32_t size = 2; size <= P2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride < (uint32_t)E) {
        // compile-time register indices only (a runtime index would move v[]
        // to scratch memory)
#pragma unroll
        for (int sc = E / 2; sc >= 1; sc >>= 1) {
          if (stride == (uint32_t)sc) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
              if (!(e & sc)) {
                const uint32_t i = t * E + e;
                const uint64_t x = v[e], y = v[e + sc];
                const bool desc = !(i & size);
                const bool sw = desc ? (x < y) : (x > y);
                v[e] = sw ? y : x;
                v[e + sc] = sw ? x : y;
              }
            }
          }
        }
      } else {
        const uint32_t ls = stride / E;  // distance in threads
        if (ls >= 64) {
#pragma unroll
          for (int e = 0; e < E; ++e)
            if (t < nthr) a[t * E + e] = v[e];
          __syncthreads();
        }
        if (ls >= 64 || wave_busy) {
#pragma unroll
          for (int e = 0; e < E; ++e) {
            uint64_t y;
            if (ls < 64) y = __shfl_xor(v[e], (int)ls, 64);
            else y = (t < nthr) ? a[(t ^ ls) * E + e] : 0ull;
            const uint32_t i = t * E + e;
            const bool lower = !(i & stride);
            const bool desc = !(i & size);
            const bool keep_max = (lower == desc);
            v[e] = keep_max ? (v[e] > y ? v[e] : y) : (v[e] < y ? v[e] : y);
          }
        }
        if (ls >= 64) __syncthreads();
      }
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (t < nthr) a[t * E + e] = v[e];
  __syncthreads();
}

// Descending bitonic sort of NT * E 64-bit keys held in registers, thread t
// holding keys [t E, t E + E) (NT a multiple of 64).  A stage at distance
// `stride` runs in registers when stride < E, by lane shuffles when the
// partner thread is in the same wave (stride / E < 64), and through `xs`
// (LDS, NT * E keys, laid out key-major so that a wave's accesses fall in
// distinct banks) otherwise: of the 78 stages of 4096 keys in 256 x 16,
// three cross waves.  No barrier except around those stages.
template <int NT, int E>
RK_DEV void bitonic_sort_regs(uint64_t (&v)[E], uint64_t* xs) {
  constexpr uint32_t N = (uint32_t)NT * E;
  const uint32_t t = threadIdx.x;
  for (uint32_t size = 2; size <= N; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride < (uint32_t)E) {
#pragma unroll
        for (int sc = E / 2; sc >= 1; sc >>= 1) {
          if (stride == (uint32_t)sc) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
              if (!(e & sc)) {
                const bool desc = !((t * E + e) & size);
                const uint64_t x = v[e], y = v[e + sc];
                const bool sw = desc ? (x < y) : (x > y);
                v[e] = sw ? y : x;
                v[e + sc] = sw ? x : y;
              }
            }
          }
        }
      } else {
        // All E partner keys are fetched before any is used, so the E
        // exchanges of a stage are in flight together.
        const uint32_t ls = stride / E;  // distance in threads
        uint64_t y[E];
        if (ls < 64) {
#pragma unroll
          for (int e = 0; e < E; ++e) y[e] = __shfl_xor(v[e], (int)ls, 64);
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e) xs[(uint32_t)e * NT + t] = v[e];
          __syncthreads();
#pragma unroll
          for (int e = 0; e < E; ++e) y[e] = xs[(uint32_t)e * NT + (t ^ ls)];
          __syncthreads();
        }
        // i & stride and i & size are the same for all E keys of a thread
        // here (stride >= E), so one direction per thread
        const uint32_t i0 = t * E;
        const bool keep_max = (!(i0 & stride)) == (!(i0 & size));
#pragma unroll
        for (int e = 0; e < E; ++e)
          v[e] = keep_max ? (v[e] > y[e] ? v[e] : y[e]) : (v[e] < y[e] ? v[e] : y[e]);
      }
    }
  }
}

// Compare-exchange of x (lower index) and y in a bitonic stage.
RK_DEV void bitonic_cas(uint64_t& x, uint64_t& y, bool desc) {
  if (desc ? (x < y) : (x > y)) {
    const uint64_t t = x;
    x = y;
    y = t;
  }
}

// LDS bitonic sort for the largest sizes, where keeping E keys per thread in
// registers would spill.  Stages go through LDS two at a time: a thread
// loads the four keys b, b+h, b+2h, b+3h (b with bits h and 2h clear) and
// applies the stage at distance 2h and then the one at h in registers, so
// every key is read and written once per two stages, with one barrier.  A
// size with an odd number of stages ends with a single stage at distance 1.
template <int NT>
RK_DEV void bitonic_sort_lds(uint64_t* a, uint32_t P2) {
  for (uint32_t size = 2; size <= P2; size <<= 1) {
    uint32_t stride = size >> 1;
    for (; stride >= 2; stride >>= 2) {
      const uint32_t h = stride >> 1;
      const int lh = __builtin_ctz(h);
      for (uint32_t g = threadIdx.x; g < (P2 >> 2); g += NT) {
        const uint32_t b = (g & (h - 1)) | ((g >> lh) << (lh + 2));
        const bool desc = (b & size) == 0;  // the four keys share it: h, 2h < size
        uint64_t x0 = a[b], x1 = a[b + h], x2 = a[b + 2 * h], x3 = a[b + 3 * h];
        bitonic_cas(x0, x2, desc);
        bitonic_cas(x1, x3, desc);
        bitonic_cas(x0, x1, desc);
        bitonic_cas(x2, x3, desc);
        a[b] = x0;
        a[b + h] = x1;
        a[b + 2 * h] = x2;
        a[b + 3 * h] = x3;
      }
      __syncthreads();
    }
    if (stride == 1) {
      for (uint32_t i = threadIdx.x; i < (P2 >> 1); i += NT) {
        const uint32_t lo = 2 * i;
        const bool desc = (lo & size) == 0;
        uint64_t x = a[lo], y = a[lo + 1];
        bitonic_cas(x, y, desc);
        a[lo] = x;
        a[lo + 1] = y;
      }
      __syncthreads();
    }
  }
}

// Measured on gfx950 (one workgroup per CU): a single wave sorts up to 64
// keys faster with shuffles than the whole workgroup does through LDS; from
// 128 keys on, the dependent shuffle chain loses to LDS stages, whose
// barriers the other waves' work overlaps.
template <int NT>
RK_DEV void block_bitonic_desc(uint64_t* a, uint32_t P2) {
  if (P2 <= 64) bitonic_sort_e<NT, 1>(a, P2);
  else bitonic_sort_lds<NT>(a, P2);
}

// Descending sort of m distinct 64-bit keys a[0, m) (LDS), using tmp[0, m)
// (LDS, disjoint from a): one most-significant-digit pass groups the keys
// by bucket into tmp, then each key goes back to a at its bucket's start
// plus the number of the bucket's keys above it.  Registers hold no keys
// across the passes, so the sort costs the calling kernel no occupancy.
//
// Buckets follow the distribution of the keys' squeezed images.  The
// squeezed image of a key is its high 32-bit half without the bits that
// every key shares there, followed by its low half without its shared bits.
// Both halves compare as unsigned integers, high half first, so squeezing
// keeps the order; it removes the constant top bits of the position half of
// a composite key (value key << 32 | ~position; rows are shorter than 2^32),
// so the copies of a repeated value spread over buckets by position when
// the values are close in key space.  The images' range is cut linearly
// into 256 coarse bins; each coarse bin gets fine buckets in proportion to
// its keys, and a key's fine bucket in its coarse bin is linear in its
// offset inside the bin.  So the buckets adapt to clustered keys (the
// float keys of a row sit in a few exponents, in two ranges far apart
// when both signs occur) where a single linear or top-bits cut would leave
// most buckets empty.
//
// Ranks are counted in grouped order, so the 64 lanes of a wave mostly
// share one bucket, its loop and its (broadcast) LDS reads: about
// m / 64 + sum over buckets of size^2 / 64 wave steps, where a bitonic
// network takes log^2 m / 2 dependent stages over all m keys.  Every pass
// issues a thread's LDS loads four at a time, before any store.  Returns
// false, keys untouched, when the sum of squared bucket sizes exceeds
// BUCKET_RANK_BUDGET * m (a few values far apart in key space, each
// repeated many times; or a far outlier): the caller sorts another way.
// m <= BUCKET_SORT_MAX.  bins: 2048 words of LDS; ctab: 256 words of LDS;
// scratch64: 3 NT / 64 words; scratch: NT / 64 words.
constexpr uint32_t BUCKET_RANK_BUDGET = 128;
constexpr uint32_t BUCKET_SORT_MAX = 4096;

template <int NT>
RK_DEV bool bucket_sort_desc(uint64_t* a, uint64_t* tmp, uint32_t m, uint32_t* bins,
                             uint32_t* ctab, uint64_t* scratch64, uint32_t* scratch) {
  constexpr int NB = 2048, NC = 256, PER = NB / NT, UN = 4;
  static_assert(NB % NT == 0, "buckets per thread");
  const uint32_t t = threadIdx.x;
  // Loads UN keys per thread (keys past m read a[0], within the keys'
  // range; callers mask them out by index).
  auto load4 = [&](const uint64_t* src, uint32_t i0, uint64_t (&x)[UN]) {
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      const uint32_t i = i0 + (uint32_t)u * NT;
      x[u] = src[i < m ? i : 0u];
    }
  };
  // One pass: the bits that vary, and the smallest and largest key.  The
  // squeezed image is monotone in the key, so the images' range is the
  // images of these two.
  const uint64_t x0 = a[0];
  uint64_t diff = 0, kmin = ~0ull, kmax = 0;
  for (uint32_t i0 = t; i0 < m; i0 += NT * UN) {
    uint64_t x[UN];
    load4(a, i0, x);
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      diff |= x[u] ^ x0;
      kmin = x[u] < kmin ? x[u] : kmin;
      kmax = x[u] > kmax ? x[u] : kmax;
    }
  }
  for (uint32_t i = t; i < (uint32_t)NB; i += NT) bins[i] = 0;
  for (uint32_t i = t; i < (uint32_t)NC; i += NT) ctab[i] = 0;
  // its barriers also order the clearing
  block_or_min_max64<NT>(diff, kmin, kmax, scratch64);
  const uint32_t dh = (uint32_t)(diff >> 32), dl = (uint32_t)diff;
  const int kb = dh ? 32 - __builtin_clz(dh) : 0;  // varying bits of the high half
  const int pb = dl ? 32 - __builtin_clz(dl) : 0;  // and of the low half
  const uint64_t kmask = (1ull << kb) - 1ull, pmask = (1ull << pb) - 1ull;
  auto squeeze = [&](uint64_t x) { return (((x >> 32) & kmask) << pb) | (x & pmask); };
  const uint64_t lo = squeeze(kmin), hi = squeeze(kmax);
  // Position of an image in [lo, hi] as the 64-bit fixed-point number
  // (d >> s) * mul = coarse bin (high 32 bits) . fraction (low 32 bits),
  // mul = floor(NC 2^32 / r), r = (span >> s) + 1: the offset is cut to its
  // top 32 bits (s), so r <= 2^32 and the product stays below NC 2^32;
  // monotone in the image.  A span below NB takes the offset as the bucket.
  const uint64_t span = hi - lo;
  const bool direct = span < (uint64_t)NB;
  const int bl = span ? 64 - __builtin_clzll(span) : 0;
  const int s = bl > 32 ? bl - 32 : 0;
  // mul = floor(NC 2^32 / r): a double quotient, corrected to the floor by
  // exact integer products (mul * r <= NC 2^32 + r < 2^64)
  uint64_t mul = 0;
  if (!direct) {
    const uint64_t r = (span >> s) + 1ull, num = (uint64_t)NC << 32;
    mul = (uint64_t)((double)num / (double)r);
    while (mul * r > num) --mul;
    while ((mul + 1) * r <= num) ++mul;
  }
  auto fixed = [&](uint64_t x) { return ((squeeze(x) - lo) >> s) * mul; };
  if (!direct) {
    for (uint32_t i0 = t; i0 < m; i0 += NT * UN) {
      uint64_t x[UN];
      load4(a, i0, x);
#pragma unroll
      for (int u = 0; u < UN; ++u)
        if (i0 + (uint32_t)u * NT < m) atomicAdd(&ctab[(uint32_t)(fixed(x[u]) >> 32)], 1u);
    }
    __syncthreads();
    // fine buckets of coarse bin c: 1 + floor(n_c (NB - NC) / m) when it has
    // keys (at most NB in all), packed as first bucket | count << 12
    if (t < 64) {
      uint32_t f[NC / 64], tot = 0;
#pragma unroll
      for (int j = 0; j < NC / 64; ++j) {
        const uint32_t n = ctab[t * (NC / 64) + j];
        f[j] = n ? 1u + n * (uint32_t)(NB - NC) / m : 0u;
        tot += f[j];
      }
      uint32_t run = wave_incl_scan(tot) - tot;
#pragma unroll
      for (int j = 0; j < NC / 64; ++j) {
        ctab[t * (NC / 64) + j] = run | (f[j] << 12);
        run += f[j];
      }
    }
    __syncthreads();
  }
  auto bucket = [&](uint64_t x) -> uint32_t {
    if (direct) return (uint32_t)(squeeze(x) - lo);
    const uint64_t fx = fixed(x);
    const uint32_t e = ctab[(uint32_t)(fx >> 32)];
    return (e & 0xfffu) + (uint32_t)(((uint64_t)(uint32_t)fx * (e >> 12)) >> 32);
  };
  for (uint32_t i0 = t; i0 < m; i0 += NT * UN) {
    uint64_t x[UN];
    load4(a, i0, x);
#pragma unroll
    for (int u = 0; u < UN; ++u)
      if (i0 + (uint32_t)u * NT < m) atomicAdd(&bins[bucket(x[u])], 1u);
  }
  __syncthreads();
  // Bucket starts in descending bucket order: thread t scans buckets
  // NB-1-t*PER down to NB-PER-t*PER; the squared sizes decide the method.
  // One 64-bit scan carries the key counts (low half, at most m) and the
  // squared sizes (high half, at most m^2), whose total is all it needs.
  uint32_t c[PER], tot = 0, sq = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    c[j] = bins[NB - 1 - (t * PER + j)];
    tot += c[j];
    sq += c[j] * c[j];  // <= m^2 < 2^32
  }
  uint64_t all;
  const uint64_t pre = block_excl_scan64<NT>(((uint64_t)sq << 32) | tot, scratch64, all);
  if ((uint32_t)(all >> 32) > BUCKET_RANK_BUDGET * m) return false;  // uniform
  uint32_t run = (uint32_t)pre;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    bins[NB - 1 - (t * PER + j)] = run;  // start: the bucket's cursor
    run += c[j];
  }
  __syncthreads();
  // grouped by bucket into tmp (any order inside one)
  for (uint32_t i0 = t; i0 < m; i0 += NT * UN) {
    uint64_t x[UN];
    load4(a, i0, x);
    uint32_t bk[UN];
#pragma unroll
    for (int u = 0; u < UN; ++u) bk[u] = bucket(x[u]);
#pragma unroll
    for (int u = 0; u < UN; ++u)
      if (i0 + (uint32_t)u * NT < m) tmp[atomicAdd(&bins[bk[u]], 1u)] = x[u];
  }
  __syncthreads();
  // bins[b] is now the end of bucket b, which bucket b - 1 starts at.  Each
  // thread ranks the keys at grouped positions t, t + NT, ... and stores
  // them at their place in a.
  for (uint32_t i0 = t; i0 < m; i0 += NT * UN) {
    uint64_t x[UN];
    load4(tmp, i0, x);
    uint32_t st[UN], en[UN];
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      const uint32_t bk = bucket(x[u]);
      st[u] = bk == NB - 1 ? 0u : bins[bk + 1];
      en[u] = i0 + (uint32_t)u * NT < m ? bins[bk] : st[u];  // no keys to rank against
    }
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      uint32_t r = st[u];
      for (uint32_t j = st[u]; j < en[u]; ++j) r += tmp[j] > x[u] ? 1u : 0u;
      if (i0 + (uint32_t)u * NT < m) a[r] = x[u];
    }
  }
  __syncthreads();
  return true;
}

RK_DEV uint32_t next_pow2(uint32_t x) {
  return x <= 1 ? 1u : (1u << (32 - __builtin_clz(x - 1)));
}

}  // namespace rk
