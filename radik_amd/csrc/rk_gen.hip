// Seeded input generators of the benchmark harness (the counterpart of
// radik/test/test_radik.cu:22-55 and :89-124, and radik/test/zipf.hpp:8-46).
//
// Every element is a pure function of (seed, stream, index): a counter-based
// SplitMix64 stream (Steele, Lea & Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014: state += golden gamma, output mixed by the
// variant-13 finalizer), so a row can be generated anywhere, in any order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/radik_amd.h"
#include "../../include/radik_amd_harness.h"

namespace {

constexpr uint64_t GAMMA = 0x9E3779B97F4A7C15ull;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// The j-th 64-bit draw of element i.
__device__ inline uint64_t draw(uint64_t base, uint64_t i, uint64_t j) {
  return mix64(base + (i * 64 + j + 1) * GAMMA);
}

__device__ inline double u53(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

// Zipf(N, s) by rejection sampling from a continuous bound (the method the
// reference's zipf.hpp:8-46 implements; J. Crease, "Rejection sampling the
// Zipf distribution"), with at most 31 attempts (then rank 1).
__device__ inline double zipf_rank(uint64_t base, uint64_t i, double N, double s) {
  const double t = (pow(N, 1.0 - s) - s) / (1.0 - s);
  for (int a = 0; a < 31; ++a) {
    const double p = u53(draw(base, i, 2 * a + 1));
    const double y = u53(draw(base, i, 2 * a + 2));
    const double pt = p * t;
    const double invB = pt <= 1.0 ? pt : pow(pt * (1.0 - s) + s, 1.0 / (1.0 - s));
    const double x = floor(invB + 1.0);
    const double top = pow(x, -s);
    const double bottom = x <= 1.0 ? 1.0 / t : pow(invB, -s) / t;
    if (y < top / (bottom * t)) return x > N ? N : x;
  }
  return 1.0;
}

__device__ inline float base_value(int kind, uint64_t base, uint64_t i, double a, double b) {
  switch (kind) {
    case RK_GEN_UNIFORM:  // U[0,1) on the 2^-24 grid: exact in fp32
      return (float)((double)(draw(base, i, 0) >> 40) * 0x1.0p-24);
    case RK_GEN_RANGE:  // U[a, b)
      return (float)(a + (b - a) * u53(draw(base, i, 0)));
    case RK_GEN_NORMAL: {  // Box-Muller in double (a = mean, b = stddev)
      const double u1 = ((double)(draw(base, i, 0) >> 11) + 1.0) * 0x1.0p-53;
      const double u2 = u53(draw(base, i, 1));
      return (float)(a + b * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
    case RK_GEN_ZIPF:  // rank in [1, b] with exponent a, mapped to rank / b
      return (float)(zipf_rank(base, i, b, a) / b);
    case RK_GEN_NEARCONST:  // 1 + m * 2^-23, m uniform in [0, 2^13)
      return 1.0f + (float)(draw(base, i, 0) >> 51) * 0x1.0p-23f;
    default:  // RK_GEN_CONST
      return (float)a;
  }
}

__device__ inline uint32_t to_bits(int dtype, float f) {
  const uint32_t u = __float_as_uint(f);
  if (dtype == RK_F32) return u;
  if (dtype == RK_BF16) {  // round to nearest even
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  }
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)f);
}

__global__ void k_gen(int dtype, int kind, uint64_t base, int64_t n, int64_t row_len, double a,
                      double b, double dup, void* out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += stride) {
    // duplicates are drawn within the element's own row
    const int64_t row0 = row_len > 0 ? (g / row_len) * row_len : 0;
    const uint64_t local = (uint64_t)(g - row0);
    uint64_t src = (uint64_t)g;
    if (dup > 0.0 && local > 0) {
      const uint64_t h = draw(base, (uint64_t)g, 63);
      if (u53(h) < dup) src = (uint64_t)row0 + mix64(h) % local;
    }
    const uint32_t bits = to_bits(dtype, base_value(kind, base, src, a, b));
    if (dtype == RK_F32) static_cast<uint32_t*>(out)[g] = bits;
    else static_cast<uint16_t*>(out)[g] = (uint16_t)bits;
  }
}

// Streaming-read probe: every workgroup reads one contiguous slice with
// 16-byte buffer loads (the load form of the top-k filter, 8 in flight per
// lane) and folds it into one word, so the kernel moves exactly `bytes` from
// HBM: the achievable read bandwidth the filter is compared against, and the
// known-size reference that scales the profiler's fetch counter.
__global__ __launch_bounds__(256) void k_probe_read(const uint4* p, int64_t nvec,
                                                    unsigned int* out) {
  // Workgroups take 256 KB blocks grid-stride, so the blocks read at any
  // moment lie close together; non-temporal loads (nothing is re-read); two
  // load sets in flight per wave (the next set is issued before the current
  // one is consumed).  The fastest read pattern measured on MI355X, at 2-3
  // workgroups per CU.
  constexpr int U = 2;
  constexpr int64_t BLK = 16384;  // vectors per block
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  const uint32_t wb = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t acc = 0;
  for (int64_t b0 = (int64_t)blockIdx.x * BLK; b0 < nvec; b0 += (int64_t)gridDim.x * BLK) {
    const uint32_t n = (uint32_t)((nvec - b0) < BLK ? (nvec - b0) : BLK);
    const uint64_t a = (uint64_t)(p + b0);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(((uint64_t)hi << 32) | lo), (short)0,
        (int)__builtin_amdgcn_readfirstlane(n * 16u), 0x00020000);
    // offsets past n * 16 bytes (at most 2 x 512 vectors beyond) read zeros
    v4u x[U], y[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      x[u] = __builtin_amdgcn_raw_buffer_load_b128(r, (int)((wb + lane + u * 256) * 16u), 0, 2);
    for (uint32_t jb = wb; jb < n; jb += 256 * U) {
#pragma unroll
      for (int u = 0; u < U; ++u)
        y[u] = __builtin_amdgcn_raw_buffer_load_b128(
            r, (int)((jb + 256 * U + lane + u * 256) * 16u), 0, 2);
#pragma unroll
      for (int u = 0; u < U; ++u) acc ^= x[u].x ^ x[u].y ^ x[u].z ^ x[u].w;
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = y[u];
    }
  }
  if (acc == 0x9E3779B9u) atomicXor(out, acc);  // keeps the loads live
}

}  // namespace

extern "C" int rk_probe_read(const void* data, int64_t bytes, int32_t blocks_per_cu, void* out4,
                             void* stream) {
  if (!data || !out4 || bytes < 0 || ((uintptr_t)data % 16) || blocks_per_cu < 0)
    return RK_E_ARG;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return RK_E_HIP;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return RK_E_HIP;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_probe_read, 256, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int bpc = blocks_per_cu == 0 ? 3 : (blocks_per_cu < per_cu ? blocks_per_cu : per_cu);
  const int64_t grid = (int64_t)cus * bpc;
  hipLaunchKernelGGL(k_probe_read, dim3((unsigned)grid), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const uint4*>(data), bytes / 16,
                     static_cast<unsigned int*>(out4));
  return hipGetLastError() == hipSuccess ? RK_OK : RK_E_HIP;
}

extern "C" int rk_gen_fill(int dtype, int kind, uint64_t seed, uint64_t stream_id, int64_t n,
                           int64_t row_len, double a, double b, double dup_frac, void* out,
                           void* stream) {
  if (dtype != RK_F32 && dtype != RK_F16 && dtype != RK_BF16) return RK_E_UNSUPPORTED;
  if (kind < 0 || kind > RK_GEN_CONST || n < 0 || !out) return RK_E_ARG;
  if (n == 0) return RK_OK;
  const uint64_t base = mix64(seed * GAMMA ^ mix64(stream_id + 0x632BE59BD9B4E019ull));
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_gen, dim3((unsigned)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     dtype, kind, base, n, row_len, a, b, dup_frac, out);
  return hipGetLastError() == hipSuccess ? RK_OK : RK_E_HIP;
}
