/*
 * radik_amd harness — seeded input generators used by the benchmark and the
 * parity tests (the counterpart of radik/test/test_radik.cu:22-55, :89-124
 * and radik/test/zipf.hpp:8-46, which use cuRAND and std::default_random_engine).
 * Every element is a pure function of (seed, stream_id, index).
 */
#ifndef RADIK_AMD_HARNESS_H_
#define RADIK_AMD_HARNESS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RK_GEN_UNIFORM = 0,   /* U[0,1) on the 2^-24 grid (exact in fp32) */
  RK_GEN_NORMAL = 1,    /* Box-Muller, mean a, stddev b */
  RK_GEN_ZIPF = 2,      /* Zipf(s = a) rank over [1, b], value rank / b */
  RK_GEN_NEARCONST = 3, /* 1 + m * 2^-23, m uniform in [0, 2^13) */
  RK_GEN_RANGE = 4,     /* U[a, b) */
  RK_GEN_CONST = 5      /* every value = a */
};

/* Fills n elements (dtype RK_F32 / RK_F16 / RK_BF16, device pointer `out`).
 * With dup_frac > 0, each element after the first of its row (rows of
 * row_len elements; row_len <= 0 means one row) is, with probability
 * dup_frac, a copy of a uniformly chosen earlier element of that row.
 * 16-bit outputs are the fp32 value rounded to nearest even. */
int rk_gen_fill(int dtype, int kind, uint64_t seed, uint64_t stream_id, int64_t n,
                int64_t row_len, double a, double b, double dup_frac, void* out,
                void* stream);

/* Streaming-read probe: reads `bytes` (rounded down to 16) from the
 * 16-byte-aligned device buffer `data` once, with `blocks_per_cu` persistent
 * workgroups per CU (0: 3; capped at what is resident), and writes nothing
 * but (rarely) one word to `out4`.  Timed around, the fastest of a few grids
 * gives the achievable HBM read bandwidth of the box. */
int rk_probe_read(const void* data, int64_t bytes, int32_t blocks_per_cu, void* out4,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RADIK_AMD_HARNESS_H_ */
