// lepic_philox.h -- Philox-4x32-10 counter-based generator (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), written
// from the published round function.  Each (key, counter) gives four
// independent 32-bit words; draws never depend on launch order.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

struct philox4 {
  uint32_t v[4];
};

__host__ __device__ inline philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                 uint32_t c3, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0; k1 += W1;
  }
  philox4 o;
  o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
  return o;
}

// Uniform double in [0,1) with 53 random bits from two 32-bit words.
__host__ __device__ inline double u01_53(uint32_t a, uint32_t b) {
  uint64_t x = ((uint64_t)a << 32) | (uint64_t)b;
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// A stream view: key = (k0,k1), counter prefix (c1,c2,c3), block index in c0.
// Each block of four words gives two uniforms, words (0,1) then (2,3); the
// second is kept as a double, so the stream holds no indexed word buffer
// (which the compiler would place in LDS or scratch).
struct PhiloxStream {
  uint32_t k0, k1, c1, c2, c3;
  uint32_t idx;
  double spare;
  bool has_spare;
  __host__ __device__ void init(uint64_t key, uint32_t a, uint32_t b, uint32_t c) {
    k0 = (uint32_t)key; k1 = (uint32_t)(key >> 32);
    c1 = a; c2 = b; c3 = c;
    idx = 0; has_spare = false; spare = 0.0;
  }
  __host__ __device__ double uniform() {
    if (has_spare) {
      has_spare = false;
      return spare;
    }
    philox4 r = philox4x32_10(idx++, c1, c2, c3, k0, k1);
    spare = u01_53(r.v[2], r.v[3]);
    has_spare = true;
    return u01_53(r.v[0], r.v[1]);
  }
};
