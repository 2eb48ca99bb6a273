// Philox4x32-10 (Salmon et al., SC'11) and the particle stream layout.
//
// key = (seed_lo, seed_hi); counter = (sid_lo, sid_hi, index, purpose).
// A particle's stream id (sid) is 64 bits; children and dark-process streams
// derive theirs from the parent stream (PURPOSE_CHILD / PURPOSE_DARK), so a
// shower's random numbers do not depend on launch shapes or queue order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace petite {

enum Purpose : uint32_t {
  PURPOSE_STEP = 0,        // substep k: x hard test, y -> U(6,20), z MCS radius, w MCS azimuth (sign: bit 0 of w)
  PURPOSE_CLOSE = 2,       // closing increment: x fraction, y MCS radius, z MCS azimuth (sign: bit 0 of z)
  PURPOSE_INTERACT = 3,    // u0 process choice, u1 kinematic azimuth, u2/u3 extra
  PURPOSE_VEGAS_A = 4,     // candidate j: y0..y3 (ndim <= 3: u3 = accept uniform; ndim 4: low bits)
  PURPOSE_CHILD = 6,       // index k: child stream id
  PURPOSE_DARK = 7,        // index = dark process code: dark stream id
  PURPOSE_BOUND_E = 8,     // bound-electron momentum accept/reject pairs
  PURPOSE_MAP_TRAIN = 9,   // map training sample j of (process, energy): y0..y3
};

struct U4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

__host__ __device__ inline U4 philox10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__host__ __device__ inline U4 draw_block(uint64_t seed, uint64_t sid, uint32_t index, uint32_t purpose) {
  return philox10(U4{(uint32_t)sid, (uint32_t)(sid >> 32), index, purpose}, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__host__ __device__ inline uint64_t derive_sid(uint64_t seed, uint64_t sid, uint32_t index, uint32_t purpose) {
  U4 o = draw_block(seed, sid, index, purpose);
  return (uint64_t)o.x | ((uint64_t)o.y << 32);
}

// u in (0,1), exactly representable in float32: ((x >> 9) + 0.5) * 2^-23
__host__ __device__ inline float u01f(uint32_t x) { return ((float)(x >> 9) + 0.5f) * 0x1.0p-23f; }
__host__ __device__ inline double u01d(uint32_t x) { return ((double)(x >> 9) + 0.5) * 0x1.0p-23; }

}  // namespace petite
