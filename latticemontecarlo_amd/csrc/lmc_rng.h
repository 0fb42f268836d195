// Counter-based random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11). Every draw is a pure function of
// (seed, counter, replica, tag), so a replica's stream does not depend on how replicas are
// batched, and a run resumed from a stored step counter continues the same stream.
//
// Streams:
//   KMC step s of replica r:      counter (s, r, 0)            -> u1 in (0,1] (time), u2 in [0,1)
//   CMC trial t, pair attempt a:  counter (t, r, 2a+1)         -> two site indices in [0, N)
//   CMC trial t, Metropolis:      counter (t, r, 0xFFFFFFFE)   -> u in [0,1)
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define LMC_HD __host__ __device__ __forceinline__
#else
#define LMC_HD inline
#endif

namespace lmc {

struct U4 {
  uint32_t x, y, z, w;
};

LMC_HD U4 Philox4x32_10(uint64_t seed, uint64_t counter, uint32_t replica, uint32_t tag) {
  uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
  uint32_t c0 = static_cast<uint32_t>(counter), c1 = static_cast<uint32_t>(counter >> 32);
  uint32_t c2 = replica, c3 = tag;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
    c1 = static_cast<uint32_t>(p1);
    c3 = static_cast<uint32_t>(p0);
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return {c0, c1, c2, c3};
}

LMC_HD uint64_t Join64(uint32_t lo, uint32_t hi) {
  return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}
// 53-bit uniforms.
LMC_HD double UnitOpenLeft(uint64_t b) { return static_cast<double>((b >> 11) + 1) * 0x1.0p-53; }
LMC_HD double UnitOpenRight(uint64_t b) { return static_cast<double>(b >> 11) * 0x1.0p-53; }
// floor(b * n / 2^64): an index in [0, n).
LMC_HD uint64_t MulHi64(uint64_t b, uint64_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(b, n);
#else
  return static_cast<uint64_t>((static_cast<unsigned __int128>(b) * n) >> 64);
#endif
}

}  // namespace lmc
