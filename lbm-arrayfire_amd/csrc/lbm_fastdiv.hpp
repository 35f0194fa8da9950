// lbm_fastdiv.hpp — exact unsigned 32-bit division by an invariant divisor through one high
// multiply, shared by the host (which computes the constants) and the kernels (which divide node
// indices by nx and ny).  Free of HIP headers, so a host-only program can include it.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define LBM_HOST_DEVICE __host__ __device__
#else
#define LBM_HOST_DEVICE
#endif

namespace lbm {

// Granlund & Montgomery, "Division by invariant integers using multiplication", 1994, fig. 4.1:
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, and for every 32-bit n
// n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0) with t = mulhi(m, n).
// out = {m, min(l, 1), max(l - 1, 0)}; 1 <= d < 2^31.
inline void make_fastdiv(uint32_t d, uint32_t out[3]) {
  int l = 0;
  while ((1ull << l) < (unsigned long long)d) ++l;
  out[0] = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
  out[1] = (uint32_t)(l < 1 ? l : 1);
  out[2] = (uint32_t)(l > 0 ? l - 1 : 0);
}

LBM_HOST_DEVICE inline uint32_t fastdiv(uint32_t n, uint32_t m, uint32_t s1, uint32_t s2) {
  const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
  return (t + ((n - t) >> s1)) >> s2;
}

}  // namespace lbm
