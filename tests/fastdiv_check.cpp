// Checks lbm::make_fastdiv / lbm::fastdiv (the kernels' division of node indices by nx and ny).
// Usage: fastdiv_check FULL_DIVISOR...  Exhaustive ranges are checked against a quotient counted
// up one remainder at a time; sampled indices against the hardware divide.
//  - every divisor 1..4096 (all lattice widths and heights the tests and serving mixes use, and
//    more), each over n in [0, 2^18) and over 2^16 hashed n spread across all 32-bit values;
//  - each FULL_DIVISOR over every 32-bit n.
// Prints the first mismatch and exits 1, else prints the count of checks and exits 0.
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "lbm_fastdiv.hpp"

static bool check(uint32_t d, uint32_t n, const uint32_t c[3], uint32_t want) {
  const uint32_t q = lbm::fastdiv(n, c[0], c[1], c[2]);
  if (q != want) {
    std::printf("MISMATCH d=%u n=%u fastdiv=%u want=%u\n", d, n, q, want);
    return false;
  }
  return true;
}
static bool check(uint32_t d, uint32_t n, const uint32_t c[3]) { return check(d, n, c, n / d); }

// Every n in [0, end) (end = 0: all 2^32), against the counted quotient.
static bool sweep(uint32_t d, uint32_t end, const uint32_t c[3]) {
  uint32_t q = 0, r = 0, n = 0;
  do {
    if (!check(d, n, c, q)) return false;
    if (++r == d) { r = 0; ++q; }
  } while (++n != end);
  return true;
}

int main(int argc, char **argv) {
  uint64_t checks = 0;
  for (uint32_t d = 1; d <= 4096; ++d) {
    uint32_t c[3];
    lbm::make_fastdiv(d, c);
    if (!sweep(d, 1u << 18, c)) return 1;
    uint32_t h = d * 0x9E3779B9u;
    for (int k = 0; k < (1 << 16); ++k) {
      h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
      if (!check(d, h, c) || !check(d, 0xFFFFFFFFu - (uint32_t)k, c)) return 1;
    }
    checks += (1u << 18) + 2u * (1u << 16);
  }
  for (int a = 1; a < argc; ++a) {
    const uint32_t d = (uint32_t)std::strtoul(argv[a], nullptr, 10);
    uint32_t c[3];
    lbm::make_fastdiv(d, c);
    if (!sweep(d, 0, c)) return 1;
    checks += 1ull << 32;
  }
  std::printf("ok %llu\n", (unsigned long long)checks);
  return 0;
}
