// sr_probe.hip -- runs v_cvt_sr_f16_f32 and v_cvt_sr_bf16_f32 on 2^20 (x, r) pairs (one pair per
// thread, the grid sized exactly to the arrays) and writes x, r and both results to the file named
// by argv[1]; make_sr_fixture.py turns that into the fixture sr_conversions.npz.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <vector>
#include <random>
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef __bf16 b2 __attribute__((ext_vector_type(2)));
__global__ void k(const float *x, const uint32_t *r, uint16_t *oh, uint16_t *ob, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const h2 h = __builtin_amdgcn_cvt_sr_f16_f32((h2){0, 0}, x[i], r[i], false);
  const b2 b = __builtin_amdgcn_cvt_sr_bf16_f32((b2){0, 0}, x[i], r[i], false);
  oh[i] = __builtin_bit_cast(uint16_t, h[0]);
  ob[i] = __builtin_bit_cast(uint16_t, b[0]);
}
int main(int argc, char **argv) {
  const int n = 1 << 20;
  std::vector<float> x(n);
  std::vector<uint32_t> r(n);
  std::mt19937_64 g(12345);
  for (int i = 0; i < n; ++i) {
    // exponents spread over half-normal, half-subnormal and tiny ranges, both signs
    const int kind = i % 8;
    float e;
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    if (kind < 4) e = std::ldexp(1.0f + u(g), (int)(g() % 30) - 14);        // 2^-14 .. 2^15
    else if (kind < 6) e = std::ldexp(1.0f + u(g), (int)(g() % 12) - 26);   // half subnormal range
    else if (kind == 6) e = std::ldexp(1.0f + u(g), (int)(g() % 40) - 60);  // far below
    else e = (float)(int)(g() % 4096) * 0x1p-24f;                            // multiples of 2^-24
    x[i] = (g() & 1) ? -e : e;
    r[i] = (uint32_t)g();
  }
  x[0] = 0.0f; x[1] = -0.0f; x[2] = 65504.0f; x[3] = 65519.0f; x[4] = 65520.0f; x[5] = 1e30f;
  float *dx; uint32_t *dr; uint16_t *dh, *db;
  if (hipMalloc(&dx, n * 4) || hipMalloc(&dr, n * 4) || hipMalloc(&dh, n * 2) || hipMalloc(&db, n * 2)) return 2;
  if (hipMemcpy(dx, x.data(), n * 4, hipMemcpyHostToDevice) || hipMemcpy(dr, r.data(), n * 4, hipMemcpyHostToDevice))
    return 2;
  k<<<n / 256, 256>>>(dx, dr, dh, db, n);
  if (hipDeviceSynchronize() != hipSuccess) return 3;
  std::vector<uint16_t> oh(n), ob(n);
  if (hipMemcpy(oh.data(), dh, n * 2, hipMemcpyDeviceToHost) || hipMemcpy(ob.data(), db, n * 2, hipMemcpyDeviceToHost))
    return 4;
  FILE *f = fopen(argv[1], "wb");
  fwrite(x.data(), 4, n, f); fwrite(r.data(), 4, n, f); fwrite(oh.data(), 2, n, f); fwrite(ob.data(), 2, n, f);
  fclose(f);
  printf("srprobe wrote %d cases\n", n);
  return 0;
}
