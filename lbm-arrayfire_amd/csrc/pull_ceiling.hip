// pull_ceiling.hip — the attainable rate of the D3Q27 (and D3Q19) step's memory access pattern,
// measured with no arithmetic.  Every node of a 256^3 lattice pulls its 27 populations from x - c_i (periodic)
// and stores them, in the library's padded SoA layout (population_stride), with 4-byte and 2-byte
// populations (D3Q19: its 19 directions, non-temporal stores); the same 27 arrays copied node by
// node without the neighbour shifts (the stream count alone); and a plain copy of the same number
// of bytes, 16 bytes per lane, one element per thread over a full grid (the device's streaming
// rate: 6.2-6.3 TB/s on MI355X, where one 4-byte element per lane reaches 3.9 TB/s and a
// grid-stride loop over 1,024 workgroups 5.4 TB/s).  Each with ordinary and with non-temporal
// stores (the step kernels use the latter on lattices this large).
// Prints one JSON line of GB/s, counted as the step kernels count them (2 Q s bytes per node).
// bench.py reports these beside the D3Q27 kernels' rates: they bound what any step kernel in this
// layout can reach on the device it runs on.
#include <cstdio>

#include "lbm_device.hpp"

using namespace lbm;

namespace {

constexpr int N3 = 256;

template <class T, bool NT, int Q = 27>
__global__ void __launch_bounds__(256) k_pull(const T *__restrict__ in, T *__restrict__ out, int64_t stride) {
  const int x = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63), y = (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
  const int z = (int)blockIdx.z;
  const int xm = x == 0 ? N3 - 1 : x - 1, xp = x == N3 - 1 ? 0 : x + 1;
  const int ym = y == 0 ? N3 - 1 : y - 1, yp = y == N3 - 1 ? 0 : y + 1;
  const int zm = z == 0 ? N3 - 1 : z - 1, zp = z == N3 - 1 ? 0 : z + 1;
  T v[Q];
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr int cx = Lat<Q>::cx[i], cy = Lat<Q>::cy[i], cz = Lat<Q>::cz[i];
    const int sx = cx == 1 ? xm : (cx == -1 ? xp : x);
    const int sy = cy == 1 ? ym : (cy == -1 ? yp : y);
    const int sz = cz == 1 ? zm : (cz == -1 ? zp : z);
    v[i] = in[(size_t)i * (size_t)stride + (size_t)sx + (size_t)N3 * ((size_t)sy + (size_t)N3 * (size_t)sz)];
  });
  const size_t n = (size_t)x + (size_t)N3 * ((size_t)y + (size_t)N3 * (size_t)z);
  static_for<Q>([&](auto I) {
    T *dst = out + (size_t)decltype(I)::value * (size_t)stride + n;
    if constexpr (NT) __builtin_nontemporal_store(v[decltype(I)::value], dst);
    else *dst = v[decltype(I)::value];
  });
}

typedef float v4f __attribute__((ext_vector_type(4)));

// one 16-byte element per thread; the grid covers the n4 elements exactly
template <bool NT>
__global__ void __launch_bounds__(256) k_copy(const float4 *__restrict__ in, float4 *__restrict__ out) {
  const v4f *src = (const v4f *)in;
  v4f *dst = (v4f *)out;
  const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
  if constexpr (NT) __builtin_nontemporal_store(src[k], dst + k);
  else dst[k] = src[k];
}

// the 27 population arrays copied node by node (every node reads and writes its own slot of each)
__global__ void __launch_bounds__(256) k_copy27(const float *__restrict__ in, float *__restrict__ out, int64_t stride) {
  const size_t n = (size_t)blockIdx.x * 256 + threadIdx.x;
  float v[27];
  static_for<27>([&](auto I) { v[decltype(I)::value] = in[(size_t)decltype(I)::value * (size_t)stride + n]; });
  static_for<27>([&](auto I) {
    __builtin_nontemporal_store(v[decltype(I)::value], out + (size_t)decltype(I)::value * (size_t)stride + n);
  });
}

bool ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "pull_ceiling: %s: %s\n", what, hipGetErrorString(e));
  return false;
}

// Average seconds per launch over 20 launches after 3 warm-up launches, alternating buffers.
template <class L>
double time_launches(L &&launch) {
  for (int w = 0; w < 3; ++w) launch(w & 1);
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event")) return -1;
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < 20; ++r) launch(r & 1);
  (void)hipEventRecord(e1, 0);
  if (!ok(hipEventSynchronize(e1), "launches")) return -1;
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms * 1e-3 / 20;
}

template <class T, bool NT, int Q = 27>
double pull_gbs() {
  const int64_t N = (int64_t)N3 * N3 * N3, stride = population_stride(N, (int)sizeof(T));
  const size_t bytes = (size_t)stride * Q * sizeof(T);
  T *a = nullptr, *b = nullptr;
  if (!ok(hipMalloc(&a, bytes), "hipMalloc") || !ok(hipMalloc(&b, bytes), "hipMalloc")) return -1;
  if (!ok(hipMemset(a, 0, bytes), "hipMemset") || !ok(hipMemset(b, 0, bytes), "hipMemset")) return -1;
  const double t = time_launches([&](int odd) {
    hipLaunchKernelGGL((k_pull<T, NT, Q>), dim3(N3 / 64, N3 / 4, N3), dim3(256), 0, 0, odd ? b : a, odd ? a : b, stride);
  });
  (void)hipFree(a);
  (void)hipFree(b);
  return t > 0 ? 2.0 * (double)N * Q * sizeof(T) / t / 1e9 : -1;
}

template <bool NT>
double copy_gbs(size_t bytes) {
  float4 *a = nullptr, *b = nullptr;
  if (!ok(hipMalloc(&a, bytes), "hipMalloc") || !ok(hipMalloc(&b, bytes), "hipMalloc")) return -1;
  if (!ok(hipMemset(a, 0, bytes), "hipMemset") || !ok(hipMemset(b, 0, bytes), "hipMemset")) return -1;
  const size_t n4 = bytes / sizeof(float4);  // a multiple of 256 (checked by the caller)
  const double t = time_launches([&](int odd) {
    hipLaunchKernelGGL(k_copy<NT>, dim3((unsigned)(n4 / 256)), dim3(256), 0, 0, odd ? b : a, odd ? a : b);
  });
  (void)hipFree(a);
  (void)hipFree(b);
  return t > 0 ? 2.0 * (double)bytes / t / 1e9 : -1;
}

double copy27_gbs() {
  const int64_t N = (int64_t)N3 * N3 * N3, stride = population_stride(N, 4);
  const size_t bytes = (size_t)stride * 27 * 4;
  float *a = nullptr, *b = nullptr;
  if (!ok(hipMalloc(&a, bytes), "hipMalloc") || !ok(hipMalloc(&b, bytes), "hipMalloc")) return -1;
  if (!ok(hipMemset(a, 0, bytes), "hipMemset") || !ok(hipMemset(b, 0, bytes), "hipMemset")) return -1;
  const double t = time_launches([&](int odd) {
    hipLaunchKernelGGL(k_copy27, dim3((unsigned)(N / 256)), dim3(256), 0, 0, odd ? b : a, odd ? a : b, stride);
  });
  (void)hipFree(a);
  (void)hipFree(b);
  return t > 0 ? 2.0 * (double)N * 27 * 4 / t / 1e9 : -1;
}

}  // namespace

int main() {
  const double p4 = pull_gbs<float, false>(), p2 = pull_gbs<uint16_t, false>();
  const double p4n = pull_gbs<float, true>(), p2n = pull_gbs<uint16_t, true>();
  const double q4n = pull_gbs<float, true, 19>(), q2n = pull_gbs<uint16_t, true, 19>();
  const size_t cb = (size_t)N3 * N3 * N3 * 27 * sizeof(float);  // 4096 * 27 * 16 B * 256 ...: n4 % 256 == 0
  static_assert(((size_t)N3 * N3 * N3 * 27 * sizeof(float) / 16) % 256 == 0, "copy grid covers the buffer");
  const double c = copy_gbs<false>(cb), cn = copy_gbs<true>(cb), c27 = copy27_gbs();
  if (p4 < 0 || p2 < 0 || c < 0 || p4n < 0 || p2n < 0 || cn < 0 || q4n < 0 || q2n < 0 || c27 < 0) return 1;
  std::printf("{\"lattice\": \"256^3, padded SoA (D3Q27; D3Q19 where named)\", \"pull_4byte_gbs\": %.1f, \"pull_2byte_gbs\": %.1f, "
              "\"copy_gbs\": %.1f, \"pull_4byte_nt_gbs\": %.1f, \"pull_2byte_nt_gbs\": %.1f, \"copy_nt_gbs\": %.1f, "
              "\"d3q19_pull_4byte_nt_gbs\": %.1f, \"d3q19_pull_2byte_nt_gbs\": %.1f, \"copy27_nt_gbs\": %.1f, "
              "\"copy_bytes\": %zu}\n",
              p4, p2, c, p4n, p2n, cn, q4n, q2n, c27, cb);
  return 0;
}
