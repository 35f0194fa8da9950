// lbm_shanchen.hip — two-component Shan-Chen D2Q9 (cpp/d2q9_shanchen.cpp:39-181), with the
// reference's work-in-progress defects corrected (streaming before the collision, each fluid's
// interaction from the other fluid's density, interaction sums restarted every step, momenta in
// the common velocity), and its C ABI (lbm_sc_*).
//
// Per step, one launch on the handle's stream (k_sc_step): each workgroup pulls both fluids'
// populations of a 64-column tile (periodic), keeps them in registers and puts both densities of
// the tile and of a one-node ring around it in LDS; then, per node, the moments, the common
// velocity u' = (w1 j1 + w2 j2) / (w1 rho1 + w2 rho2), the interaction
// u_s = u' - (G / omega_s) sum_{i>=1} w_i rho_t(x - c_i) c_i from the neighbours' densities in LDS,
// BGK relaxation of each fluid towards f_eq(rho_s, u_s), store.  HBM traffic per node update:
// 2 x 2 x 9 s, both fluids' populations read and written once (the ring's pulls are the
// neighbouring tiles' populations and mostly hit L2).
//
// Arithmetic follows the single-fluid kernels: FP64 term by term, FP32 with the contracted
// equilibrium, the relaxation fma(omega, f_eq - f, f) and the reciprocal of the common
// velocity's denominator; sums run over ascending i.  The oracle (oracle_sc_run) evaluates the
// same sequence.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "lbm_device.hpp"
#include "lbm_handle.hpp"

namespace lbm {
namespace {

struct ScParams {
  const void *fin;
  void *fout;
  int32_t nx, ny;
  int64_t N, stride;  // populations of fluid 2 start at plane 9
  double omega1, omega2, gw1, gw2;  // gw_s = G / omega_s
};

template <class T>
__device__ __forceinline__ void sc_pull(const ScParams &p, int x, int y, int base, T *f) {
  const T *in = (const T *)p.fin;
  const int xm = x == 0 ? p.nx - 1 : x - 1, xp = x == p.nx - 1 ? 0 : x + 1;
  const int ym = y == 0 ? p.ny - 1 : y - 1, yp = y == p.ny - 1 ? 0 : y + 1;
  static_for<9>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr int cx = Lat<9>::cx[i], cy = Lat<9>::cy[i];
    const int sx = cx == 1 ? xm : (cx == -1 ? xp : x);
    const int sy = cy == 1 ? ym : (cy == -1 ? yp : y);
    f[i] = in[(size_t)(base + i) * (size_t)p.stride + (size_t)sx + (size_t)p.nx * (size_t)sy];
  });
}

template <class T>
__device__ __forceinline__ T sc_sum(const T *f) {
  T r = (T)-0.0;
  static_for<9>([&](auto I) { r = r + f[decltype(I)::value]; });
  return r;
}

// wrapped coordinate of a in [0, n) for a in [-n, 2n) and any n >= 1 (tiles reach 64 nodes past
// the last row or column, which exceeds 2n only for n < 64: the general remainder then)
__device__ __forceinline__ int sc_wrap(int a, int n) {
  if (a < 0) a += n;
  if (a >= n) a = (a < 2 * n) ? a - n : a % n;
  return a;
}

// One step in one launch.  Workgroup = NT threads on a SC_TX x (NT / 64 RPT) tile: thread t owns
// column t & 63 and rows (t >> 6) + (NT / 64) r.  Phase 1 pulls both fluids' populations of the owned nodes into
// registers and writes their densities to LDS; threads t < 2 (SC_TX + 2) + 2 TY also compute the
// densities of one ring node around the tile (pulled and summed, nothing kept).  Phase 2 collides
// the owned nodes with the 8 neighbour densities of each fluid read from LDS.  Tile slots past
// the lattice's last row or column hold the densities of the wrapped coordinate, so the ring of
// a partial tile is what its last valid nodes need.  Populations are read from HBM once per node
// (ring pulls mostly hit L2: they are the neighbour tiles' nodes, on the same XCD by xcd_tile).
constexpr int SC_TX = 64;

template <class T, int RPT, int NT>
__global__ void __launch_bounds__(NT) k_sc_step(ScParams p, int tiles_x, int ntiles) {
  constexpr int ROWS = NT / 64, TY = ROWS * RPT, LX = SC_TX + 2, LY = TY + 2;
  static_assert(NT % 64 == 0 && 2 * LX + 2 * TY <= NT, "one ring node per thread at most");
  __shared__ T lr[2][LY][LX];
  const int tile = xcd_tile((int)blockIdx.x, ntiles);
  const int x0 = (tile % tiles_x) * SC_TX, y0 = (tile / tiles_x) * TY;
  const int t = (int)threadIdx.x, lx = t & 63, ly = t >> 6;
  const int x = x0 + lx, gx = x < p.nx ? x : sc_wrap(x, p.nx);
  T f[RPT][9], g[RPT][9];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int y = y0 + ly + ROWS * r, gy = y < p.ny ? y : sc_wrap(y, p.ny);
    sc_pull<T>(p, gx, gy, 0, f[r]);
    sc_pull<T>(p, gx, gy, 9, g[r]);
    lr[0][ly + ROWS * r + 1][lx + 1] = sc_sum(f[r]);
    lr[1][ly + ROWS * r + 1][lx + 1] = sc_sum(g[r]);
  }
  if (t < 2 * LX + 2 * TY) {
    int hx, hy;
    if (t < LX) { hx = t; hy = 0; }
    else if (t < 2 * LX) { hx = t - LX; hy = LY - 1; }
    else if (t < 2 * LX + TY) { hx = 0; hy = t - 2 * LX + 1; }
    else { hx = LX - 1; hy = t - 2 * LX - TY + 1; }
    const int hgx = sc_wrap(x0 - 1 + hx, p.nx), hgy = sc_wrap(y0 - 1 + hy, p.ny);
    T h[9];
    sc_pull<T>(p, hgx, hgy, 0, h);
    lr[0][hy][hx] = sc_sum(h);
    sc_pull<T>(p, hgx, hgy, 9, h);
    lr[1][hy][hx] = sc_sum(h);
  }
  __syncthreads();
  if (x >= p.nx) return;
  T *out = (T *)p.fout;
  const T w1 = (T)p.omega1, w2 = (T)p.omega2;
  const T om1 = (T)1 - w1, om2 = (T)1 - w2;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    const int y = y0 + ly + ROWS * r;
    if (y >= p.ny) break;
    const int cxl = lx + 1, cyl = ly + ROWS * r + 1;
    const T r1 = lr[0][cyl][cxl], r2 = lr[1][cyl][cxl];
    T j1x = (T)-0.0, j1y = (T)-0.0, j2x = (T)-0.0, j2y = (T)-0.0;
    static_for<9>([&](auto I) {
      constexpr int i = decltype(I)::value;
      j1x = j1x + c_times<Lat<9>::cx[i]>(f[r][i]);
      j1y = j1y + c_times<Lat<9>::cy[i]>(f[r][i]);
      j2x = j2x + c_times<Lat<9>::cx[i]>(g[r][i]);
      j2y = j2y + c_times<Lat<9>::cy[i]>(g[r][i]);
    });
    const T den = w1 * r1 + w2 * r2;
    T upx, upy;
    if constexpr (sizeof(T) == 4) {
      const T inv = recip_rn(den);
      upx = (w1 * j1x + w2 * j2x) * inv;
      upy = (w1 * j1y + w2 * j2y) * inv;
    } else {
      upx = (w1 * j1x + w2 * j2x) / den;
      upy = (w1 * j1y + w2 * j2y) / den;
    }
    // sum_{i>=1} w_i rho_s(x - c_i) c_i of each fluid, ascending i
    T c1x = (T)-0.0, c1y = (T)-0.0, c2x = (T)-0.0, c2y = (T)-0.0;
    static_for<8>([&](auto I) {
      constexpr int i = decltype(I)::value + 1;
      constexpr int cx = Lat<9>::cx[i], cy = Lat<9>::cy[i];
      const T a = (T)Lat<9>::w[i] * lr[0][cyl - cy][cxl - cx];
      const T b = (T)Lat<9>::w[i] * lr[1][cyl - cy][cxl - cx];
      if constexpr (cx == 1) { c1x = c1x + a; c2x = c2x + b; }
      if constexpr (cx == -1) { c1x = c1x - a; c2x = c2x - b; }
      if constexpr (cy == 1) { c1y = c1y + a; c2y = c2y + b; }
      if constexpr (cy == -1) { c1y = c1y - a; c2y = c2y - b; }
    });
    const T u1x = upx - (T)p.gw1 * c2x, u1y = upy - (T)p.gw1 * c2y;
    const T u2x = upx - (T)p.gw2 * c1x, u2y = upy - (T)p.gw2 * c1y;
    const T a1 = usq_of<9>(u1x, u1y, (T)0), a2 = usq_of<9>(u2x, u2y, (T)0);
    const size_t n = (size_t)x + (size_t)p.nx * (size_t)y;
    static_for<9>([&](auto I) {
      constexpr int i = decltype(I)::value;
      const T e1 = feq_i<9, i>(r1, u1x, u1y, (T)0, a1), e2 = feq_i<9, i>(r2, u2x, u2y, (T)0, a2);
      T o1, o2;
      if constexpr (sizeof(T) == 4) {
        o1 = __builtin_fmaf(w1, e1 - f[r][i], f[r][i]);
        o2 = __builtin_fmaf(w2, e2 - g[r][i], g[r][i]);
      } else {
        o1 = w1 * e1 + om1 * f[r][i];
        o2 = w2 * e2 + om2 * g[r][i];
      }
      out[(size_t)i * (size_t)p.stride + n] = o1;
      out[(size_t)(9 + i) * (size_t)p.stride + n] = o2;
    });
  }
}

// Tile shapes {rows per thread, threads}; LBM_SC_SHAPE=<index> picks one at create.  Default 0,
// the fastest measured at 1024 x 1024: FP32 shape 3 (2 x 512: 64 x 16 tiles, 81 VGPRs, 5 waves per
// SIMD), FP64 shape 8 (2 x 1024: 64 x 32 tiles, 123 VGPRs, 4 waves per SIMD).
struct ScShape { int rpt, nt; };
constexpr ScShape kScShapes[] = {{0, 0}, {2, 256}, {4, 256}, {2, 512}, {4, 512}, {8, 256}, {1, 512}, {1, 1024}, {2, 1024}};
constexpr int kScNumShapes = (int)(sizeof(kScShapes) / sizeof(kScShapes[0]));

template <class T, int RPT, int NT>
hipError_t launch_sc_step(const ScParams &p, int tiles_x, int ty, int ny, hipStream_t s) {
  const int nt = tiles_x * ((ny + ty - 1) / ty);
  hipLaunchKernelGGL((k_sc_step<T, RPT, NT>), dim3((unsigned)nt), dim3(NT), 0, s, p, tiles_x, nt);
  return hipGetLastError();
}

template <class T>
hipError_t launch_sc_shape(int shape, const ScParams &p, int tiles_x, int ny, hipStream_t s) {
  if (shape == 0) shape = sizeof(T) == 4 ? 3 : 8;
  switch (shape) {
    case 1: return launch_sc_step<T, 2, 256>(p, tiles_x, 8, ny, s);
    case 2: return launch_sc_step<T, 4, 256>(p, tiles_x, 16, ny, s);
    case 3: return launch_sc_step<T, 2, 512>(p, tiles_x, 16, ny, s);
    case 4: return launch_sc_step<T, 4, 512>(p, tiles_x, 32, ny, s);
    case 5: return launch_sc_step<T, 8, 256>(p, tiles_x, 32, ny, s);
    case 6: return launch_sc_step<T, 1, 512>(p, tiles_x, 8, ny, s);
    case 7: return launch_sc_step<T, 1, 1024>(p, tiles_x, 16, ny, s);
    default: return launch_sc_step<T, 2, 1024>(p, tiles_x, 32, ny, s);
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_sc_fields(ScParams p, double *rho1, double *rho2, double *ux, double *uy) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= p.N) return;
  const T *in = (const T *)p.fin;
  T r1 = (T)-0.0, r2 = (T)-0.0, jx = (T)-0.0, jy = (T)-0.0;
  static_for<9>([&](auto I) {
    constexpr int i = decltype(I)::value;
    const T a = in[(size_t)i * (size_t)p.stride + (size_t)n], b = in[(size_t)(9 + i) * (size_t)p.stride + (size_t)n];
    r1 = r1 + a;
    r2 = r2 + b;
    jx = jx + c_times<Lat<9>::cx[i]>(a) + c_times<Lat<9>::cx[i]>(b);
    jy = jy + c_times<Lat<9>::cy[i]>(a) + c_times<Lat<9>::cy[i]>(b);
  });
  rho1[n] = (double)r1;
  rho2[n] = (double)r2;
  ux[n] = (double)jx / ((double)r1 + (double)r2);
  uy[n] = (double)jy / ((double)r1 + (double)r2);
}

template <class T>
__global__ void __launch_bounds__(256) k_sc_init(ScParams p, const double *rho1, const double *rho2) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= p.N) return;
  T *a = (T *)p.fin, *b = (T *)p.fout;
  static_for<9>([&](auto I) {
    constexpr int i = decltype(I)::value;
    const T w = (T)Lat<9>::w[i];
    const T f = w * (T)rho1[n], g = w * (T)rho2[n];
    a[(size_t)i * (size_t)p.stride + (size_t)n] = f;
    a[(size_t)(9 + i) * (size_t)p.stride + (size_t)n] = g;
    b[(size_t)i * (size_t)p.stride + (size_t)n] = f;
    b[(size_t)(9 + i) * (size_t)p.stride + (size_t)n] = g;
  });
}

}  // namespace
}  // namespace lbm

struct lbm_sc_ctx {
  lbm_sc_desc d{};
  int device = 0;
  int64_t N = 0, stride = 0;
  int esize = 8;
  void *f[2] = {nullptr, nullptr};
  double *fields = nullptr;  // 4 N, lbm_sc_get_fields staging
  int cur = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  lbm::ScParams params() const {
    lbm::ScParams p{};
    p.fin = f[cur];
    p.fout = f[cur ^ 1];
    p.nx = d.nx;
    p.ny = d.ny;
    p.N = N;
    p.stride = stride;
    p.omega1 = d.omega1;
    p.omega2 = d.omega2;
    p.gw1 = d.g / d.omega1;
    p.gw2 = d.g / d.omega2;
    return p;
  }
  size_t buf_bytes() const { return (size_t)stride * 18 * (size_t)esize; }
  unsigned grid() const { return (unsigned)((N + 255) / 256); }
  int shape = 0;  // index into kScShapes
  int tiles_x() const { return (d.nx + lbm::SC_TX - 1) / lbm::SC_TX; }
};

namespace {

lbm_status sc_fail(lbm_status s, const std::string &m) { return lbm::set_error(s, m); }

#define SCCHK(call)                                                                               \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return sc_fail(LBM_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class F>
hipError_t sc_typed(const lbm_sc_ctx *h, F &&f) {
  return h->d.storage == LBM_F64 ? f((double)0) : f((float)0);
}

lbm_status sc_steps(lbm_sc_handle h, int64_t nsteps) {
  for (int64_t t = 0; t < nsteps; ++t) {
    const lbm::ScParams p = h->params();
    SCCHK(sc_typed(h, [&](auto z) {
      using T = decltype(z);
      return lbm::launch_sc_shape<T>(h->shape, p, h->tiles_x(), h->d.ny, h->stream);
    }));
    h->cur ^= 1;
  }
  return LBM_OK;
}

}  // namespace

extern "C" {

lbm_status lbm_sc_create(const lbm_sc_desc *d, lbm_sc_handle *out) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!out) return sc_fail(LBM_EINVAL, "out is NULL");
    *out = nullptr;
    if (!d) return sc_fail(LBM_EINVAL, "descriptor is NULL");
    if (d->nx < 1 || d->ny < 1) return sc_fail(LBM_EINVAL, "lattice dimensions must be >= 1");
    if ((uint64_t)d->nx * (uint64_t)d->ny >= (1ull << 31)) return sc_fail(LBM_EINVAL, "lattice too large");
    if (d->storage != LBM_F64 && d->storage != LBM_F32) return sc_fail(LBM_EINVAL, "Shan-Chen storage must be F64 or F32");
    if (!(d->omega1 > 0 && d->omega1 < 2 && d->omega2 > 0 && d->omega2 < 2))
      return sc_fail(LBM_EINVAL, "relaxation rates must lie in (0, 2)");
    if (!std::isfinite(d->g)) return sc_fail(LBM_EINVAL, "G is not finite");
    lbm_sc_ctx *h = new (std::nothrow) lbm_sc_ctx();
    if (!h) return sc_fail(LBM_ENOMEM, "host allocation failed");
    h->d = *d;
    h->N = (int64_t)d->nx * d->ny;
    h->esize = d->storage == LBM_F64 ? 8 : 4;
    h->stride = lbm::population_stride(h->N, h->esize);
    if (const char *e = std::getenv("LBM_SC_SHAPE")) {
      const int v = std::atoi(e);
      if (v < 0 || v >= lbm::kScNumShapes) {
        delete h;
        return sc_fail(LBM_EINVAL, "LBM_SC_SHAPE out of range");
      }
      h->shape = v;
    }
    auto cleanup = [&](lbm_status s, const std::string &m) {
      lbm_sc_destroy(h);
      return sc_fail(s, m);
    };
    if (hipGetDevice(&h->device) != hipSuccess) return cleanup(LBM_EDEVICE, "hipGetDevice failed");
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
      return cleanup(LBM_EDEVICE, "hipStreamCreate failed");
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
      return cleanup(LBM_EDEVICE, "hipEventCreate failed");
    for (int b = 0; b < 2; ++b)
      if (hipMalloc(&h->f[b], h->buf_bytes()) != hipSuccess) return cleanup(LBM_ENOMEM, "population buffer allocation failed");
    if (hipMalloc((void **)&h->fields, 4 * (size_t)h->N * sizeof(double)) != hipSuccess)
      return cleanup(LBM_ENOMEM, "field buffer allocation failed");
    *out = h;
    return LBM_OK;
  });
}

lbm_status lbm_sc_init(lbm_sc_handle h, const double *rho1, const double *rho2) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || !rho1 || !rho2) return sc_fail(LBM_EINVAL, "handle or density is NULL");
    SCCHK(hipSetDevice(h->device));
    const size_t nb = (size_t)h->N * sizeof(double);
    SCCHK(hipMemcpyAsync(h->fields, rho1, nb, hipMemcpyHostToDevice, h->stream));
    SCCHK(hipMemcpyAsync(h->fields + h->N, rho2, nb, hipMemcpyHostToDevice, h->stream));
    h->cur = 0;
    const lbm::ScParams p = h->params();
    SCCHK(sc_typed(h, [&](auto z) {
      using T = decltype(z);
      hipLaunchKernelGGL(lbm::k_sc_init<T>, dim3(h->grid()), dim3(256), 0, h->stream, p, h->fields, h->fields + h->N);
      return hipGetLastError();
    }));
    SCCHK(hipStreamSynchronize(h->stream));
    return LBM_OK;
  });
}

lbm_status lbm_sc_step(lbm_sc_handle h, int64_t nsteps) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return sc_fail(LBM_EINVAL, "handle is NULL");
    if (nsteps < 0) return sc_fail(LBM_EINVAL, "nsteps < 0");
    SCCHK(hipSetDevice(h->device));
    return sc_steps(h, nsteps);
  });
}

lbm_status lbm_sc_step_timed(lbm_sc_handle h, int64_t nsteps, double *ms) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || !ms) return sc_fail(LBM_EINVAL, "handle or ms is NULL");
    SCCHK(hipSetDevice(h->device));
    SCCHK(hipEventRecord(h->ev0, h->stream));
    const lbm_status st = sc_steps(h, nsteps);
    if (st != LBM_OK) return st;
    SCCHK(hipEventRecord(h->ev1, h->stream));
    SCCHK(hipEventSynchronize(h->ev1));
    float t = 0;
    SCCHK(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = t;
    return LBM_OK;
  });
}

lbm_status lbm_sc_get_fields(lbm_sc_handle h, double *rho1, double *rho2, double *ux, double *uy) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return sc_fail(LBM_EINVAL, "handle is NULL");
    SCCHK(hipSetDevice(h->device));
    lbm::ScParams p = h->params();
    double *F = h->fields;
    SCCHK(sc_typed(h, [&](auto z) {
      using T = decltype(z);
      hipLaunchKernelGGL(lbm::k_sc_fields<T>, dim3(h->grid()), dim3(256), 0, h->stream, p, F, F + h->N, F + 2 * h->N,
                         F + 3 * h->N);
      return hipGetLastError();
    }));
    double *dst[4] = {rho1, rho2, ux, uy};
    for (int k = 0; k < 4; ++k)
      if (dst[k]) SCCHK(hipMemcpyAsync(dst[k], F + (size_t)k * h->N, (size_t)h->N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SCCHK(hipStreamSynchronize(h->stream));
    return LBM_OK;
  });
}

lbm_status lbm_sc_get_populations(lbm_sc_handle h, double *f) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || !f) return sc_fail(LBM_EINVAL, "handle or buffer is NULL");
    SCCHK(hipSetDevice(h->device));
    SCCHK(hipStreamSynchronize(h->stream));
    std::vector<char> tmp((size_t)h->stride * 18 * (size_t)h->esize);
    SCCHK(hipMemcpy(tmp.data(), h->f[h->cur], tmp.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < 18; ++i)
      for (int64_t n = 0; n < h->N; ++n) {
        const size_t src = (size_t)i * (size_t)h->stride + (size_t)n;
        f[(size_t)i * (size_t)h->N + (size_t)n] =
            h->esize == 8 ? ((const double *)tmp.data())[src] : (double)((const float *)tmp.data())[src];
      }
    return LBM_OK;
  });
}

lbm_status lbm_sc_set_populations(lbm_sc_handle h, const double *f) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h || !f) return sc_fail(LBM_EINVAL, "handle or buffer is NULL");
    SCCHK(hipSetDevice(h->device));
    SCCHK(hipStreamSynchronize(h->stream));
    std::vector<char> tmp((size_t)h->stride * 18 * (size_t)h->esize, 0);
    for (int i = 0; i < 18; ++i)
      for (int64_t n = 0; n < h->N; ++n) {
        const size_t dst = (size_t)i * (size_t)h->stride + (size_t)n;
        const double v = f[(size_t)i * (size_t)h->N + (size_t)n];
        if (h->esize == 8) ((double *)tmp.data())[dst] = v;
        else ((float *)tmp.data())[dst] = (float)v;
      }
    SCCHK(hipMemcpy(h->f[h->cur], tmp.data(), tmp.size(), hipMemcpyHostToDevice));
    return LBM_OK;
  });
}

lbm_status lbm_sc_destroy(lbm_sc_handle h) {
  return lbm::abi_guard([&]() -> lbm_status {
    if (!h) return LBM_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (int b = 0; b < 2; ++b)
      if (h->f[b]) (void)hipFree(h->f[b]);
    if (h->fields) (void)hipFree(h->fields);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return LBM_OK;
  });
}

}  // extern "C"
