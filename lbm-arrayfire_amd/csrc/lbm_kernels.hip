// lbm_kernels.hip — node-per-thread kernels: the fused step (ordinary stores; the step body is in
// lbm_step.hpp), its batched form, field extraction, initial states and the mass /
// kinetic-energy monitors, and their host-side launchers.
#include <algorithm>

#include "lbm_step.hpp"

namespace lbm {

// rho, u the reference holds after a step: moments of the pulled populations with the overrides
// (REFERENCE), or of the boundary-completed pulled populations (PHYSICAL; 0 at solid nodes).
template <int Q, int ST, int BC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_fields(StepParams p, double *rho_o, double *ux_o,
                                                          double *uy_o, double *uz_o) {
  const NodePos q = node_of_thread(p);
  if (!q.valid) return;
  double r[4];
  fields_of_node<Q, ST, BC>(p, q, r);
  if (rho_o) rho_o[q.n] = r[0];
  if (ux_o) ux_o[q.n] = r[1];
  if (uy_o) uy_o[q.n] = r[2];
  if (uz_o) uz_o[q.n] = r[3];
}

// An initial state restarts the step count of the stochastic rounding (both words of the pair).
template <class P>
__device__ __forceinline__ void reset_step_count(const P &p) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *p.step_out = 0;
    *const_cast<uint32_t *>(p.step_in) = 0;
  }
}

// f = f_eq(rho, u) (cpp/d2q9_bgk_lid.cpp:100-107) into both buffers.
template <int Q, int ST>
__global__ void __launch_bounds__(BLOCK_THREADS) k_init_fields(StepParams p, const double *rho_i,
                                                               const double *ux_i, const double *uy_i,
                                                               const double *uz_i) {
  using C = typename Codec<ST, Q>::C;
  reset_step_count(p);
  const NodePos q = node_of_thread(p);
  if (!q.valid) return;
  const C r = rho_i ? (C)rho_i[q.n] : (C)1, vx = ux_i ? (C)ux_i[q.n] : (C)0,
          vy = uy_i ? (C)uy_i[q.n] : (C)0, vz = (uz_i && Lat<Q>::D == 3) ? (C)uz_i[q.n] : (C)0;
  const C usq = usq_of<Q>(vx, vy, vz);
  C out[Q];
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    out[i] = feq_i<Q, i>(r, vx, vy, vz, usq);
  });
  store_all_f<Q, ST>(p, q.n, out);
  StepParams p2 = p;
  p2.fout = const_cast<void *>(p.fin);
  store_all_f<Q, ST>(p2, q.n, out);
}

// f = f_eq(rho0, u0 + noise), the noise drawn on the device.  SplitMix64 is counter based: draw
// k of the stream seeded with s is mix(s + (k + 1) * 0x9e3779b97f4a7c15), so node n, component a
// (draw k = n * D + a) needs no sequential state.  The double arithmetic is the host contract's.
__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

template <int Q, int ST>
__global__ void __launch_bounds__(BLOCK_THREADS) k_init_equilibrium(StepParams p, double rho0, double u0x,
                                                                    double u0y, double u0z, uint64_t seed,
                                                                    double amp) {
  using C = typename Codec<ST, Q>::C;
  reset_step_count(p);
  const NodePos q = node_of_thread(p);
  if (!q.valid) return;
  constexpr int D = Lat<Q>::D;
  double u[3] = {u0x, u0y, u0z};
  if (!(p.type[q.n] & NODE_SOLID))
    for (int a = 0; a < D; ++a) {
      const double r = (double)(splitmix64_at(seed, (uint64_t)q.n * D + a) >> 11) * 0x1.0p-53;
      u[a] += amp * (2.0 * r - 1.0);
    }
  const C r = (C)rho0, vx = (C)u[0], vy = (C)u[1], vz = D == 3 ? (C)u[2] : (C)0;
  const C usq = usq_of<Q>(vx, vy, vz);
  C out[Q];
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    out[i] = feq_i<Q, i>(r, vx, vy, vz, usq);
  });
  store_all_f<Q, ST>(p, q.n, out);
  StepParams p2 = p;
  p2.fout = const_cast<void *>(p.fin);
  store_all_f<Q, ST>(p2, q.n, out);
}

hipError_t launch_init_equilibrium(const Config &c, const StepParams &p, double rho0, const double *u0,
                                   uint64_t seed, double amp, hipStream_t s) {
  return with_storage(c, [&](auto Qc, auto STc) -> hipError_t {
    hipLaunchKernelGGL((k_init_equilibrium<decltype(Qc)::value, decltype(STc)::value>), dim3(p.ntiles),
                       dim3(BLOCK_THREADS), 0, s, p, rho0, u0[0], u0[1], u0[2], seed, amp);
    return hipGetLastError();
  });
}

// MRT (D3Q27): f = rho0/27 then one collision at rho = rho0, u = 0, no bounce-back
// (rust/examples/d3q27_mrt_lid.rs:228-287).  BGK, and D3Q19 MRT: f = f_eq(rho0, 0).
template <int Q, int ST, int CO>
__global__ void __launch_bounds__(BLOCK_THREADS) k_init_reference(StepParams p, double rho0) {
  using C = typename Codec<ST, Q>::C;
  reset_step_count(p);
  const NodePos q = node_of_thread(p);
  if (!q.valid) return;
  C out[Q];
  if constexpr (CO == CO_BGK || Q == 19) {
    const C usq = usq_of<Q>((C)0, (C)0, (C)0);  // the u-term of f_eq's storage type (FP32: A = 1)
    static_for<Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      out[i] = feq_i<Q, i>((C)rho0, (C)0, (C)0, (C)0, usq);
    });
  } else {
    C f[Q];
    static_for<Q>([&](auto I) { f[decltype(I)::value] = (C)rho0 / (C)27; });
    collide<Q, CO>(p, f, (C)rho0, (C)0, (C)0, (C)0, out, (C)0, (C)1);
  }
  store_all_f<Q, ST>(p, q.n, out);
  StepParams p2 = p;
  p2.fout = const_cast<void *>(p.fin);
  store_all_f<Q, ST>(p2, q.n, out);
}

// Wavefront + workgroup reduction of (mass, kinetic energy, nonfinite) in FP64 over the
// non-solid nodes of one tile; one partial triple per tile (deterministic order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

template <int Q, int ST>
__global__ void __launch_bounds__(BLOCK_THREADS) k_monitors(StepParams p, double *partials) {
  using CD = Codec<ST, Q>;
  using T = typename CD::T;
  __shared__ double red[3][BLOCK_THREADS / 64];
  const NodePos q = node_of_thread(p);
  double mass = 0, ke = 0, bad = 0;
  if (q.valid && !(p.type[q.n] & NODE_SOLID)) {
    const T *fin = (const T *)p.fin;
    double r = 0, jx = 0, jy = 0, jz = 0;
    bool finite = true;
    static_for<Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      double v = (double)CD::template ld<i>(fin + (size_t)i * (size_t)p.stride, q.n) * (1.0 / CD::S);
      if constexpr (CD::shifted) v += Lat<Q>::w[i];  // f = g + w_i, exact in FP64
      finite = finite && __builtin_isfinite(v);
      r += v;
      jx += (double)Lat<Q>::cx[i] * v;
      jy += (double)Lat<Q>::cy[i] * v;
      jz += (double)Lat<Q>::cz[i] * v;
    });
    if (finite) {
      mass = r;
      ke = (jx * jx + jy * jy + jz * jz) / (2.0 * r);
    } else {
      bad = 1;
    }
  }
  mass = wave_sum(mass);
  ke = wave_sum(ke);
  bad = wave_sum(bad);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { red[0][wv] = mass; red[1][wv] = ke; red[2][wv] = bad; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0;
    for (int k = 0; k < BLOCK_THREADS / 64; ++k) { a += red[0][k]; b += red[1][k]; c += red[2][k]; }
    const int tile = xcd_tile((int)blockIdx.x, p.ntiles);
    partials[3 * tile + 0] = a;
    partials[3 * tile + 1] = b;
    partials[3 * tile + 2] = c;
  }
}

// Force on the solid nodes inside box (x0, x1, y0, y1, z0, z1, half-open) by momentum exchange:
// every link from a non-solid node x into such a solid node x + c_i carries c_i (f_i(x) + the
// population it returns), i.e. c_i (2 f_i(x) - 6 w_i rho_bc c_i.u_wall) (second term for moving
// walls), f = the post-collision populations the state holds.  FP64 partial sums per tile.
template <int Q, int ST>
__global__ void __launch_bounds__(BLOCK_THREADS) k_force(StepParams p, int x0, int x1, int y0, int y1, int z0,
                                                         int z1, double *partials) {
  using CD = Codec<ST, Q>;
  using T = typename CD::T;
  __shared__ double red[3][BLOCK_THREADS / 64];
  const NodePos q = node_of_thread(p);
  double fx = 0, fy = 0, fz = 0;
  if (q.valid && !(p.type[q.n] & NODE_SOLID)) {
    const T *fin = (const T *)p.fin;
    const double rw = (double)p.rho_bc;
    static_for<Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      constexpr int cx = Lat<Q>::cx[i], cy = Lat<Q>::cy[i], cz = Lat<Q>::cz[i];
      if constexpr (i != 0) {
        int sx = q.x + cx, sy = q.y + cy, sz = q.z + cz;
        sx = sx < 0 ? sx + p.nx : (sx >= p.nx ? sx - p.nx : sx);
        sy = sy < 0 ? sy + p.ny : (sy >= p.ny ? sy - p.ny : sy);
        sz = sz < 0 ? sz + p.nz : (sz >= p.nz ? sz - p.nz : sz);
        const uint8_t ts = p.type[(uint32_t)sx + (uint32_t)p.nx * ((uint32_t)sy + (uint32_t)p.ny * (uint32_t)sz)];
        if ((ts & NODE_SOLID) && sx >= x0 && sx < x1 && sy >= y0 && sy < y1 && sz >= z0 && sz < z1) {
          double fi = (double)CD::template ld<i>(fin + (size_t)i * (size_t)p.stride, q.n) * (1.0 / CD::S);
          if constexpr (CD::shifted) fi += Lat<Q>::w[i];
          double m = 2.0 * fi;
          if (ts & NODE_MOVING)
            m -= 6.0 * Lat<Q>::w[i] * rw * (cx * p.u_wall[0] + cy * p.u_wall[1] + cz * p.u_wall[2]);
          fx += cx * m;
          fy += cy * m;
          fz += cz * m;
        }
      }
    });
  }
  fx = wave_sum(fx);
  fy = wave_sum(fy);
  fz = wave_sum(fz);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { red[0][wv] = fx; red[1][wv] = fy; red[2][wv] = fz; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0;
    for (int k = 0; k < BLOCK_THREADS / 64; ++k) { a += red[0][k]; b += red[1][k]; c += red[2][k]; }
    const size_t t = blockIdx.x;
    partials[3 * t + 0] = a;
    partials[3 * t + 1] = b;
    partials[3 * t + 2] = c;
  }
}

__global__ void __launch_bounds__(BLOCK_THREADS) k_reduce_partials(const double *partials, int n, double *out) {
  __shared__ double red[3][BLOCK_THREADS / 64];
  double a = 0, b = 0, c = 0;
  for (int k = threadIdx.x; k < n; k += BLOCK_THREADS) {
    a += partials[3 * k + 0];
    b += partials[3 * k + 1];
    c += partials[3 * k + 2];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { red[0][wv] = a; red[1][wv] = b; red[2][wv] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0, y = 0, z = 0;
    for (int k = 0; k < BLOCK_THREADS / 64; ++k) { x += red[0][k]; y += red[1][k]; z += red[2][k]; }
    out[0] = x;
    out[1] = y;
    out[2] = z;
  }
}

template <int Q>
__global__ void __launch_bounds__(BLOCK_THREADS) k_prepare_types(int bc_mode, int nx, int ny, int nz,
                                                                 const uint8_t *in, uint8_t *out, uint32_t *links) {
  const int64_t n = (int64_t)blockIdx.x * BLOCK_THREADS + threadIdx.x;
  if (n >= (int64_t)nx * ny * nz) return;
  uint8_t t = in[n] & 0x1f;
  uint32_t mask = 0;  // wall link mask (StepParams::links)
  if (bc_mode == BC_PHYS && !(t & NODE_SOLID)) {
    const int x = (int)(n % nx), y = (int)((n / nx) % ny), z = (int)(n / ((int64_t)nx * ny));
    static_for<Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      int sx = x - Lat<Q>::cx[i], sy = y - Lat<Q>::cy[i], sz = z - Lat<Q>::cz[i];
      sx = sx < 0 ? sx + nx : (sx >= nx ? sx - nx : sx);
      sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
      sz = sz < 0 ? sz + nz : (sz >= nz ? sz - nz : sz);
      const uint8_t ts = in[sx + (int64_t)nx * (sy + (int64_t)ny * sz)];
      if (i != 0 && (ts & NODE_SOLID)) {
        mask |= 1u << (i - 1);
        if (ts & NODE_MOVING) t |= NODE_WALLMOV;
      }
    });
    if (mask) t |= NODE_SLOW | NODE_WALLNB;
    if (t & (NODE_INLET | NODE_OUTLET)) t |= NODE_SLOW;
  }
  out[n] = t;
  if (links) links[n] = mask;
}

hipError_t launch_prepare_types(int Q, int bc_mode, int nx, int ny, int nz, const uint8_t *in, uint8_t *out,
                                uint32_t *links, hipStream_t s) {
  const int64_t N = (int64_t)nx * ny * nz;
  const dim3 grid((unsigned)((N + BLOCK_THREADS - 1) / BLOCK_THREADS));
  if (Q == 9) hipLaunchKernelGGL(k_prepare_types<9>, grid, dim3(BLOCK_THREADS), 0, s, bc_mode, nx, ny, nz, in, out, links);
  else if (Q == 19) hipLaunchKernelGGL(k_prepare_types<19>, grid, dim3(BLOCK_THREADS), 0, s, bc_mode, nx, ny, nz, in, out, links);
  else hipLaunchKernelGGL(k_prepare_types<27>, grid, dim3(BLOCK_THREADS), 0, s, bc_mode, nx, ny, nz, in, out, links);
  return hipGetLastError();
}

// Gathers rho/u of selected nodes from the four field arrays (sensors).
__global__ void k_gather_fields(const double *fields, int64_t N, const int64_t *nodes, int n, double *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t v = nodes[k];
  for (int a = 0; a < 4; ++a) out[(size_t)a * n + k] = fields[(size_t)a * N + v];
}

hipError_t launch_gather_fields(const double *fields, int64_t N, const int64_t *nodes, int n, double *out,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_gather_fields, dim3((n + 255) / 256), dim3(256), 0, s, fields, N, nodes, n, out);
  return hipGetLastError();
}

hipError_t launch_step_scalar(const Config &c, const StepParams &p, hipStream_t s) {
  if (c.nt) return launch_step_scalar_nt(c, p, s);  // lbm_kernels_nt.hip
  return with_full(c, [&](auto Qc, auto STc, auto BCc, auto COc) -> hipError_t {
    hipLaunchKernelGGL((k_step<decltype(Qc)::value, decltype(STc)::value, decltype(BCc)::value,
                               decltype(COc)::value, false>),
                       dim3(p.ntiles), dim3(BLOCK_THREADS), 0, s, p);
    return hipGetLastError();
  });
}

hipError_t launch_step_batch_scalar(const Config &c, const BatchArgs &b, int grid, hipStream_t s) {
  if (b.nt) return launch_step_batch_scalar_nt(c, b, grid, s);  // lbm_kernels_nt.hip
  return with_full(c, [&](auto Qc, auto STc, auto BCc, auto COc) -> hipError_t {
    constexpr int Q = decltype(Qc)::value, ST = decltype(STc)::value, BC = decltype(BCc)::value,
                  CO = decltype(COc)::value;
    if (b.one_tile)
      hipLaunchKernelGGL((k_step_batch<Q, ST, BC, CO, true>), dim3(grid), dim3(BLOCK_THREADS), 0, s, b);
    else
      hipLaunchKernelGGL((k_step_batch<Q, ST, BC, CO, false>), dim3(grid), dim3(BLOCK_THREADS), 0, s, b);
    return hipGetLastError();
  });
}

hipError_t launch_fields(const Config &c, const StepParams &p, double *rho, double *ux, double *uy,
                         double *uz, hipStream_t s) {
  return with_storage(c, [&](auto Qc, auto STc) -> hipError_t {
    constexpr int Q = decltype(Qc)::value, ST = decltype(STc)::value;
    if (c.bc == BC_REF)
      hipLaunchKernelGGL((k_fields<Q, ST, BC_REF>), dim3(p.ntiles), dim3(BLOCK_THREADS), 0, s, p, rho,
                         ux, uy, uz);
    else
      hipLaunchKernelGGL((k_fields<Q, ST, BC_PHYS>), dim3(p.ntiles), dim3(BLOCK_THREADS), 0, s, p, rho,
                         ux, uy, uz);
    return hipGetLastError();
  });
}

hipError_t launch_init_fields(const Config &c, const StepParams &p, const double *rho, const double *ux,
                              const double *uy, const double *uz, hipStream_t s) {
  return with_storage(c, [&](auto Qc, auto STc) -> hipError_t {
    hipLaunchKernelGGL((k_init_fields<decltype(Qc)::value, decltype(STc)::value>), dim3(p.ntiles),
                       dim3(BLOCK_THREADS), 0, s, p, rho, ux, uy, uz);
    return hipGetLastError();
  });
}

hipError_t launch_init_reference(const Config &c, const StepParams &p, double rho0, hipStream_t s) {
  Config c2 = c;
  c2.bc = BC_REF;
  return with_full(c2, [&](auto Qc, auto STc, auto, auto COc) -> hipError_t {
    hipLaunchKernelGGL((k_init_reference<decltype(Qc)::value, decltype(STc)::value, decltype(COc)::value>),
                       dim3(p.ntiles), dim3(BLOCK_THREADS), 0, s, p, rho0);
    return hipGetLastError();
  });
}

hipError_t launch_force(const Config &c, const StepParams &p, const int *box, double *partials, double *out,
                        hipStream_t s) {
  hipError_t e = with_storage(c, [&](auto Qc, auto STc) -> hipError_t {
    hipLaunchKernelGGL((k_force<decltype(Qc)::value, decltype(STc)::value>), dim3(p.ntiles), dim3(BLOCK_THREADS),
                       0, s, p, box[0], box[1], box[2], box[3], box[4], box[5], partials);
    return hipGetLastError();
  });
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK_THREADS), 0, s, partials, p.ntiles, out);
  return hipGetLastError();
}

__global__ void __launch_bounds__(BLOCK_THREADS) k_check_reciprocal(uint32_t lo, uint32_t hi,
                                                                    unsigned long long *count) {
  unsigned long long bad = 0;
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK_THREADS;
  for (uint64_t b = (uint64_t)lo + blockIdx.x * BLOCK_THREADS + threadIdx.x; b < hi; b += stride) {
    const float x = __uint_as_float((uint32_t)b);
    bad += __float_as_uint(recip_rn(x)) != __float_as_uint(1.0f / x);
  }
  if (bad) atomicAdd(count, bad);
}

hipError_t launch_check_reciprocal(uint32_t lo, uint32_t hi, unsigned long long *count, hipStream_t s) {
  if (hi <= lo) return hipSuccess;
  const uint64_t n = (uint64_t)hi - lo, per_block = (uint64_t)BLOCK_THREADS * 64;
  const unsigned grid = (unsigned)std::min<uint64_t>((n + per_block - 1) / per_block, 65536);
  hipLaunchKernelGGL(k_check_reciprocal, dim3(grid), dim3(BLOCK_THREADS), 0, s, lo, hi, count);
  return hipGetLastError();
}

hipError_t launch_monitors(const Config &c, const StepParams &p, double *partials, double *out,
                           hipStream_t s) {
  hipError_t e = with_storage(c, [&](auto Qc, auto STc) -> hipError_t {
    hipLaunchKernelGGL((k_monitors<decltype(Qc)::value, decltype(STc)::value>), dim3(p.ntiles),
                       dim3(BLOCK_THREADS), 0, s, p, partials);
    return hipGetLastError();
  });
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(BLOCK_THREADS), 0, s, partials, p.ntiles, out);
  return hipGetLastError();
}

}  // namespace lbm

namespace lbm {
hipError_t launch_step(const Config &c, const StepParams &p, hipStream_t s) {
  if (uses_mfma(c)) return launch_step_mrt_mfma(c, p, s);
  return launch_step_scalar(c, p, s);
}

hipError_t launch_step_batch(const Config &c, const BatchArgs &b, int grid, hipStream_t s) {
  if (uses_mfma(c)) return launch_step_batch_mrt_mfma(c, b, grid, s);
  return launch_step_batch_scalar(c, b, grid, s);
}
}  // namespace lbm
