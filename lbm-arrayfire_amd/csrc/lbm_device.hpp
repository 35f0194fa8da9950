// lbm_device.hpp — device building blocks of the fused pull stream + collide step (gfx950).
//
// One thread owns one node.  It pulls its Q populations from the neighbours (periodic wrap as
// the reference's af::shift, cpp/d2q9_bgk_lid.cpp:18-28), applies the boundary rules, takes the
// moments, collides and writes Q populations: 2*Q*s bytes of HBM traffic per node update.
//
// FP64 arithmetic follows the reference's operation order (cpp/d2q9_bgk_lid.cpp:131-153,
// rust/examples/d3q27_mrt_lid.rs:311-387) term by term, with every product and sum rounded
// separately (-ffp-contract=off).  FP32 arithmetic uses explicit FMAs in the equilibrium, the
// BGK relaxation and the MRT moment transforms (the contraction the reference's FP32 GPU builds
// get from their JIT compilers).  The CPU oracle follows the same sequences, so both precisions
// reproduce it bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "lbm_internal.hpp"
#include "mrt_tables.h"

namespace lbm {

template <int V> using IC = std::integral_constant<int, V>;

template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(IC<I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// ---- storage codecs ----------------------------------------------------------------------
template <int ST, int Q> struct Codec;
// Each codec converts one population between storage (T) and compute (C) types:
//   ld<I>(p, k)      the compute value of population I stored at p[k] (dec<I>(raw<I>(p, k)));
//   raw<I>(p, k)     the stored form at p[k];  dec<I>(t)  its compute value;
//   enc<I>(v, r)     the stored form of compute value v (r: random bits, shifted codecs only);
//   enc_f<I>(f)      the stored form of a population f given unshifted, rounded to nearest (the
//                    initial states);
//   st<I>(p, k, v, r)  p[k] = enc<I>(v, r).
// shifted: the compute value is the deviation g = f - w_i, not f (the 2-byte storages).
// S: the compute scale.  Shifted codecs hand out S g (S an exact power of two), and every step
// computes in those units: each operation on S-scaled operands rounds to S times its unscaled
// result, so a step in scaled units equals the unscaled step bit for bit, while loads and stores
// need no scaling multiply.  Quantities that do not scale (rho, u) stay unscaled; rho S and the
// density deviation S drho enter the equilibria.
template <int Q> struct Codec<ST_F64, Q> {
  using T = double;
  using C = double;
  static constexpr bool shifted = false;
  static constexpr float S = 1.0f;
  template <int I> __device__ static __forceinline__ C ld(const T *p, uint32_t k) { return p[k]; }
  template <int I> __device__ static __forceinline__ T raw(const T *p, uint32_t k) { return p[k]; }
  template <int I> __device__ static __forceinline__ C dec(T v) { return v; }
  template <int I> __device__ static __forceinline__ T enc(C v, uint32_t) { return v; }
  template <int I> __device__ static __forceinline__ T enc_f(C v) { return v; }
  template <int I> __device__ static __forceinline__ void st(T *p, uint32_t k, C v, uint32_t) { p[k] = v; }
};
template <int Q> struct Codec<ST_F32, Q> {
  using T = float;
  using C = float;
  static constexpr bool shifted = false;
  static constexpr float S = 1.0f;
  template <int I> __device__ static __forceinline__ C ld(const T *p, uint32_t k) { return p[k]; }
  template <int I> __device__ static __forceinline__ T raw(const T *p, uint32_t k) { return p[k]; }
  template <int I> __device__ static __forceinline__ C dec(T v) { return v; }
  template <int I> __device__ static __forceinline__ T enc(C v, uint32_t) { return v; }
  template <int I> __device__ static __forceinline__ T enc_f(C v) { return v; }
  template <int I> __device__ static __forceinline__ void st(T *p, uint32_t k, C v, uint32_t) { p[k] = v; }
};
// Shifted half storage: the stored value is (f - w_i) * 2^15 in IEEE half.  The exact
// power-of-two scale keeps the small deviations of the low-weight D3Q27 directions out of the
// half-precision subnormal range, so every stored value carries 11 significant bits.
// The compute value is the stored half itself, 2^15 g (S = 2^15), stored with
// stochastic rounding (v_cvt_sr_f16_f32: round up with probability equal to the fraction of the
// half ulp that g exceeds the value below, from the random bits r).  A population whose state
// does not change from one step to the next would otherwise round the same way every step, and in
// a steady region those frozen errors act as a source: on a 256^2 lid cavity at tau = 1 the
// velocity drifted to 1.2e-2 of the FP64 run after 8,000 steps with rounding to nearest, against
// 3.8e-4 (and not growing) with stochastic rounding.  Initial states are rounded to nearest.
// The 2-byte codecs load through a 32-bit byte offset from the population's (uniform) base, so
// the loads take the scalar-base address form and no 64-bit address arithmetic per population
// (-53 VALU instructions per node in the D3Q27 MRT F16S step, +2 to +4 % on the F16S workloads);
// lbm_create limits 2-byte lattices to 2^31 - 1 nodes so the offsets are exact.
constexpr float F16S_SCALE = 32768.0f, F16S_INV_SCALE = 1.0f / 32768.0f;
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf162_t __attribute__((ext_vector_type(2)));
template <int Q> struct Codec<ST_F16S, Q> {
  using T = uint16_t;
  using C = float;
  static constexpr bool shifted = true;
  static constexpr float S = F16S_SCALE;
  template <int I> __device__ static __forceinline__ T raw(const T *p, uint32_t k) {
    return *(const T *)((const char *)p + (uint32_t)(k * (uint32_t)sizeof(T)));
  }
  template <int I> __device__ static __forceinline__ C dec(T v) { return (float)__builtin_bit_cast(_Float16, v); }
  template <int I> __device__ static __forceinline__ C ld(const T *p, uint32_t k) { return dec<I>(raw<I>(p, k)); }
  // the conversion writes the low half of its destination and keeps the high half, which no
  // store reads: the destination starts as the value's own register (no zeroing move)
  template <int I> __device__ static __forceinline__ T enc(C v, uint32_t r) {
    const half2_t h = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(half2_t, v), v, r, false);
    return __builtin_bit_cast(uint16_t, h[0]);
  }
  // (f - w_i) * 2^15 as one FMA: the scale is a power of two, so this rounds like the separate
  // difference and product
  template <int I> __device__ static __forceinline__ T enc_f(C v) {
    constexpr float ws = -(float)Lat<Q>::w[I] * F16S_SCALE;
    return __builtin_bit_cast(uint16_t, (_Float16)__builtin_fmaf(v, F16S_SCALE, ws));
  }
  template <int I> __device__ static __forceinline__ void st(T *p, uint32_t k, C v, uint32_t r) { p[k] = enc<I>(v, r); }
};
// Shifted bfloat16 storage: the stored value is f - w_i (stochastically rounded by the steps,
// v_cvt_sr_bf16_f32; to nearest in the initial states).
template <int Q> struct Codec<ST_BF16S, Q> {
  using T = uint16_t;
  using C = float;
  static constexpr bool shifted = true;
  static constexpr float S = 1.0f;
  template <int I> __device__ static __forceinline__ T raw(const T *p, uint32_t k) {
    return *(const T *)((const char *)p + (uint32_t)(k * (uint32_t)sizeof(T)));
  }
  template <int I> __device__ static __forceinline__ C dec(T v) { return __builtin_bit_cast(float, (uint32_t)v << 16); }
  template <int I> __device__ static __forceinline__ C ld(const T *p, uint32_t k) { return dec<I>(raw<I>(p, k)); }
  template <int I> __device__ static __forceinline__ T enc(C v, uint32_t r) {
    const bf162_t h = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(bf162_t, v), v, r, false);
    return __builtin_bit_cast(uint16_t, h[0]);
  }
  template <int I> __device__ static __forceinline__ T enc_f(C v) {
    return __builtin_bit_cast(uint16_t, (__bf16)(v - (float)Lat<Q>::w[I]));
  }
  template <int I> __device__ static __forceinline__ void st(T *p, uint32_t k, C v, uint32_t r) { p[k] = enc<I>(v, r); }
};

// ---- random bits of the stochastic rounding ------------------------------------------------
// Node n's rounding bits in step t (t counts the lattice's steps from its initial state):
// mix(n ^ key(t)), key(t) = mix(t + 0x9E3779B9), mix the "lowbias32" integer hash (C. Wellons);
// population i takes them rotated left by 7 i (mod 32).  The conversions use the top bits (13 for
// half, 16 for bfloat16).  A function of (n, t, i) alone, so a node's result does not depend on
// how the lattice is cut into tiles, launches or batches.
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ uint32_t step_key(uint32_t t) { return mix32(t + 0x9E3779B9u); }
__device__ __forceinline__ uint32_t node_bits(uint32_t n, uint32_t key) { return mix32(n ^ key); }
template <int I>
__device__ __forceinline__ uint32_t pop_bits(uint32_t h) {
  constexpr int k = (7 * I) & 31;
  if constexpr (k == 0) return h;
  else return (h << k) | (h >> (32 - k));
}

// The step counter of a shifted-storage lattice: read from p.step_in by every workgroup; the
// thread that owns node 0 of the lattice's first tile writes the next value to p.step_out (the
// other slot of the lattice's pair, the one the next launch reads).  Returns key(t).
template <bool SH, class P>
__device__ __forceinline__ uint32_t step_begin(const P &p, int tile, uint32_t steps) {
  if constexpr (!SH) {
    return 0;
  } else {
    const uint32_t t = *p.step_in;
    if (tile == 0 && threadIdx.x == 0) *p.step_out = t + steps;
    return step_key(t);
  }
}

// ---- XCD-aware tile order ---------------------------------------------------------------
// Workgroups are dealt round-robin to the 8 XCDs (workgroup b runs on XCD b % 8).  Remapping
// b -> tile so that each XCD owns one contiguous range of tiles keeps a slab of the lattice, and
// the neighbour rows its pulls touch, in that XCD's L2 from one step to the next.
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int q = ntiles >> 3, r = ntiles & 7, k = b & 7, j = b >> 3;
  return k * q + (k < r ? k : r) + j;
}

// Tile order of lattices that stream from HBM every step (too large for the caches): close to
// natural order, row after row, so the eight XCDs work on neighbouring tiles at any moment (the XCD
// slabs above, which keep a cache-resident lattice's neighbourhood in one XCD's L2, cost them 6 %: a
// bare 256^3 D3Q27 pull at 5.33 against 5.67 TB/s).  Within each row of tiles the x position is
// rotated by the row index, so that the tiles holding the x faces (inlet, outlet, walls) do not all
// land on one XCD (rows of 8 tiles would send every x = 0 tile to XCD 0).
__device__ __forceinline__ int stream_tile(int b, int gx) {
  const int r = b / gx, c = b - r * gx;
  return r * gx + (c + r) % gx;
}

struct NodePos {
  int x, y, z;
  uint32_t n;
  bool valid;
};

template <class P>
__device__ __forceinline__ NodePos node_of_tile(const P &p, int tile) {
  NodePos q;
  if (p.flat) {  // 256 consecutive node indices per tile (rows not a multiple of 64 wide)
    const uint32_t n = (uint32_t)tile * (uint32_t)BLOCK_THREADS + threadIdx.x;
    const uint32_t yz = fastdiv(n, p.div_nx[0], p.div_nx[1], p.div_nx[2]);
    const uint32_t z = fastdiv(yz, p.div_ny[0], p.div_ny[1], p.div_ny[2]);
    q.x = (int)(n - yz * (uint32_t)p.nx);
    q.y = (int)(yz - z * (uint32_t)p.ny);
    q.z = (int)z;
    q.n = n;
    q.valid = (int64_t)n < p.N;
    return q;
  }
  const int bx = tile % p.gx, t2 = tile / p.gx, by = t2 % p.gy, bz = t2 / p.gy;
  q.x = bx * TILE_X + (int)(threadIdx.x & (TILE_X - 1));
  q.y = by * TILE_Y + (int)(threadIdx.x / TILE_X);
  q.z = bz;
  q.valid = q.x < p.nx && q.y < p.ny;
  q.n = (uint32_t)q.x + (uint32_t)p.nx * ((uint32_t)q.y + (uint32_t)p.ny * (uint32_t)q.z);
  return q;
}

template <class P>
__device__ __forceinline__ NodePos node_of_thread(const P &p) {
  return node_of_tile(p, xcd_tile((int)blockIdx.x, p.ntiles));
}

// Batched launches: a persistent workgroup walks one contiguous chunk of the batch's tile range.
// The table is read through the constant address space: it does not change while a launch runs
// (the host rewrites it between launches, in stream order), so every read of it is a uniform
// scalar load.  The step functions take the member's parameter block by constant-space reference
// (they are templates on the parameter type), so, as with a kernel argument, the compiler loads
// each field where it is used and may load it again instead of holding it.  Through generic
// pointers the parameters landed in VGPRs (D3Q27 MRT: 165 VGPRs against 81 for the
// single-lattice kernel); copied into a local they spilled SGPRs into VGPR lanes (101 against 69).
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using const_as = const __attribute__((address_space(4))) T *;
#else  // the host pass only parses device code; it has no constant address space
template <class T> using const_as = const T *;
#endif
// MULTI walks the multi-step tiles (prefix_multi) instead of the single-step ones.  ONE: the grid
// has one workgroup per tile (no loop, so no state carried across tiles).
template <bool MULTI = false, bool ONE = false, class F>
__device__ __forceinline__ void for_batch_tiles(const BatchArgs &b, F &&f) {
  const_as<BatchTable> tab = (const_as<BatchTable>)b.tab;
  const int nmembers = tab->nmembers, total = MULTI ? tab->total_multi : tab->total_tiles;
  const_as<int32_t> prefix = (const_as<int32_t>)(MULTI ? tab->prefix_multi : tab->prefix);
  const_as<StepParams> members = (const_as<StepParams>)tab->members[b.phase];
  const int nb = (int)gridDim.x;
  // one workgroup per tile: the grid may hold more workgroups than tiles (it is sized with some
  // room, so membership changes need no new launch shape); the spare ones leave at once, and the
  // tiles are dealt over the tile count so every XCD gets its share.  Groups too large for the
  // caches take their tiles in stream order (stream_tile within each member), the others in XCD
  // slabs.
  const bool stream = !MULTI && tab->stream;
  if (ONE && (int)blockIdx.x >= total) return;
  const int blk = ONE ? (stream ? (int)blockIdx.x : xcd_tile((int)blockIdx.x, total)) : xcd_tile((int)blockIdx.x, nb);
  const int chunk = ONE ? 1 : (total + nb - 1) / nb;
  const int t0 = blk * chunk;
  const int t1 = t0 + chunk < total ? t0 + chunk : total;
  if (t0 >= t1) return;
  int lo = 0, hi = nmembers - 1;  // member holding t0
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= t0) lo = mid;
    else hi = mid - 1;
  }
  if constexpr (ONE) {
    const int local = t0 - prefix[lo];
    f(members[lo], stream ? stream_tile(local, members[lo].gx) : local);
    return;
  }
  int m = lo;
  for (int t = t0; t < t1; ++t) {
    while (t >= prefix[m + 1]) ++m;
    f(members[m], t - prefix[m]);  // a constant-space reference: fields are re-loaded as needed
  }
}

// ---- pulls -------------------------------------------------------------------------------
// The periodic source x - c_i (mod dims) of each population of node q: src(IC<i>) is its node
// index.  Source rows are the node's row plus wrapped y and z steps (modulo 2^32, exact once
// summed): additions instead of one 32-bit multiply (a quarter-rate instruction) per (y, z) pair.
template <int Q>
struct PullSources {
  uint32_t row, dym, dyp, dzm, dzp;
  int xm, x, xp;
  template <class P>
  __device__ __forceinline__ PullSources(const P &p, const NodePos &q) {
    const uint32_t nx = (uint32_t)p.nx, nxy = nx * (uint32_t)p.ny;
    row = q.n - (uint32_t)q.x;
    dym = q.y == 0 ? nxy - nx : 0u - nx;
    dyp = q.y == p.ny - 1 ? nx - nxy : nx;
    dzm = q.z == 0 ? nxy * (uint32_t)(p.nz - 1) : 0u - nxy;
    dzp = q.z == p.nz - 1 ? 0u - nxy * (uint32_t)(p.nz - 1) : nxy;
    xm = q.x == 0 ? p.nx - 1 : q.x - 1;
    x = q.x;
    xp = q.x == p.nx - 1 ? 0 : q.x + 1;
  }
  template <int i>
  __device__ __forceinline__ uint32_t operator()(IC<i>) const {
    constexpr int cx = Lat<Q>::cx[i], cy = Lat<Q>::cy[i], cz = Lat<Q>::cz[i];
    const int sx = cx == 1 ? xm : (cx == -1 ? xp : x);
    uint32_t r = row;
    if constexpr (cz == 1) r += dzm;
    if constexpr (cz == -1) r += dzp;
    if constexpr (cy == 1) r += dym;
    if constexpr (cy == -1) r += dyp;
    return r + (uint32_t)sx;
  }
};

// Periodic pull: f_i(x) = F_i(x - c_i mod dims).
template <int Q, int ST, class P>
__device__ __forceinline__ void pull_periodic(const P &p, const NodePos &q,
                                              typename Codec<ST, Q>::C *f) {
  using CD = Codec<ST, Q>;
  using T = typename CD::T;
  const T *fin = (const T *)p.fin;
  const PullSources<Q> src(p, q);
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    f[i] = CD::template ld<i>(fin + (size_t)i * (size_t)p.stride, src(I));
  });
}

template <class C, class P>
__device__ __forceinline__ C inlet_velocity(const P &p, int D, int y, int z) {
  if (p.inlet_profile == 0) return (C)p.u_in;
  const C H = (C)(p.ny - 2), eta = (C)y - (C)0.5;
  C u = (((C)4 * (C)p.u_in) * eta * (H - eta)) / (H * H);
  if (D == 3) {
    const C Hz = (C)(p.nz - 2), ez = (C)z - (C)0.5;
    u = u * (((C)4 * ez * (Hz - ez)) / (Hz * Hz));
  }
  return u;
}

// c u for a lattice component c in {-1, 0, 1}: +-u, and for c = 0 an exact -0 (which the
// compiler folds out of the sums below; 0 u would cost a multiply and add the same value up to the
// sign of a zero).
template <int c, class C>
__device__ __forceinline__ C c_times(C u) {
  if constexpr (c == 0) return (C)-0.0;
  else if constexpr (c == 1) return u;
  else return -u;
}

template <int Q, class C>
__device__ __forceinline__ void moments(const C *f, C &rho, C &ux, C &uy, C &uz);
template <int Q>
__device__ __forceinline__ void moments_s(const float *g, float &drho, float &rho, float &ux, float &uy, float &uz,
                                          float S);
template <int Q, int i, class C>
__device__ __forceinline__ C cdotu(C ux, C uy, C uz);
template <int Q, int i, class C>
__device__ __forceinline__ C feq_i(C rho, C ux, C uy, C uz, C usq);
template <int Q, class C>
__device__ __forceinline__ C usq_of(C ux, C uy, C uz);
template <int Q>
__device__ __forceinline__ float m15_of(float ux, float uy, float uz);
template <int Q, int i>
__device__ __forceinline__ float geq_i(float drho, float rho, float ux, float uy, float uz, float m15);
template <int Q, int i>
__device__ __forceinline__ float peq_i(float ux, float uy, float uz, float m15);
template <int Q, int i, class C>
__device__ __forceinline__ C qeq_i(C ux, C uy, C uz, C usq);

// Non-equilibrium bounce-back closure of the unknowns (c_x == SGN) of an x-face node with
// prescribed rho, u; transverse momentum corrected through N_a = sum_{c_x=0} f c_a - (2/3) rho u_a
// (reduces to Zou-He's formulas on D2Q9).
template <int Q, int SGN, class C>
__device__ __forceinline__ void nebb_close(C *f, C rho, C ux, C uy, C uz) {
  C ny_ = (C)-0.0, nz_ = (C)-0.0;  // sums from -0 over the nonzero c_a terms, as in moments()
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (Lat<Q>::cx[i] == 0) {
      if constexpr (Lat<Q>::cy[i] != 0) ny_ = ny_ + c_times<Lat<Q>::cy[i]>(f[i]);
      if constexpr (Lat<Q>::cz[i] != 0) nz_ = nz_ + c_times<Lat<Q>::cz[i]>(f[i]);
    }
  });
  ny_ = ny_ - (((C)2 / (C)3) * rho) * uy;
  nz_ = nz_ - (((C)2 / (C)3) * rho) * uz;
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (Lat<Q>::cx[i] == SGN) {
      constexpr int io = Lat<Q>::opp[i];
      const C cu = cdotu<Q, i>(ux, uy, uz);
      f[i] = (f[io] + (((C)6 * (C)Lat<Q>::w[i]) * rho) * cu) -
             ((C)18 * (C)Lat<Q>::w[i]) * (c_times<Lat<Q>::cy[i]>(ny_) + c_times<Lat<Q>::cz[i]>(nz_));
    }
  });
}

// PHYSICAL-mode boundary rules applied to the periodically pulled populations f of a SLOW node
// at (x, y, z) of type t: half-way bounce-back on links whose source is solid (plus the
// moving-wall momentum term 6 w_i rho_w c_i.u_w); inlet nodes either close their c_x = +1
// unknowns for the prescribed velocity (Zou-He) or take f_eq(rho_e, u_in) plus the
// non-equilibrium part of the downstream neighbour in every direction (extrapolated, p.inlet_closure
// = 1); outlet nodes take f_eq(rho_bc, u_nb) plus the non-equilibrium part of the upstream
// neighbour in every direction (zero normal gradient of u and f_neq).  The state is read through
// accessors, so the single-step kernels (global memory) and the multi-step kernel (an LDS tile)
// share one statement of the rules:
//   src_type(IC<i>)  type of the node this node pulls population i from (node - c_i)
//   own(IC<j>)       population j of this node in the state being advanced
//   west(IC<j>)      population j of the node at x - 1;  west_type()  its type
//   east(IC<j>)      population j of the node at x + 1;  east_type()  its type
// Half-way bounce-back of the links whose source is solid: f_i = f_opp(i) of this node, plus
// 6 w_i rho_w c_i.u_w when the source is a moving wall (rho_w taken in compute units, S rho_bc).
// MaskedCodec (not void: lattices stepped from global memory, whose storage codec it is): lm, the
// node's wall link mask, marks the links whose source is solid (bit i - 1 for population i), so
// no source type is read for a resting wall.  The masked links reload the node's own populations
// in place of what the periodic pull read from the solid slots; moving-wall links, on nodes of
// type NODE_WALLMOV (t) only, are told apart through src_type.  Unmasked (the multi-step kernel's
// LDS tile): every link's source type is read.
template <int Q>
constexpr uint32_t links_with_c(int a, int sg) {  // link mask bits (i - 1) of the directions with c_a = sg
  uint32_t m = 0;
  for (int i = 1; i < Q; ++i)
    if ((a == 0 ? Lat<Q>::cx[i] : a == 1 ? Lat<Q>::cy[i] : Lat<Q>::cz[i]) == sg) m |= 1u << (i - 1);
  return m;
}
template <int Q, class MaskedCodec, class C, class SrcType, class Own, class P>
__device__ __forceinline__ void wall_links(const P &p, C *f, C S, uint8_t t, uint32_t lm, SrcType &&src_type,
                                           Own &&own) {
  constexpr bool MASKED = !std::is_void<MaskedCodec>::value;
  const C rw = (C)p.rho_bc * S;
  auto moving_term = [&](auto I) -> C {
    constexpr int i = decltype(I)::value;
    C cu = (C)Lat<Q>::cx[i] * (C)p.u_wall[0] + (C)Lat<Q>::cy[i] * (C)p.u_wall[1];
    if (Lat<Q>::D == 3) cu = cu + (C)Lat<Q>::cz[i] * (C)p.u_wall[2];
    return (((C)6 * (C)Lat<Q>::w[i]) * rw) * cu;
  };
  if constexpr (MASKED) {
    // own(IC<j>) hands out population j of the node in storage form here (MaskedCodec::dec: its
    // compute value).  The links are reloaded by group: the nine directions with c_a = s for
    // (a, s) = (x, +1), (x, -1), (y, +1), ... -- the links of one planar wall are one group.  A
    // group that no lane of the wavefront still needs is skipped, within a group every load is
    // issued before any is decoded (one memory round trip per group), and the links a group
    // reloaded are cleared from the lane's mask for the groups after it.
    uint32_t todo = lm;
    static_for<6>([&](auto Gk) {
      constexpr int a = (int)decltype(Gk)::value / 2, sg = ((int)decltype(Gk)::value % 2) ? -1 : 1;
      constexpr uint32_t gm = links_with_c<Q>(a, sg);
      if constexpr (gm == 0) return;  // D2Q9: no c_z
      const uint32_t mine = todo & gm;
      if (__builtin_amdgcn_ballot_w64(mine != 0) == 0) return;
      decltype(own(IC<0>{})) r[Q];
      static_for<Q - 1>([&](auto I) {
        constexpr int i = decltype(I)::value + 1;
        if constexpr ((gm >> (i - 1)) & 1u) {
          r[i] = 0;
          if ((mine >> (i - 1)) & 1u) r[i] = own(IC<Lat<Q>::opp[i]>{});
        }
      });
      static_for<Q - 1>([&](auto I) {
        constexpr int i = decltype(I)::value + 1;
        if constexpr ((gm >> (i - 1)) & 1u) {  // all loads before any decode
          auto v = r[i];
          asm volatile("" : "+v"(v));
          r[i] = v;
        }
      });
      static_for<Q - 1>([&](auto I) {
        constexpr int i = decltype(I)::value + 1;
        if constexpr ((gm >> (i - 1)) & 1u)
          if ((mine >> (i - 1)) & 1u) f[i] = MaskedCodec::template dec<Lat<Q>::opp[i]>(r[i]);
      });
      todo &= ~gm;
    });
    if (t & NODE_WALLMOV) {
      static_for<Q - 1>([&](auto I) {
        constexpr int i = decltype(I)::value + 1;
        if ((lm >> (i - 1)) & 1u)
          if (src_type(IC<i>{}) & NODE_MOVING) f[i] = f[i] + moving_term(IC<i>{});
      });
    }
    return;
  }
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (i != 0) {
      const uint8_t ts = src_type(I);
      if (ts & NODE_SOLID) {
        C v = own(IC<Lat<Q>::opp[i]>{});
        if (ts & NODE_MOVING) v = v + moving_term(I);
        f[i] = v;
      }
    }
  });
}

// Non-equilibrium extrapolation (Guo, Zheng & Shi, Chinese Physics 11 (2002) 366) from a
// neighbour whose populations nbr(IC<j>) the node would have pulled had the closure not replaced
// them: f_i = f_eq,i(rho_b, u_b) + (g_i - f_eq,i(rho_g, u_g)) with g the neighbour's populations
// and rho_g, u_g their moments.  rho_b < 0 takes rho_b = rho_g (velocity inlet: the two equilibria
// share w_i rho_g, and their difference is taken inside the bracket, g_i + w_i rho_g (qeq_i(u_b) -
// qeq_i(u_g)), about half the arithmetic of the two full equilibria), otherwise u_b = u_g
// (pressure outlet: f_eq is linear in rho, so g_i + (rho_b - rho_g) f_eq,i(1, u_g), one FMA in
// FP32).  Without a fluid neighbour the node takes f_eq(rho_b, u_b) (rho_b = rho_bc).
// SH: populations in the shifted representation (f - w_i, in compute units S); the same rules,
// with the equilibria as deviations from w_i and the density differences taken between density
// deviations (all S-scaled).
template <int Q, bool SH = false, class C, class Nbr, class NbrType, class P>
__device__ __forceinline__ void extrapolate(const P &p, C *f, C S, bool has_nb, C rho_b, bool set_u, C ubx, Nbr &&nbr,
                                            NbrType &&nbr_type) {
  C g[Q], rn = 1, drn = 0, vx = 0, vy = 0, vz = 0;
  bool nb = false;  // the neighbour exists and is not solid
  if (has_nb) {
    // its type and its populations are read together, one round trip (the populations are in
    // bounds whatever the type; a solid neighbour's go unused)
    const uint8_t nt = nbr_type();
    static_for<Q>([&](auto I) { g[decltype(I)::value] = nbr(I); });
    nb = !(nt & NODE_SOLID);
  }
  if (nb) {
    if constexpr (SH) moments_s<Q>(g, drn, rn, vx, vy, vz, S);
    else moments<Q>(g, rn, vx, vy, vz);
  }
  if constexpr (SH) {
    const float drbc = (float)(p.rho_bc - 1.0) * S, rbc = (float)p.rho_bc * S;
    if (set_u) {  // velocity inlet
      const float mb = m15_of<Q>(ubx, 0.0f, 0.0f);
      if (nb) {
        // rho_b = rho_g: the density deviations cancel, f_i = g_i + w_i rho_g (P_b,i - P_g,i)
        const float rs = rn * S, mg = m15_of<Q>(vx, vy, vz);
        static_for<Q>([&](auto I) {
          constexpr int i = decltype(I)::value;
          const float d = peq_i<Q, i>(ubx, 0.0f, 0.0f, mb) - peq_i<Q, i>(vx, vy, vz, mg);
          f[i] = __builtin_fmaf((float)Lat<Q>::w[i] * rs, d, g[i]);
        });
      } else {
        static_for<Q>([&](auto I) {
          constexpr int i = decltype(I)::value;
          f[i] = geq_i<Q, i>(drbc, rbc, ubx, 0.0f, 0.0f, mb);
        });
      }
      return;
    }
    const float usq = usq_of<Q>(vx, vy, vz), mg = m15_of<Q>(vx, vy, vz);
    const float dr = drbc - drn;  // rho_b - rho_g (rho_b = rho_bc for the outlet)
    static_for<Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      if (nb) f[i] = __builtin_fmaf(dr, feq_i<Q, i>(1.0f, vx, vy, vz, usq), g[i]);
      else f[i] = geq_i<Q, i>(drbc, rbc, vx, vy, vz, mg);
    });
    return;
  }
  if (set_u) {  // velocity inlet: u_b = (ubx, 0, 0), rho_b = rho_g
    const C us_b = usq_of<Q>(ubx, (C)0, (C)0);
    if (nb) {  // f_i = g_i + w_i rho_g (qeq_i(u_b) - qeq_i(u_g))
      const C us_g = usq_of<Q>(vx, vy, vz);
      static_for<Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const C d = qeq_i<Q, i>(ubx, (C)0, (C)0, us_b) - qeq_i<Q, i>(vx, vy, vz, us_g);
        if constexpr (sizeof(C) == 4) f[i] = __builtin_fmaf((C)Lat<Q>::w[i] * rn, d, g[i]);
        else f[i] = g[i] + ((C)Lat<Q>::w[i] * rn) * d;
      });
    } else {
      static_for<Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        f[i] = feq_i<Q, i>((C)p.rho_bc, ubx, (C)0, (C)0, us_b);
      });
    }
    return;
  }
  const C usq = usq_of<Q>(vx, vy, vz);
  const C dr = rho_b - rn;
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if (nb) {
      const C e = feq_i<Q, i>((C)1, vx, vy, vz, usq);
      if constexpr (sizeof(C) == 4) f[i] = __builtin_fmaf(dr, e, g[i]);
      else f[i] = g[i] + dr * e;
    } else {
      f[i] = feq_i<Q, i>(rho_b, vx, vy, vz, usq);
    }
  });
}

// S: the compute scale of the populations (Codec::S; 1 unless shifted).  MaskedCodec: wall links
// from the node's link mask lm, own in that codec's storage form (wall_links).
template <int Q, bool SH = false, class MaskedCodec = void, class C, class SrcType, class Own, class West, class WestType,
          class East, class EastType, class P>
__device__ __forceinline__ void physical_fixups(const P &p, int x, int y, int z, uint8_t t, uint32_t lm, C *f,
                                                C S, SrcType &&src_type, Own &&own, West &&west,
                                                WestType &&west_type, East &&east, EastType &&east_type) {
  const C rw = (C)p.rho_bc;
  if (t & NODE_WALLNB)  // only nodes with a solid pull source look at their links
    wall_links<Q, MaskedCodec>(p, f, S, t, lm, src_type, own);
  if ((t & NODE_INLET) && p.inlet_closure == 1) {
    // extrapolated velocity inlet: the node holds rho of its downstream neighbour and u_in
    extrapolate<Q, SH>(p, f, S, x + 1 < p.nx, (C)-1, true, inlet_velocity<C>(p, Lat<Q>::D, y, z), east, east_type);
  } else if (t & NODE_INLET) {
    const C uin = inlet_velocity<C>(p, Lat<Q>::D, y, z);
    C s0 = (C)-0.0, sm = (C)-0.0;  // sums from -0, as in moments()
    static_for<Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      if constexpr (Lat<Q>::cx[i] == 0) s0 += f[i];
      if constexpr (Lat<Q>::cx[i] == -1) sm += f[i];
    });
    // shifted: the weights of the c_x = 0 directions plus twice those of c_x = -1 sum to 1 (the
    // sums unscaled exactly; nebb_close is homogeneous in (f, rho), so it takes S rho)
    const C rho = SH ? (((C)1 + s0 * ((C)1 / S)) + (C)2 * (sm * ((C)1 / S))) / ((C)1 - uin)
                     : (s0 + (C)2 * sm) / ((C)1 - uin);
    nebb_close<Q, 1>(f, SH ? rho * S : rho, uin, (C)0, (C)0);
  } else if (t & NODE_OUTLET) {
    // zero-gradient outlet by non-equilibrium extrapolation (Guo, Zheng & Shi 2002), all i:
    // f_i = f_eq,i(rho_bc, u_nb) + (g_i - f_eq,i(rho_nb, u_nb)), g = the upstream neighbour's
    // post-collision populations; the node then holds rho_bc and u_nb exactly.
    extrapolate<Q, SH>(p, f, S, x > 0, rw, false, (C)0, west, west_type);
  }
}

// PHYSICAL lattices stepped from global memory take their wall links from a per-node link mask
// (wall_links, MaskedCodec).  In 3D lattices the solid nodes of a wavefront that holds a fluid
// node also store the rest equilibrium f_i = w_i
// (0 in the shifted storages) every step, in the same store instructions as the fluid nodes: a
// wavefront whose solid lanes stored nothing would leave every line it writes partly unwritten,
// and partial lines cost a read-modify-write in memory (a 192^3 cavity with x-walls stepped at
// 0.63x the rate of one without).  Wavefronts of solid nodes only store nothing.  Every step
// kernel's wavefront is an aligned block of 64 consecutive node indices, so the rule reads: a
// solid node takes w_i iff its block holds a non-solid node.  Nothing reads the solid slots: the
// links that point at them are reloaded.
template <int Q>
constexpr bool rest_at_solids() { return Q >= 19; }

// PHYSICAL-mode pull from global memory.  The periodic pull is issued by every node before its
// type is known; nodes flagged SLOW then apply the boundary rules (wall neighbours after loading
// their link mask).
template <int Q, int ST, class P>
__device__ __forceinline__ void pull_physical(const P &p, const NodePos &q, uint8_t t,
                                             typename Codec<ST, Q>::C *f) {
  using CD = Codec<ST, Q>;
  using T = typename CD::T;
  pull_periodic<Q, ST>(p, q, f);
  if (!(t & NODE_SLOW)) return;
  const T *fin = (const T *)p.fin;
  constexpr bool MASKED = true;  // every PHYSICAL lattice keeps its link masks (wall_links_kept)
  uint32_t lm = 0;
  if (t & NODE_WALLNB) lm = p.links[q.n];
  physical_fixups<Q, CD::shifted, std::conditional_t<MASKED, CD, void>>(
      p, q.x, q.y, q.z, t, lm, f, (typename CD::C)CD::S,
      [&](auto I) -> uint8_t {
        constexpr int i = decltype(I)::value;
        constexpr int cx = Lat<Q>::cx[i], cy = Lat<Q>::cy[i], cz = Lat<Q>::cz[i];
        int sx = q.x - cx;
        sx = sx < 0 ? sx + p.nx : (sx >= p.nx ? sx - p.nx : sx);
        int sy = q.y - cy;
        sy = sy < 0 ? sy + p.ny : (sy >= p.ny ? sy - p.ny : sy);
        int sz = q.z - cz;
        sz = sz < 0 ? sz + p.nz : (sz >= p.nz ? sz - p.nz : sz);
        return p.type[(uint32_t)sx + (uint32_t)p.nx * ((uint32_t)sy + (uint32_t)p.ny * (uint32_t)sz)];
      },
      [&](auto J) {  // MASKED: the stored form (the wall links decode it)
        constexpr int j = decltype(J)::value;
        if constexpr (MASKED) return CD::template raw<j>(fin + (size_t)j * (size_t)p.stride, q.n);
        else return CD::template ld<j>(fin + (size_t)j * (size_t)p.stride, q.n);
      },
      [&](auto J) {
        constexpr int j = decltype(J)::value;
        return CD::template ld<j>(fin + (size_t)j * (size_t)p.stride, q.n - 1);
      },
      [&]() -> uint8_t { return p.type[q.n - 1]; },
      [&](auto J) {
        constexpr int j = decltype(J)::value;
        return CD::template ld<j>(fin + (size_t)j * (size_t)p.stride, q.n + 1);
      },
      [&]() -> uint8_t { return p.type[q.n + 1]; });
}

// The rest equilibrium in the compute representation: w_i, or 0 for the shifted storages.
template <int Q, int ST>
__device__ __forceinline__ void rest_populations(typename Codec<ST, Q>::C *out) {
  using CD = Codec<ST, Q>;
  using C = typename CD::C;
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    out[i] = CD::shifted ? (C)0 : (C)Lat<Q>::w[i];
  });
}

// ---- moments, overrides, collisions ------------------------------------------------------
// 1/x rounded to nearest for normal x with a normal reciprocal (2^-126 <= |x| < 2^125): the
// hardware reciprocal (error below one ulp) refined by one Newton step y + y (1 - x y), both FMAs
// exact in their products.  Equal bit for bit to the IEEE division 1.0f / x for every such x
// (lbm_check_reciprocal runs the exhaustive comparison on the device); 3 instructions where the
// division takes 10.  Outside that range (a density of 0, inf or NaN: a lattice that has already
// blown up) it may return NaN where the division returns 0 or inf.
__device__ __forceinline__ float recip_rn(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, y, 1.0f), y, y);
}

// rho = sum_i f_i; u = (sum_i c_i f_i) / rho, ascending i (cpp/d2q9_bgk_lid.cpp:131-138); FP32
// multiplies by the correctly rounded 1/rho instead of dividing three times.  The sums start from
// -0, which adds exactly (-0 + x == x for every x), so the compiler drops the first addition; the
// c_i = 0 terms are left out (0 f_i adds an exact zero).  Both change at most the sign of a zero.
template <int Q, class C>
__device__ __forceinline__ void moment_sums(const C *f, C &r, C &jx, C &jy, C &jz) {
  r = (C)-0.0, jx = (C)-0.0, jy = (C)-0.0, jz = (C)-0.0;
  static_for<Q>([&](auto I) { r = r + f[decltype(I)::value]; });
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (Lat<Q>::cx[i] == 1) jx = jx + f[i];
    if constexpr (Lat<Q>::cx[i] == -1) jx = jx - f[i];
  });
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (Lat<Q>::cy[i] == 1) jy = jy + f[i];
    if constexpr (Lat<Q>::cy[i] == -1) jy = jy - f[i];
  });
  if constexpr (Lat<Q>::D == 3) {
    static_for<Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      if constexpr (Lat<Q>::cz[i] == 1) jz = jz + f[i];
      if constexpr (Lat<Q>::cz[i] == -1) jz = jz - f[i];
    });
  }
}

template <int Q, class C>
__device__ __forceinline__ void moments(const C *f, C &rho, C &ux, C &uy, C &uz) {
  C r, jx, jy, jz;
  moment_sums<Q>(f, r, jx, jy, jz);
  rho = r;
  if constexpr (sizeof(C) == 4) {  // FP32: one correctly rounded reciprocal, then products
    const C inv = recip_rn(r);
    ux = jx * inv;
    uy = jy * inv;
    uz = (Lat<Q>::D == 3) ? jz * inv : (C)0;
  } else {
    ux = jx / r;
    uy = jy / r;
    uz = (Lat<Q>::D == 3) ? jz / r : (C)0;
  }
}

// Shifted representation (g_i = f_i - w_i, in compute units S): S drho = sum_i S g_i, the
// density deviation, keeps the precision of the small g; rho = 1 + drho; u = (sum_i c_i g_i) /
// rho, since sum_i c_i w_i = 0.  drho is returned in compute units, rho and u unscaled; the
// unscaling is folded into one FMA and the reciprocal (exact powers of two).
template <int Q>
__device__ __forceinline__ void moments_s(const float *g, float &drho, float &rho, float &ux, float &uy, float &uz,
                                          float S) {
  float r, jx, jy, jz;
  moment_sums<Q>(g, r, jx, jy, jz);
  drho = r;
  rho = __builtin_fmaf(r, 1.0f / S, 1.0f);
  const float inv = recip_rn(rho) * (1.0f / S);
  ux = jx * inv;
  uy = jy * inv;
  uz = (Lat<Q>::D == 3) ? jz * inv : 0.0f;
}

// REFERENCE Dirichlet overrides, in the reference's order: lid velocity, inlet x-velocity,
// solid rho = u = 0, fixed-density columns (cpp/d2q9_bgk_lid.cpp:141-146,
// rust/examples/d2q9_bgk_channel.rs:190-201, rust/examples/d3q27_mrt_lid.rs:323-330).
template <int Q, class C, class P>
__device__ __forceinline__ void overrides(const P &p, uint8_t t, int y, int z, C &rho, C &ux,
                                          C &uy, C &uz) {
  if (t & NODE_MOVING) { ux = (C)p.u_wall[0]; uy = (C)p.u_wall[1]; uz = (C)p.u_wall[2]; }
  if (t & NODE_INLET) ux = inlet_velocity<C>(p, Lat<Q>::D, y, z);
  if (t & NODE_SOLID) { ux = 0; uy = 0; uz = 0; rho = 0; }
  if (t & NODE_RHO_FIX) rho = (C)p.rho_bc;
}

// The same overrides in the shifted representation: drho (compute units S) follows rho.
template <int Q, class P>
__device__ __forceinline__ void overrides_s(const P &p, uint8_t t, int y, int z, float &drho, float &rho, float &ux,
                                            float &uy, float &uz, float S) {
  overrides<Q>(p, t, y, z, rho, ux, uy, uz);
  if (t & NODE_SOLID) drho = -S;
  if (t & NODE_RHO_FIX) drho = (float)(p.rho_bc - 1.0) * S;
}

// c_i . u evaluated as the reference does: (cx ux + cy uy) + cz uz with exact +-1/0 products.
template <int Q, int i, class C>
__device__ __forceinline__ C cdotu(C ux, C uy, C uz) {
  C e = c_times<Lat<Q>::cx[i]>(ux) + c_times<Lat<Q>::cy[i]>(uy);
  if constexpr (Lat<Q>::D == 3) e = e + c_times<Lat<Q>::cz[i]>(uz);
  return e;
}

// f_eq,i = w_i rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u).  FP64 evaluates the reference's expression
// term by term (cpp/d2q9_bgk_lid.cpp:149-151); FP32 uses the multiply-add form
// w_i rho (A + eu (3 + 4.5 eu)) with A = 1 - 1.5 u.u, three FMAs per direction.  usq_of returns u.u
// in FP64 and A in FP32.
// The bracket of f_eq,i: f_eq,i = (w_i rho) qeq_i.
template <int Q, int i, class C>
__device__ __forceinline__ C qeq_i(C ux, C uy, C uz, C usq) {
  const C eu = cdotu<Q, i>(ux, uy, uz);
  if constexpr (sizeof(C) == 4 && Lat<Q>::cx[i] == 0 && Lat<Q>::cy[i] == 0 && Lat<Q>::cz[i] == 0) {
    return usq;  // eu = 0: the FMAs below return A exactly
  } else if constexpr (sizeof(C) == 4) {
    const C t = __builtin_fmaf(4.5f, eu, 3.0f);
    return __builtin_fmaf(eu, t, usq);
  } else {
    return (((C)1 + (C)3 * eu) + (C)4.5 * (eu * eu)) - (C)1.5 * usq;
  }
}
template <int Q, int i, class C>
__device__ __forceinline__ C feq_i(C rho, C ux, C uy, C uz, C usq) {
  return ((C)Lat<Q>::w[i] * rho) * qeq_i<Q, i>(ux, uy, uz, usq);
}

template <int Q, class C>
__device__ __forceinline__ C usq_of(C ux, C uy, C uz) {
  if constexpr (sizeof(C) == 4) {
    C u2 = __builtin_fmaf(uy, uy, ux * ux);
    if constexpr (Lat<Q>::D == 3) u2 = __builtin_fmaf(uz, uz, u2);
    return __builtin_fmaf(-1.5f, u2, 1.0f);
  } else {
    C u2 = ux * ux + uy * uy;
    if constexpr (Lat<Q>::D == 3) u2 = u2 + uz * uz;
    return u2;
  }
}

// Shifted equilibrium g_eq,i = f_eq,i - w_i = w_i (drho + rho P_i), P_i = 3 c.u + 4.5 (c.u)^2
// - 1.5 u.u: every term small, so its FP32 rounding is relative to the deviation, not to w_i.
// m15_of returns -1.5 u.u.
template <int Q>
__device__ __forceinline__ float m15_of(float ux, float uy, float uz) {
  float u2 = __builtin_fmaf(uy, uy, ux * ux);
  if constexpr (Lat<Q>::D == 3) u2 = __builtin_fmaf(uz, uz, u2);
  return -1.5f * u2;
}
template <int Q, int i>
__device__ __forceinline__ float peq_i(float ux, float uy, float uz, float m15) {  // P_i
  if constexpr (Lat<Q>::cx[i] == 0 && Lat<Q>::cy[i] == 0 && Lat<Q>::cz[i] == 0) {
    return m15;
  } else {
    const float eu = cdotu<Q, i>(ux, uy, uz);
    return __builtin_fmaf(eu, __builtin_fmaf(4.5f, eu, 3.0f), m15);
  }
}
template <int Q, int i>
__device__ __forceinline__ float geq_i(float drho, float rho, float ux, float uy, float uz, float m15) {
  return (float)Lat<Q>::w[i] * __builtin_fmaf(rho, peq_i<Q, i>(ux, uy, uz, m15), drho);
}

// F = omega f_eq + (1 - omega) f  (cpp/d2q9_bgk_lid.cpp:149-153); FP32: f + omega (f_eq - f);
// SH (shifted representation): g + omega (g_eq - g).
template <int Q, bool SH = false, class C, class P>
__device__ __forceinline__ void collide_bgk(const P &p, const C *f, C rho, C ux, C uy, C uz,
                                            C *out, C drho = (C)0) {
  if constexpr (SH) {
    const float omega = (float)p.omega, m15 = m15_of<Q>(ux, uy, uz);
    static_for<Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      out[i] = __builtin_fmaf(omega, geq_i<Q, i>(drho, rho, ux, uy, uz, m15) - f[i], f[i]);
    });
    return;
  }
  const C omega = (C)p.omega, om1 = (C)1 - omega;
  const C usq = usq_of<Q>(ux, uy, uz);
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    if constexpr (sizeof(C) == 4) out[i] = __builtin_fmaf(omega, feq_i<Q, i>(rho, ux, uy, uz, usq) - f[i], f[i]);
    else out[i] = omega * feq_i<Q, i>(rho, ux, uy, uz, usq) + om1 * f[i];
  });
}

// ---- MRT (D3Q27) -------------------------------------------------------------------------
template <int CO> __device__ __forceinline__ constexpr int mcoef(int k, int i) {
  return CO == CO_MRT_RAW ? MRAW[k][i] : MGS[k][i];
}
template <class C, int CO> __device__ __forceinline__ constexpr C minv_coef(int i, int k) {
  return CO == CO_MRT_RAW ? (C)MRAW_INV24[i][k] / (C)24 : (C)MGS[k][i] / (C)MGS_NORM2[k];
}

// acc + a*c with c a compile-time coefficient; +-1 are exact adds, others a rounded product
// (FP64) or one FMA (FP32).
template <class C, int ci>
__device__ __forceinline__ C add_int_term(C acc, C a) {
  if constexpr (ci == 1) return acc + a;
  else if constexpr (ci == -1) return acc - a;
  else if constexpr (ci == 0) return acc;
  else if constexpr (sizeof(C) == 4) return __builtin_fmaf(a, (C)ci, acc);
  else return acc + a * (C)ci;
}

// out_k = sum_i M[k][i] v_i, ascending i
template <class C, int CO>
__device__ __forceinline__ void mrt_forward(const C *v, C *out) {
  static_for<27>([&](auto K) {
    constexpr int k = decltype(K)::value;
    C acc = 0;
    static_for<27>([&](auto I) {
      constexpr int i = decltype(I)::value;
      acc = add_int_term<C, mcoef<CO>(k, i)>(acc, v[i]);
    });
    out[k] = acc;
  });
}

// out_i = sum_k Minv[i][k] m_k, ascending k
template <class C, int CO>
__device__ __forceinline__ void mrt_inverse(const C *m, C *out) {
  static_for<27>([&](auto I) {
    constexpr int i = decltype(I)::value;
    C acc = 0;
    static_for<27>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr C c = minv_coef<C, CO>(i, k);
      if constexpr (c != (C)0) {
        if constexpr (sizeof(C) == 4) acc = __builtin_fmaf(m[k], c, acc);
        else acc = acc + m[k] * c;
      }
    });
    out[i] = acc;
  });
}

// Reference equilibrium-moment table (rust/examples/d3q27_mrt_lid.rs:341-383), row 26 quirk kept.
template <class C>
__device__ __forceinline__ void meq_reference(C r, C ux, C uy, C uz, C *meq) {
  const C cs2 = (C)1 / (C)3, cs4 = cs2 * cs2, cs6 = cs2 * cs2 * cs2;
  const C rvx = r * ux, rvy = r * uy, rvz = r * uz, rc2 = cs2 * r, rc4 = cs4 * r, rc6 = cs6 * r;
  const C vx2 = ux * ux, vy2 = uy * uy, vz2 = uz * uz, v2 = vx2 + vy2 + vz2;
  meq[0] = r; meq[1] = rvx; meq[2] = rvy; meq[3] = rvz;
  meq[4] = rvx * uy; meq[5] = rvx * uz; meq[6] = rvy * uz;
  meq[7] = r + rvx * ux + rvy * uy + rvz * uz;
  meq[8] = r * (vx2 - vy2); meq[9] = r * (vx2 - vz2);
  meq[10] = rc2 * ux; meq[11] = rc2 * ux; meq[12] = rc2 * uy; meq[13] = rc2 * uz;
  meq[14] = rc2 * uy; meq[15] = rc2 * uz; meq[16] = 0;
  meq[17] = rc2 * (cs2 + vx2 + vy2); meq[18] = rc2 * (cs2 + vx2 + vz2); meq[19] = rc2 * (cs2 + vy2 + vz2);
  meq[20] = rc2 * uy * uz; meq[21] = rc2 * ux * uz; meq[22] = rc2 * ux * uy;
  meq[23] = rc4 * ux; meq[24] = rc4 * uy; meq[25] = rc4 * uz;
  meq[26] = v2 + rc6;
}

// The same table less its value at rest (rho = 1, u = 0), from drho = rho - 1: the rows of the
// shifted representation, in compute units S (dr and r given as S drho and S rho; the one term
// without a density factor, u.u in row 26, is scaled here).
__device__ __forceinline__ void meq_reference_s(float dr, float r, float ux, float uy, float uz, float *meq, float S) {
  const float cs2 = 1.0f / 3.0f, cs4 = cs2 * cs2, cs6 = cs2 * cs2 * cs2;
  const float rvx = r * ux, rvy = r * uy, rvz = r * uz, rc2 = cs2 * r, rc4 = cs4 * r;
  const float vx2 = ux * ux, vy2 = uy * uy, vz2 = uz * uz, v2 = vx2 + vy2 + vz2;
  meq[0] = dr; meq[1] = rvx; meq[2] = rvy; meq[3] = rvz;
  meq[4] = rvx * uy; meq[5] = rvx * uz; meq[6] = rvy * uz;
  meq[7] = dr + rvx * ux + rvy * uy + rvz * uz;
  meq[8] = r * (vx2 - vy2); meq[9] = r * (vx2 - vz2);
  meq[10] = rc2 * ux; meq[11] = rc2 * ux; meq[12] = rc2 * uy; meq[13] = rc2 * uz;
  meq[14] = rc2 * uy; meq[15] = rc2 * uz; meq[16] = 0;
  meq[17] = cs4 * dr + rc2 * (vx2 + vy2); meq[18] = cs4 * dr + rc2 * (vx2 + vz2);
  meq[19] = cs4 * dr + rc2 * (vy2 + vz2);
  meq[20] = rc2 * uy * uz; meq[21] = rc2 * ux * uz; meq[22] = rc2 * ux * uy;
  meq[23] = rc4 * ux; meq[24] = rc4 * uy; meq[25] = rc4 * uz;
  meq[26] = v2 * S + cs6 * dr;
}

// FP32 MRT, the relaxation below (out = f - M^-1 S (M f - meq), S acting in the collision's
// moment basis) through the tensor-product structure of D3Q27:
//   mu = P f, the monomial moments sum_i cx^a cy^b cz^c f_i (a, b, c in 0..2): three passes of the
//      1D transform (f_-, f_0, f_+) -> (f_- + f_0 + f_+, f_+ - f_-, f_+ + f_-), 81 additions;
//   d = A mu - meq_raw, the raw-row deviations (the raw rows are monomials except x^2+y^2+z^2,
//      x^2-y^2, x^2-z^2; meq_raw the closed-form equilibrium moments or the reference table);
//   r = T_GS^-1 diag(s) T_GS d (Gram-Schmidt basis; T_GS integral lower triangular with 64
//      non-zeros, its inverse 63) or diag(s) d (raw basis), the relaxed raw-row deviations;
//   mu' = P M^-1 r, the monomial moments of the correction (the identity but for the three
//      mixed rows);
//   out = f - P^-1 mu': three passes of (r0, r1, r2) -> (r2 - r1, r0 - r2, r2 + r1), whose result
//      for direction c carries a factor 2^|c| taken out by the final FMA (|c| = non-zero
//      components; the scalings are exact).
// About 400 operations a node against 820 for the raw integer matrices or 1640 for M_GS, M_GS f_eq
// and M_GS^-1 directly, with the 27 rates as the only run-time coefficients.  In exact arithmetic
// all are the same map.
namespace d27 {
// direction of the D3Q27 table at grid slot (cx + 1) + 3 (cy + 1) + 9 (cz + 1)
struct Grid {
  int dir[27];
};
constexpr Grid make_grid() {
  Grid g{};
  for (int i = 0; i < 27; ++i) g.dir[(Lat<27>::cx[i] + 1) + 3 * (Lat<27>::cy[i] + 1) + 9 * (Lat<27>::cz[i] + 1)] = i;
  return g;
}
constexpr Grid kGrid = make_grid();
// slot o + k * step, k = 0..2, of the 9 lines along one axis (step 1, 3, 9)
template <int STEP>
__device__ __forceinline__ constexpr int line_base(int l) {
  return STEP == 1 ? 3 * l : (STEP == 3 ? (l % 3) + 9 * (l / 3) : l);
}
template <int STEP>
__device__ __forceinline__ void forward_pass(float *v) {
  static_for<9>([&](auto L) {
    constexpr int o = line_base<STEP>(decltype(L)::value);
    const float fm = v[o], f0 = v[o + STEP], fp = v[o + 2 * STEP];
    const float sp = fp + fm;
    v[o] = sp + f0;
    v[o + STEP] = fp - fm;
    v[o + 2 * STEP] = sp;
  });
}
template <int STEP>
__device__ __forceinline__ void inverse_pass(float *v) {
  static_for<9>([&](auto L) {
    constexpr int o = line_base<STEP>(decltype(L)::value);
    const float r0 = v[o], r1 = v[o + STEP], r2 = v[o + 2 * STEP];
    v[o] = r2 - r1;
    v[o + STEP] = r0 - r2;
    v[o + 2 * STEP] = r2 + r1;
  });
}
// the monomial basis reproduces MRAW: raw row k = sum_t coef_t monomial_t
constexpr bool raw_rows_match() {
  for (int k = 0; k < 27; ++k)
    for (int i = 0; i < 27; ++i) {
      int acc = 0;
      for (int t = 0; t < MRT_RAW_NT[k]; ++t) {
        const int j = MRT_RAW_MONO[k][t], a = j % 3, b = (j / 3) % 3, c = j / 9;
        int m = 1;
        for (int e = 0; e < a; ++e) m *= Lat<27>::cx[i];
        for (int e = 0; e < b; ++e) m *= Lat<27>::cy[i];
        for (int e = 0; e < c; ++e) m *= Lat<27>::cz[i];
        acc += MRT_RAW_COEF[k][t] * m;
      }
      if (acc != MRAW[k][i]) return false;
    }
  return true;
}
static_assert(raw_rows_match(), "monomial tables disagree with MRAW");
}  // namespace d27

// SH: f holds g = f - w_i and the raw equilibrium moments are taken less their value at rest
// (meq(rho, u) - meq(1, 0) = c_0 drho + rho (the u terms), c_0 the constant column), so the
// deviations are differences of small numbers; the passes below are linear and map g to g'.
// In compute units S: f, rho and drho come scaled by S (S the codec's compute scale).
// The monomial moments mu = P f of a node's populations (grid order: slot a + 3 b + 9 c holds
// sum_i cx^a cy^b cz^c f_i with exponent 2 for a squared component): three passes of the 1D
// transform.  mu[0] = sum f, mu[1], mu[3], mu[9] = the momentum components.
__device__ __forceinline__ void mrt27_monomials(const float *f, float *v) {
  static_for<27>([&](auto G) { v[decltype(G)::value] = f[d27::kGrid.dir[decltype(G)::value]]; });
  d27::forward_pass<1>(v);
  d27::forward_pass<3>(v);
  d27::forward_pass<9>(v);
}

// rho and u from the monomial moments: the FP32 MRT step takes them from the forward transform it
// needs anyway instead of separate sums (the summation order of the tensor-product passes, not the
// ascending one).  SH: drho in compute units S, as moments_s.
template <bool SH>
__device__ __forceinline__ void moments_mono(const float *v, float &drho, float &rho, float &ux, float &uy, float &uz,
                                             float S) {
  if constexpr (SH) {
    drho = v[0];
    rho = __builtin_fmaf(v[0], 1.0f / S, 1.0f);
    const float inv = recip_rn(rho) * (1.0f / S);
    ux = v[1] * inv;
    uy = v[3] * inv;
    uz = v[9] * inv;
  } else {
    rho = v[0];
    const float inv = recip_rn(rho);
    ux = v[1] * inv;
    uy = v[3] * inv;
    uz = v[9] * inv;
  }
}

// mono: the monomial moments of f (mrt27_monomials), or null to compute them here.
template <int CO, bool SH = false, class P>
__device__ __forceinline__ void collide_mrt_fp32(const P &p, const float *f, float rho, float ux,
                                                 float uy, float uz, float *out, float drho = 0.0f,
                                                 float S = 1.0f, const float *mono = nullptr) {
  float d[27];  // the equilibrium raw moments, then the deviations
  if (CO == CO_MRT_RAW && p.meq_mode == 0) {
    if constexpr (SH) meq_reference_s(drho, rho, ux, uy, uz, d, S);
    else meq_reference<float>(rho, ux, uy, uz, d);
  } else {
    const float phi[10] = {1.0f, ux, uy, uz, ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz};
    static_for<27>([&](auto K) {
      constexpr int k = decltype(K)::value;
      float acc = 0.0f;
      static_for<10>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr float c = (float)MEQ_RAW[k][j];
        if constexpr (c != 0.0f && !(SH && j == 0)) acc = __builtin_fmaf(c, phi[j], acc);
      });
      if constexpr (SH) {
        constexpr float c0 = (float)MEQ_RAW[k][0];
        if constexpr (c0 != 0.0f) d[k] = __builtin_fmaf(rho, acc, c0 * drho);
        else d[k] = rho * acc;
      } else {
        d[k] = rho * acc;
      }
    });
  }
  float v[27];
  if (mono) static_for<27>([&](auto G) { v[decltype(G)::value] = mono[decltype(G)::value]; });
  else mrt27_monomials(f, v);
  static_for<27>([&](auto K) {
    constexpr int k = decltype(K)::value;
    float raw = v[MRT_RAW_MONO[k][0]];
    static_for<MRT_RAW_NT[k] - 1>([&](auto T) {
      constexpr int t = decltype(T)::value + 1;
      if constexpr (MRT_RAW_COEF[k][t] == 1) raw = raw + v[MRT_RAW_MONO[k][t]];
      else raw = raw - v[MRT_RAW_MONO[k][t]];
    });
    d[k] = raw - d[k];
  });
  float r[27];
  if constexpr (CO == CO_MRT_GS) {
    float g[27];
    static_for<27>([&](auto K) {  // Gram-Schmidt moments of the deviation, relaxed
      constexpr int k = decltype(K)::value;
      float acc = -0.0f;  // adds exactly; the compiler drops it
      static_for<k + 1>([&](auto J) {  // T_GS is lower triangular
        constexpr int j = decltype(J)::value;
        acc = add_int_term<float, T_GS[k][j]>(acc, d[j]);
      });
      g[k] = p.s32[k] * acc;
    });
    static_for<27>([&](auto K) {
      constexpr int k = decltype(K)::value;
      float acc = -0.0f;
      static_for<k + 1>([&](auto J) {  // so is its inverse
        constexpr int j = decltype(J)::value;
        constexpr float c = (float)TINV_GS[k][j];
        if constexpr (c == 1.0f) acc = acc + g[j];
        else if constexpr (c == -1.0f) acc = acc - g[j];
        else if constexpr (c != 0.0f) acc = __builtin_fmaf(c, g[j], acc);
      });
      r[k] = acc;
    });
  } else {
    static_for<27>([&](auto K) { r[decltype(K)::value] = p.s32[decltype(K)::value] * d[decltype(K)::value]; });
  }
  static_for<27>([&](auto U) {  // monomial moments of the correction
    constexpr int u = decltype(U)::value;
    float acc = -0.0f;
    static_for<27>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr float c = (float)MRT_AINV[u][k];
      if constexpr (c == 1.0f) acc = acc + r[k];
      else if constexpr (c == -1.0f) acc = acc - r[k];
      else if constexpr (c != 0.0f) acc = __builtin_fmaf(c, r[k], acc);
    });
    v[u] = acc;
  });
  d27::inverse_pass<1>(v);
  d27::inverse_pass<3>(v);
  d27::inverse_pass<9>(v);
  static_for<27>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr int g = (Lat<27>::cx[i] + 1) + 3 * (Lat<27>::cy[i] + 1) + 9 * (Lat<27>::cz[i] + 1);
    constexpr int nz = (Lat<27>::cx[i] != 0) + (Lat<27>::cy[i] != 0) + (Lat<27>::cz[i] != 0);
    constexpr float w = nz == 0 ? 1.0f : (nz == 1 ? 0.5f : (nz == 2 ? 0.25f : 0.125f));
    out[i] = __builtin_fmaf(-w, v[g], f[i]);
  });
}

// m = M f; m* = m - s (m - meq); out = M^-1 m*   (rust/examples/d3q27_mrt_lid.rs:336,385-387)
template <int CO, bool SH = false, class C, class P>
__device__ __forceinline__ void collide_mrt_scalar(const P &p, const C *f, C rho, C ux, C uy,
                                                   C uz, C *out, C drho = (C)0, float S = 1.0f) {
  if constexpr (sizeof(C) == 4) {
    collide_mrt_fp32<CO, SH>(p, f, rho, ux, uy, uz, out, drho, S);
    return;
  }
  C m[27], meq[27];
  mrt_forward<C, CO>(f, m);
  if (CO == CO_MRT_RAW && p.meq_mode == 0) {
    meq_reference<C>(rho, ux, uy, uz, meq);
  } else {
    C fe[27];
    const C usq = usq_of<27>(ux, uy, uz);
    static_for<27>([&](auto I) {
      constexpr int i = decltype(I)::value;
      fe[i] = feq_i<27, i>(rho, ux, uy, uz, usq);
    });
    mrt_forward<C, CO>(fe, meq);
  }
  static_for<27>([&](auto K) {
    constexpr int k = decltype(K)::value;
    m[k] = m[k] - (m[k] - meq[k]) * (C)p.s[k];
  });
  mrt_inverse<C, CO>(m, out);
}

// Parity of the D3Q19 orthogonal rows under c -> -c.  Every row is a polynomial in c of definite
// parity, so m_k sums over the 9 opposite-pair sums (even rows, plus the rest population) or the
// 9 pair differences (odd rows), and the inverse builds both populations of a pair from one even
// and one odd sum.
namespace d19 {
constexpr int pair_rep(int p) {  // the p-th direction i with i < opp(i)
  int n = 0;
  for (int i = 1; i < 19; ++i)
    if (i < Lat<19>::opp[i]) {
      if (n == p) return i;
      ++n;
    }
  return -1;
}
constexpr bool row_parity(int k, int sign) {
  for (int i = 0; i < 19; ++i)
    if (M19[k][i] != sign * M19[k][Lat<19>::opp[i]]) return false;
  return true;
}
constexpr bool every_row_has_parity() {
  for (int k = 0; k < 19; ++k)
    if (!row_parity(k, 1) && !row_parity(k, -1)) return false;
  return true;
}
static_assert(every_row_has_parity(), "each D3Q19 orthogonal row is even or odd in c");
static_assert(pair_rep(8) > 0 && pair_rep(9) == -1, "D3Q19 has 9 opposite pairs");
}  // namespace d19

// D3Q19 MRT in the orthogonal basis M19 (d'Humieres et al. 2002): m = M19 f, projected
// equilibrium moments meq = M19 f_eq (closed form MEQ19), m* = m - s (m - meq),
// out = M19^T diag(1 / |row|^2) m*.  FP32: out = f - M19^T ((s / |row|^2) (m - meq)), FMAs, both
// transforms in the pair form above (half the terms of the dense rows).  SH: f holds g = f - w_i
// and meq is taken less its value at rest, as in collide_mrt_fp32.
template <bool SH = false, class C, class P>
__device__ __forceinline__ void collide_mrt19(const P &p, const C *f, C rho, C ux, C uy, C uz, C *out,
                                              C drho = (C)0) {
  const C phi[10] = {(C)1, ux, uy, uz, ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz};
  C d[19], ps[9], pd[9];  // FP32: sums and differences of the opposite pairs
  if constexpr (sizeof(C) == 4)
    static_for<9>([&](auto PI) {
      constexpr int i = d19::pair_rep(decltype(PI)::value), o = Lat<19>::opp[i];
      ps[decltype(PI)::value] = f[i] + f[o];
      pd[decltype(PI)::value] = f[i] - f[o];
    });
  static_for<19>([&](auto K) {
    constexpr int k = decltype(K)::value;
    C meq = (C)-0.0;
    static_for<10>([&](auto J) {
      constexpr int j = decltype(J)::value;
      constexpr C c = (C)MEQ19[k][j];
      if constexpr (c != (C)0 && !(SH && j == 0)) {
        if constexpr (sizeof(C) == 4) meq = __builtin_fmaf(c, phi[j], meq);
        else meq = meq + c * phi[j];
      }
    });
    C m = (C)-0.0;
    if constexpr (sizeof(C) == 4) {
      if constexpr (d19::row_parity(k, 1)) {
        m = add_int_term<C, M19[k][0]>(m, f[0]);
        static_for<9>([&](auto PI) { m = add_int_term<C, M19[k][d19::pair_rep(decltype(PI)::value)]>(m, ps[decltype(PI)::value]); });
      } else {
        static_for<9>([&](auto PI) { m = add_int_term<C, M19[k][d19::pair_rep(decltype(PI)::value)]>(m, pd[decltype(PI)::value]); });
      }
    } else {
      static_for<19>([&](auto I) { m = add_int_term<C, M19[k][decltype(I)::value]>(m, f[decltype(I)::value]); });
    }
    if constexpr (SH) {
      constexpr float c0 = (float)MEQ19[k][0];
      if constexpr (c0 != 0.0f) meq = __builtin_fmaf(rho, meq, c0 * drho);
      else meq = rho * meq;
    } else {
      meq = rho * meq;
    }
    if constexpr (sizeof(C) == 4) d[k] = (p.s32[k] * (1.0f / (float)M19_NORM2[k])) * (m - meq);
    else d[k] = m - (m - meq) * (C)p.s[k];  // m*
  });
  if constexpr (sizeof(C) == 4) {
    // out_i = f_i - sum_k M19[k][i] d_k: for a pair (i, opp i) the even rows give E, the odd rows
    // +-O, so out_i = f_i - (E + O) and out_opp = f_opp - (E - O)
    C e0 = (C)-0.0;
    static_for<19>([&](auto K) {
      constexpr int k = decltype(K)::value;
      if constexpr (d19::row_parity(k, 1)) e0 = add_int_term<C, M19[k][0]>(e0, d[k]);
    });
    out[0] = f[0] - e0;
    static_for<9>([&](auto PI) {
      constexpr int i = d19::pair_rep(decltype(PI)::value), o = Lat<19>::opp[i];
      C e = (C)-0.0, q = (C)-0.0;
      static_for<19>([&](auto K) {
        constexpr int k = decltype(K)::value;
        if constexpr (d19::row_parity(k, 1)) e = add_int_term<C, M19[k][i]>(e, d[k]);
        else q = add_int_term<C, M19[k][i]>(q, d[k]);
      });
      out[i] = f[i] - (e + q);
      out[o] = f[o] - (e - q);
    });
  } else {
    static_for<19>([&](auto I) {
      constexpr int i = decltype(I)::value;
      C acc = (C)-0.0;
      static_for<19>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr C c = (C)M19[k][i] / (C)M19_NORM2[k];
        if constexpr (c != (C)0) acc = acc + d[k] * c;
      });
      out[i] = acc;
    });
  }
}

// SH: f and out in the shifted representation (f - w_i) in compute units S (the codec's scale),
// drho = rho - 1 kept separately in the same units, rho unscaled.  The collisions are homogeneous
// of degree one in (f, rho, drho) but for the reference table's u.u term (meq_reference_s), so
// they take S rho.
template <int Q, int CO, bool SH = false, class C, class P>
__device__ __forceinline__ void collide(const P &p, const C *f, C rho, C ux, C uy, C uz, C *out, C drho, C S) {
  const C rs = SH ? rho * S : rho;
  if constexpr (CO == CO_BGK) collide_bgk<Q, SH>(p, f, rs, ux, uy, uz, out, drho);
  else if constexpr (Q == 19) collide_mrt19<SH>(p, f, rs, ux, uy, uz, out, drho);
  else collide_mrt_scalar<CO, SH>(p, f, rs, ux, uy, uz, out, drho, (float)S);
}

// Stores the node's Q populations: out in the compute representation, rnd the node's rounding
// bits (node_bits; shifted codecs only).
template <int Q, int ST, bool NT = false, class P>
__device__ __forceinline__ void store_all(const P &p, uint32_t n, const typename Codec<ST, Q>::C *out, uint32_t rnd) {
  using CD = Codec<ST, Q>;
  using T = typename CD::T;
  T *fout = (T *)p.fout;
  uint32_t off = n * (uint32_t)sizeof(T);
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    // 2-byte storages: the population's (uniform) base plus a 32-bit byte offset, the scalar-base
    // store form with one offset register for all Q stores (lbm_create keeps these lattices below
    // 2^31 nodes).  The empty asm hides that the offsets are equal: the compiler would otherwise
    // form fout + offset once and add each population's base to it in 64-bit vector arithmetic,
    // one instruction per store.  Wider storages take 64-bit addresses.
    T *base = fout + (size_t)i * (size_t)p.stride;
    T *dst = base + n;
    if constexpr (sizeof(T) == 2) {
      asm volatile("" : "+v"(off));
      dst = (T *)((char *)base + off);
    }
    const T v = CD::template enc<i>(out[i], pop_bits<i>(rnd));
    // NT: non-temporal stores for lattices too large for the caches (the next step reads these
    // populations from HBM anyway).  They raised the bare D3Q27 pull from 5.35 to 5.56 TB/s
    // (4-byte) and from 5.32 to 5.62 TB/s (2-byte) on MI355X, and the 3D step kernels by 3-5 %; a
    // lattice that stays in the caches between steps runs faster with ordinary stores.
    if constexpr (NT) __builtin_nontemporal_store(v, dst);
    else *dst = v;
  });
}

// Stores populations given unshifted (initial states), rounded to nearest.
template <int Q, int ST, class P>
__device__ __forceinline__ void store_all_f(const P &p, uint32_t n, const typename Codec<ST, Q>::C *f) {
  using CD = Codec<ST, Q>;
  using T = typename CD::T;
  T *fout = (T *)p.fout;
  static_for<Q>([&](auto I) {
    constexpr int i = decltype(I)::value;
    fout[(size_t)i * (size_t)p.stride + n] = CD::template enc_f<i>(f[i]);
  });
}

}  // namespace lbm
