// lbm_multistep.hip — K D2Q9 BGK time steps per launch from an LDS tile (temporal blocking).
//
// A workgroup owns an interior tile of TX x TY nodes.  It loads the tile plus a halo of K nodes
// on every side (periodic wrap, so every region position holds the state of one lattice node),
// then advances the region K times in LDS: sub-step s recomputes the positions at least s away
// from the region's edge, whose pulls (radius 1, the outlet's upstream neighbour included) all lie
// in the window of sub-step s - 1.  After K sub-steps the interior is exact and is written out.
// Each sub-step runs the per-node rules of the single-step kernel (the same pulls, boundary
// fix-ups, moments and collision, on the same values), and every intermediate population is
// rounded through the storage codec exactly as a store and a load would round it, so K sub-steps
// reproduce K single steps bit for bit.
//
// PHYSICAL lattices whose periodic wrap carries nothing (x: inlet / solid column 0 and outlet /
// solid column nx - 1; y: solid rows 0 and ny - 1) hold the halo positions beyond that wrap as
// walls: their values never reach the interior, so updating them would be wasted work.
//
// HBM traffic per launch: the region once in (halo factor (TX + 2K)(TY + 2K) / (TX TY)) and the
// interior once out, for K steps, against 2 Q s bytes per node for each single step.  The D2Q9
// 1024 x 256 lattice (18.9 MB per step) is latency-bound as single steps; this kernel trades the
// per-step launch and memory round trip for LDS traffic and redundant halo updates.
#include "lbm_device.hpp"

namespace lbm {

template <int ST, int TX_, int K, int TY, int NT, bool DB>
struct MultiShape {
  static constexpr int Q = 9, TX = TX_;
  static constexpr int RX = TX + 2 * K, RY = TY + 2 * K, RN = RX * RY;
  static constexpr int NPOS = (RN + NT - 1) / NT;  // region positions per thread
  static constexpr int NBUF = DB ? 2 : 1;         // LDS copies of the region
};

// One tile of one lattice, K steps.  DB: two LDS copies of the region, sub-step s reads copy
// (s - 1) & 1 and writes copy s & 1 (one barrier per sub-step); otherwise one copy, results staged
// in registers between two barriers.  lf / lt are the workgroup's LDS arrays (MultiShape sizes).
//
// PHYSICAL boundaries with two LDS copies: the region's boundary (SLOW) nodes are listed at load
// time and each sub-step updates the other nodes first, with no boundary code, then the listed
// ones, one per lane.  A wavefront that held one inlet or outlet lane would otherwise run the whole
// boundary path for all 64 lanes, and with one tile per compute unit the tiles at the inlet and
// outlet (every wavefront of theirs holds such a lane) set the launch time.
template <int ST, int BC, int TX, int K, int TY, int NT, bool DB>
constexpr bool multi_uses_slow_list() { return DB && BC == BC_PHYS; }

template <int ST, int BC, int TX, int K, int TY, int NT, bool DB, class P>
__device__ __forceinline__ void multi_tile(const P &p, int gxm, int tile, typename Codec<ST, 9>::C *lf,
                                           uint8_t *lt, uint32_t *lxy, int *slow_list, int *nslow) {
  using S = MultiShape<ST, TX, K, TY, NT, DB>;
  constexpr int Q = S::Q, RX = S::RX, RY = S::RY, RN = S::RN, NPOS = S::NPOS;
  constexpr bool LIST = multi_uses_slow_list<ST, BC, TX, K, TY, NT, DB>();
  using CD = Codec<ST, Q>;
  using T = typename CD::T;
  using C = typename CD::C;
  constexpr bool SH = CD::shifted;  // populations as g = f - w_i, stored with stochastic rounding
  // step t0 + s - 1 is sub-step s (the single-step kernels' key for that step)
  const uint32_t t0 = SH ? *p.step_in : 0u;
  if constexpr (SH)
    if (tile == 0 && threadIdx.x == 0) *p.step_out = t0 + K;

  __syncthreads();  // a workgroup that walks several tiles: the previous tile's reads are done
  if constexpr (LIST) {
    if (threadIdx.x == 0) *nslow = 0;
    __syncthreads();
  }
  const int bx = tile % gxm, by = tile / gxm;
  const int x0 = bx * S::TX - K, y0 = by * TY - K;  // lattice coordinates of region position (0, 0)
  const T *fin = (const T *)p.fin;

  // load: every region position, its node type and its lattice coordinates (periodic, packed
  // x | y << 16; the lattice sides are below 65536)
#pragma unroll
  for (int k = 0; k < NPOS; ++k) {
    const int pos = (int)threadIdx.x + k * NT;
    if (pos < RN) {
      int x = (x0 + pos % RX) % p.nx, y = (y0 + pos / RX) % p.ny;
      x = x < 0 ? x + p.nx : x;
      y = y < 0 ? y + p.ny : y;
      const uint32_t n = (uint32_t)x + (uint32_t)p.nx * (uint32_t)y;
      static_for<Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        lf[q * RN + pos] = CD::template ld<q>(fin + (size_t)q * (size_t)p.stride, n);
      });
      uint8_t t = p.type[n];
      if constexpr (BC == BC_PHYS) {  // held across a wrap that carries nothing (never updated)
        const int ux = x0 + pos % RX, uy = y0 + pos / RX;  // unwrapped
        if (((p.wrap_cut & 1) && (ux < 0 || ux >= p.nx)) || ((p.wrap_cut & 2) && (uy < 0 || uy >= p.ny)))
          t = NODE_SOLID;
      }
      lt[pos] = t;
      lxy[pos] = (uint32_t)x | ((uint32_t)y << 16);
      if constexpr (LIST)
        if (t & NODE_SLOW) slow_list[atomicAdd(nslow, 1)] = pos;
    }
  }
  __syncthreads();
  const int ns = LIST ? *nslow : 0;
  if constexpr (LIST) {
    // Each thread annotates the list entries it updates in every sub-step (so no barrier): bits
    // 0-10 the position, 11-18 solid pull sources i = 1..8 (bit 10 + i), 19-26 moving-wall ones
    // (bit 18 + i), 27 the west neighbour is solid, 28 inlet, 29 outlet, 30 wall neighbour, 31 the
    // east neighbour is solid.
    static_assert(RN <= 2048, "list entries hold the position in 11 bits");
    for (int e = (int)threadIdx.x; e < ns; e += NT) {
      const int pos = slow_list[e];
      const uint8_t t = lt[pos];
      const int i = pos % RX, j = pos / RX;
      uint32_t w = (uint32_t)pos | ((t & NODE_INLET) ? 1u << 28 : 0u) | ((t & NODE_OUTLET) ? 1u << 29 : 0u) |
                   ((t & NODE_WALLNB) ? 1u << 30 : 0u);
      if (i >= 1 && i < RX - 1 && j >= 1 && j < RY - 1) {  // edge positions are never updated
        static_for<Q - 1>([&](auto I) {
          constexpr int q = decltype(I)::value + 1;
          const uint8_t ts = lt[pos - Lat<Q>::cx[q] - Lat<Q>::cy[q] * RX];
          if (ts & NODE_SOLID) w |= (1u << (10 + q)) | ((ts & NODE_MOVING) ? 1u << (18 + q) : 0u);
        });
        if (lt[pos - 1] & NODE_SOLID) w |= 1u << 27;
        if (lt[pos + 1] & NODE_SOLID) w |= 1u << 31;
      }
      slow_list[e] = (int)w;
    }
  }

  // One node update of region position pos from the LDS copy src into out; false when the node
  // is not updated (PHYSICAL walls keep their populations).
  auto update = [&](const C *src, int pos, C *out, auto with_fixups) -> bool {
    const uint8_t t = lt[pos];
    C f[Q], rho, ux, uy, uz, drho = 0;
    static_for<Q>([&](auto I) {
      constexpr int q = decltype(I)::value;
      f[q] = src[q * RN + pos - Lat<Q>::cx[q] - Lat<Q>::cy[q] * RX];
    });
    if constexpr (BC == BC_REF) {
      const int yin = (t & NODE_INLET) ? (int)(lxy[pos] >> 16) : 0;
      if constexpr (SH) {
        moments_s<Q>(f, drho, rho, ux, uy, uz, CD::S);
        overrides_s<Q>(p, t, yin, 0, drho, rho, ux, uy, uz, CD::S);
      } else {
        moments<Q>(f, rho, ux, uy, uz);
        overrides<Q>(p, t, yin, 0, rho, ux, uy, uz);
      }
      collide<Q, CO_BGK, SH>(p, f, rho, ux, uy, uz, out, drho, (C)CD::S);
      if (t & NODE_SOLID) {  // full-way bounce-back: out[opp(j)] = f[j], j >= 1
        static_for<Q - 1>([&](auto J) {
          constexpr int q = decltype(J)::value + 1;
          out[Lat<Q>::opp[q]] = f[q];
        });
      }
      return true;
    } else {
      if (t & NODE_SOLID) return false;
      if (decltype(with_fixups)::value && (t & NODE_SLOW)) {
        const uint32_t xy = lxy[pos];
        physical_fixups<Q, SH>(
            p, (int)(xy & 0xffffu), (int)(xy >> 16), 0, t, 0u, f, (C)CD::S,
            [&](auto I) -> uint8_t {
              constexpr int q = decltype(I)::value;
              return lt[pos - Lat<Q>::cx[q] - Lat<Q>::cy[q] * RX];
            },
            [&](auto J) -> C { return src[decltype(J)::value * RN + pos]; },
            [&](auto J) -> C { return src[decltype(J)::value * RN + pos - 1]; },
            [&]() -> uint8_t { return lt[pos - 1]; },
            [&](auto J) -> C { return src[decltype(J)::value * RN + pos + 1]; },
            [&]() -> uint8_t { return lt[pos + 1]; });
      }
      if constexpr (SH) moments_s<Q>(f, drho, rho, ux, uy, uz, CD::S);
      else moments<Q>(f, rho, ux, uy, uz);
      collide<Q, CO_BGK, SH>(p, f, rho, ux, uy, uz, out, drho, (C)CD::S);
      return true;
    }
  };
  // the rounding a store and a reload apply; rnd: the node's rounding bits in this sub-step
  auto round_trip = [&](auto I, C v, uint32_t rnd) -> C {
    constexpr int q = decltype(I)::value;
    T w;
    CD::template st<q>(&w, 0, v, pop_bits<q>(rnd));
    return CD::template ld<q>(&w, 0);
  };
  // node_bits of region position pos in the sub-step with key k
  auto bits_of = [&](int pos, uint32_t k) -> uint32_t {
    if constexpr (!SH) return 0u;
    const uint32_t xy = lxy[pos];
    return node_bits((xy & 0xffffu) + (uint32_t)p.nx * (xy >> 16), k);
  };
  auto keep = [&](C *dst, int pos, const C *out, uint32_t k) {
    const uint32_t rnd = bits_of(pos, k);
    static_for<Q>([&](auto I) { dst[decltype(I)::value * RN + pos] = round_trip(I, out[decltype(I)::value], rnd); });
  };

  T *fout = (T *)p.fout;
  // rows dealt top-down when the held rows are at the region's end, so that they come first
  const bool flip = BC == BC_PHYS && (p.wrap_cut & 2) && y0 + RY > p.ny;
  // sub-step K updates exactly the interior: write the part inside the lattice
  auto write_out = [&](int i, int j, const C *out, uint32_t k) {
    const int x = x0 + i, y = y0 + j;
    if (x < p.nx && y < p.ny) {
      const uint32_t n = (uint32_t)x + (uint32_t)p.nx * (uint32_t)y;
      const uint32_t rnd = SH ? node_bits(n, k) : 0u;
      static_for<Q>([&](auto I) {
        constexpr int q = decltype(I)::value;
        CD::template st<q>(fout + (size_t)q * (size_t)p.stride, n, out[q], pop_bits<q>(rnd));
      });
    }
  };
  // Sub-step s updates the window [s, RX - s) x [s, RY - s), its positions dealt to consecutive
  // threads (row-major), so a sub-step runs ceil(window / 64) wavefronts and the last one writes
  // whole tile rows.
  static_for<K>([&](auto SI) {
    constexpr int s = decltype(SI)::value + 1;
    constexpr int WX = RX - 2 * s, W = WX * (RY - 2 * s), NW = (W + NT - 1) / NT;
    const uint32_t key = SH ? step_key(t0 + (uint32_t)(s - 1)) : 0u;
    const C *src = lf + (DB ? ((s - 1) & 1) * Q * RN : 0);
    C *dst = lf + (DB ? (s & 1) * Q * RN : 0);
    if constexpr (LIST) {
      // boundary nodes first, dealt from thread 0 up: their long dependent chains then run on the
      // oldest wavefronts while the others update the plain nodes (positions of one sub-step are
      // independent)
      for (int e = (int)threadIdx.x; e < ns; e += NT) {
        const uint32_t w = (uint32_t)slow_list[e];
        const int pos = (int)(w & 0x7ffu);
        const int i = pos % RX, j = pos / RX;
        if (i < s || i >= RX - s || j < s || j >= RY - s) continue;
        // SLOW nodes are never solid; every LDS read below is issued before the first use
        C f[Q], own[Q], out[Q], rho, ux, uy, uz;
        static_for<Q>([&](auto I) {
          constexpr int q = decltype(I)::value;
          f[q] = src[q * RN + pos - Lat<Q>::cx[q] - Lat<Q>::cy[q] * RX];
          own[q] = src[q * RN + pos];
        });
        const uint32_t xy = lxy[pos];
        const uint8_t t = (uint8_t)(((w >> 28) & 1u ? NODE_INLET : 0) | ((w >> 29) & 1u ? NODE_OUTLET : 0) |
                                    ((w >> 30) & 1u ? NODE_WALLNB : 0));
        physical_fixups<Q, SH>(
            p, (int)(xy & 0xffffu), (int)(xy >> 16), 0, t, 0u, f, (C)CD::S,
            [&](auto I) -> uint8_t {
              constexpr int q = decltype(I)::value;
              return (uint8_t)(((w >> (10 + q)) & 1u ? NODE_SOLID : 0) | ((w >> (18 + q)) & 1u ? NODE_MOVING : 0));
            },
            [&](auto J) -> C { return own[decltype(J)::value]; },
            [&](auto J) -> C { return src[decltype(J)::value * RN + pos - 1]; },
            [&]() -> uint8_t { return (w >> 27) & 1u ? NODE_SOLID : 0; },
            [&](auto J) -> C { return src[decltype(J)::value * RN + pos + 1]; },
            [&]() -> uint8_t { return (w >> 31) & 1u ? NODE_SOLID : 0; });
        C drho = 0;
        if constexpr (SH) moments_s<Q>(f, drho, rho, ux, uy, uz, CD::S);
        else moments<Q>(f, rho, ux, uy, uz);
        collide<Q, CO_BGK, SH>(p, f, rho, ux, uy, uz, out, drho, (C)CD::S);
        if constexpr (s == K) write_out(i, j, out, key);
        else keep(dst, pos, out, key);
      }
    }
    C res[DB ? 1 : NW][Q];
    int wpos[NW];
    bool upd[NW];
    static_for<NW>([&](auto KI) {
      constexpr int k = decltype(KI)::value;
      // a partial last pass is dealt from the last thread down, so the first wavefronts (which
      // carry the boundary list and, with flip, the held rows) get no position in it
      constexpr bool REV = LIST && k == NW - 1 && W % NT != 0;
      const unsigned e = (REV ? (unsigned)(NT - 1) - threadIdx.x : threadIdx.x) + (unsigned)(k * NT);
      const int i = s + (int)(e % (unsigned)WX);
      const int j = flip ? RY - 1 - s - (int)(e / (unsigned)WX) : s + (int)(e / (unsigned)WX);
      const int pos = i + j * RX;
      wpos[k] = pos;
      upd[k] = (W % NT == 0 || k < NW - 1 || e < (unsigned)W);
      if (LIST && upd[k] && (lt[pos] & NODE_SLOW)) upd[k] = false;  // done in the pass above
      if (!upd[k]) return;
      C *out = res[DB ? 0 : k];
      upd[k] = update(src, pos, out, std::integral_constant<bool, !LIST>{});
      if (!upd[k]) return;
      if constexpr (s == K) write_out(i, j, out, key);
      else if constexpr (DB) keep(dst, pos, out, key);
    });
    if constexpr (s < K) {
      __syncthreads();  // DB: the copy written is complete; otherwise: every read of sub-step s is done
      if constexpr (!DB) {
        static_for<NW>([&](auto KI) {
          constexpr int k = decltype(KI)::value;
          if (upd[k]) keep(dst, wpos[k], res[k], key);
        });
        __syncthreads();
      }
    }
  });
}

template <int ST, int BC, int TX, int K, int TY, int NT, bool DB>
__global__ void __launch_bounds__(NT) k_step_multi(StepParams p, int gxm, int ntm) {
  using S = MultiShape<ST, TX, K, TY, NT, DB>;
  constexpr bool LIST = multi_uses_slow_list<ST, BC, TX, K, TY, NT, DB>();
  __shared__ typename Codec<ST, 9>::C lf[S::NBUF * S::Q * S::RN];
  __shared__ uint8_t lt[S::RN];
  __shared__ uint32_t lxy[S::RN];
  __shared__ int slow_list[LIST ? S::RN : 1];
  __shared__ int nslow;
  multi_tile<ST, BC, TX, K, TY, NT, DB>(p, gxm, xcd_tile((int)blockIdx.x, ntm), lf, lt, lxy, slow_list, &nslow);
}

// Batched: every member of one D2Q9 BGK group, K steps, in the default launch shape.  Member m
// owns multi-step tiles [prefix_multi[m], prefix_multi[m+1]).
constexpr int BATCH_TX = 64, BATCH_K = 5, BATCH_TY = 16, BATCH_NT = 1024;
constexpr bool BATCH_DB = true;
template <int ST, int BC>
__global__ void __launch_bounds__(BATCH_NT) k_step_multi_batch(BatchArgs b) {
  using S = MultiShape<ST, BATCH_TX, BATCH_K, BATCH_TY, BATCH_NT, BATCH_DB>;
  constexpr bool LIST = multi_uses_slow_list<ST, BC, BATCH_TX, BATCH_K, BATCH_TY, BATCH_NT, BATCH_DB>();
  __shared__ typename Codec<ST, 9>::C lf[S::NBUF * S::Q * S::RN];
  __shared__ uint8_t lt[S::RN];
  __shared__ uint32_t lxy[S::RN];
  __shared__ int slow_list[LIST ? S::RN : 1];
  __shared__ int nslow;
  for_batch_tiles<true>(b, [&](const auto &p, int tile) {
    multi_tile<ST, BC, BATCH_TX, BATCH_K, BATCH_TY, BATCH_NT, BATCH_DB>(p, (p.nx + BATCH_TX - 1) / BATCH_TX, tile,
                                                                       lf, lt, lxy, slow_list, &nslow);
  });
}

namespace {

// Launch shapes: interior TX x TY, steps per launch K, threads NT, double-buffered LDS.
struct MultiVariant {
  int TX, K, TY, NT;
  bool DB;
};
#define LBM_MULTI_SHAPES(X)                                                                         \
  X(0, 64, 4, 16, 1024, true) X(1, 64, 3, 16, 1024, true) X(2, 32, 4, 32, 1024, true)              \
  X(3, 64, 4, 16, 1024, false) X(4, 64, 4, 8, 1024, false) X(5, 64, 4, 8, 512, false)              \
  X(6, 32, 4, 16, 1024, false) X(7, 32, 4, 16, 512, false) X(8, 64, 2, 8, 1024, true)               \
  X(9, 64, 5, 16, 1024, true)
#define LBM_MULTI_ENTRY(i, tx, k, ty, nt, db) {tx, k, ty, nt, db},
constexpr MultiVariant kMultiVariants[] = {LBM_MULTI_SHAPES(LBM_MULTI_ENTRY)};

template <int ST, int BC, int TX, int K, int TY, int NT, bool DB>
hipError_t launch_multi_shape(const StepParams &p, hipStream_t s) {
  if constexpr (ST == ST_F64 && DB) {
    return hipErrorInvalidValue;  // two FP64 copies of a region exceed the 160 KB of LDS
  } else {
    const int gxm = (p.nx + TX - 1) / TX, gym = (p.ny + TY - 1) / TY;
    hipLaunchKernelGGL((k_step_multi<ST, BC, TX, K, TY, NT, DB>), dim3(gxm * gym), dim3(NT), 0, s, p, gxm,
                       gxm * gym);
    return hipGetLastError();
  }
}

template <int ST, int BC>
hipError_t launch_multi_variant(int v, const StepParams &p, hipStream_t s) {
#define LBM_MULTI_CASE(i, tx, k, ty, nt, db) \
  case i: return launch_multi_shape<ST, BC, tx, k, ty, nt, db>(p, s);
  switch (v) { LBM_MULTI_SHAPES(LBM_MULTI_CASE) }
#undef LBM_MULTI_CASE
  return hipErrorInvalidValue;
}

}  // namespace

static_assert(kMultiVariants[MULTI_DEFAULT].TX == BATCH_TX && kMultiVariants[MULTI_DEFAULT].K == BATCH_K &&
                  kMultiVariants[MULTI_DEFAULT].TY == BATCH_TY && kMultiVariants[MULTI_DEFAULT].NT == BATCH_NT &&
                  kMultiVariants[MULTI_DEFAULT].DB == BATCH_DB,
              "batched multi-step launches run the default shape");

int multi_variant_count() { return (int)(sizeof(kMultiVariants) / sizeof(kMultiVariants[0])); }

int multi_tiles(int variant, int nx, int ny) {
  if (variant < 0 || variant >= multi_variant_count()) return 0;
  const MultiVariant &v = kMultiVariants[variant];
  return ((nx + v.TX - 1) / v.TX) * ((ny + v.TY - 1) / v.TY);
}

hipError_t launch_step_multi_batch(const Config &c, const BatchArgs &b, int grid, hipStream_t s) {
  if (multi_steps_per_launch(c, MULTI_DEFAULT) != BATCH_K) return hipErrorInvalidValue;
  auto byBC = [&](auto STc) -> hipError_t {
    constexpr int ST = decltype(STc)::value;
    if (c.bc == BC_REF)
      hipLaunchKernelGGL((k_step_multi_batch<ST, BC_REF>), dim3(grid), dim3(BATCH_NT), 0, s, b);
    else
      hipLaunchKernelGGL((k_step_multi_batch<ST, BC_PHYS>), dim3(grid), dim3(BATCH_NT), 0, s, b);
    return hipGetLastError();
  };
  switch (c.storage) {
    case ST_F32: return byBC(IC<ST_F32>{});
    case ST_F16S: return byBC(IC<ST_F16S>{});
    case ST_BF16S: return byBC(IC<ST_BF16S>{});
  }
  return hipErrorInvalidValue;
}

int multi_steps_per_launch(const Config &c, int variant) {
  if (c.Q != 9 || c.collision != CO_BGK) return 0;  // and nx, ny < 65536
  if (variant < 0 || variant >= multi_variant_count()) return 0;
  if (c.storage == ST_F64 && kMultiVariants[variant].DB) return 0;  // FP64: one-copy shapes only
  return kMultiVariants[variant].K;
}

hipError_t launch_step_multi(const Config &c, int variant, const StepParams &p, hipStream_t s) {
  if (multi_steps_per_launch(c, variant) == 0 || p.nz != 1) return hipErrorInvalidValue;
  auto byBC = [&](auto STc) -> hipError_t {
    constexpr int ST = decltype(STc)::value;
    return c.bc == BC_REF ? launch_multi_variant<ST, BC_REF>(variant, p, s)
                          : launch_multi_variant<ST, BC_PHYS>(variant, p, s);
  };
  switch (c.storage) {
    case ST_F64: return byBC(IC<ST_F64>{});
    case ST_F32: return byBC(IC<ST_F32>{});
    case ST_F16S: return byBC(IC<ST_F16S>{});
    case ST_BF16S: return byBC(IC<ST_BF16S>{});
  }
  return hipErrorInvalidValue;
}

}  // namespace lbm
