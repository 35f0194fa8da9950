// lbm_step.hpp — the node-per-thread fused step (tile body and kernel) and the host dispatch
// over (lattice, storage, boundary mode, collision), shared by the translation units that
// instantiate the step kernel with ordinary and with non-temporal population stores.
#pragma once

#include "lbm_device.hpp"

namespace lbm {

// One fused time step: pull stream + boundary handling + collision (+ full-way bounce-back at
// solid nodes in REFERENCE mode), A -> B buffers.  Replaces cpp/d2q9_bgk_lid.cpp:124-155.
// Shifted storages compute on g = f - w_i throughout and store with stochastic rounding.
template <int Q, int ST, int BC, int CO, bool NT, class P>
__device__ __forceinline__ void step_tile(const P &p, int tile) {
  using CD = Codec<ST, Q>;
  using C = typename CD::C;
  constexpr bool SH = CD::shifted;
  const uint32_t key = step_begin<SH>(p, tile, 1);
  const NodePos q = node_of_tile(p, tile);
  if (!q.valid) return;
  const uint8_t t = p.type[q.n];
  C f[Q], out[Q], rho, ux, uy, uz, drho = 0;
  // D3Q27 MRT with FP32 arithmetic: rho and u from the monomial moments of the collision's forward
  // transform (80 fewer additions a node than the separate sums)
  constexpr bool MONO = Q == 27 && CO != CO_BGK && sizeof(C) == 4;
  float mono[MONO ? 27 : 1];
  if constexpr (BC == BC_REF) {
    pull_periodic<Q, ST>(p, q, f);
    if constexpr (MONO) {
      mrt27_monomials(f, mono);
      moments_mono<SH>(mono, drho, rho, ux, uy, uz, CD::S);
      if constexpr (SH) overrides_s<Q>(p, t, q.y, q.z, drho, rho, ux, uy, uz, CD::S);
      else overrides<Q>(p, t, q.y, q.z, rho, ux, uy, uz);
      collide_mrt_fp32<CO, SH>(p, f, SH ? rho * CD::S : rho, ux, uy, uz, out, drho, CD::S, mono);
    } else if constexpr (SH) {
      moments_s<Q>(f, drho, rho, ux, uy, uz, CD::S);
      overrides_s<Q>(p, t, q.y, q.z, drho, rho, ux, uy, uz, CD::S);
    } else {
      moments<Q>(f, rho, ux, uy, uz);
      overrides<Q>(p, t, q.y, q.z, rho, ux, uy, uz);
    }
    if constexpr (!MONO) collide<Q, CO, SH>(p, f, rho, ux, uy, uz, out, drho, (C)CD::S);
    if (t & NODE_SOLID) {  // F(REFLECTED) = BOUNCEDBACK: out[opp(j)] = f[j], j >= 1
      static_for<Q - 1>([&](auto J) {
        constexpr int j = decltype(J)::value + 1;
        out[Lat<Q>::opp[j]] = f[j];
      });
    }
  } else {
    pull_physical<Q, ST>(p, q, t, f);
    // the wavefront holds a fluid node (a wavefront is an aligned block of 64 node indices); after
    // the pull, which is issued before the node type is known
    const bool any_fluid = __builtin_amdgcn_ballot_w64(!(t & NODE_SOLID)) != 0;
    if (t & NODE_SOLID) {
      // 3D: the rest equilibrium, stored with the fluid lanes (rest_at_solids); a wavefront of
      // solid nodes only leaves its lines untouched.  2D: not written.
      if constexpr (!rest_at_solids<Q>()) return;
      if (!any_fluid) return;
      rest_populations<Q, ST>(out);
    } else if constexpr (MONO) {
      mrt27_monomials(f, mono);
      moments_mono<SH>(mono, drho, rho, ux, uy, uz, CD::S);
      collide_mrt_fp32<CO, SH>(p, f, SH ? rho * CD::S : rho, ux, uy, uz, out, drho, CD::S, mono);
    } else {
      if constexpr (SH) moments_s<Q>(f, drho, rho, ux, uy, uz, CD::S);
      else moments<Q>(f, rho, ux, uy, uz);
      collide<Q, CO, SH>(p, f, rho, ux, uy, uz, out, drho, (C)CD::S);
    }
  }
  store_all<Q, ST, NT>(p, q.n, out, SH ? node_bits(q.n, key) : 0u);
}

// rho, u the reference holds after a step at node q (cpp/d2q9_bgk_lid.cpp:131-146): moments of the
// pulled populations with the overrides (REFERENCE), or of the boundary-completed pulled
// populations (PHYSICAL; 0 at solid nodes).  out = {rho, ux, uy, uz}; the shifted storages report
// rho as 1 + drho in FP64.
template <int Q, int ST, int BC, class P>
__device__ __forceinline__ void fields_of_node(const P &p, const NodePos &q, double *out) {
  using CD = Codec<ST, Q>;
  using C = typename CD::C;
  const uint8_t t = p.type[q.n];
  C f[Q], rho = 0, ux = 0, uy = 0, uz = 0;
  if constexpr (CD::shifted) {
    C drho = -CD::S;
    if constexpr (BC == BC_REF) {
      pull_periodic<Q, ST>(p, q, f);
      moments_s<Q>(f, drho, rho, ux, uy, uz, CD::S);
      overrides_s<Q>(p, t, q.y, q.z, drho, rho, ux, uy, uz, CD::S);
    } else if (!(t & NODE_SOLID)) {
      pull_physical<Q, ST>(p, q, t, f);
      moments_s<Q>(f, drho, rho, ux, uy, uz, CD::S);
    }
    out[0] = 1.0 + (double)drho * (1.0 / CD::S);
  } else {
    if constexpr (BC == BC_REF) {
      pull_periodic<Q, ST>(p, q, f);
      moments<Q>(f, rho, ux, uy, uz);
      overrides<Q>(p, t, q.y, q.z, rho, ux, uy, uz);
    } else if (!(t & NODE_SOLID)) {
      pull_physical<Q, ST>(p, q, t, f);
      moments<Q>(f, rho, ux, uy, uz);
    }
    out[0] = (double)rho;
  }
  out[1] = (double)ux;
  out[2] = (double)uy;
  out[3] = (double)uz;
}

// NT: populations written with non-temporal stores (lattices too large for the caches), whose
// tiles run in the skewed natural order (stream_tile).
template <int Q, int ST, int BC, int CO, bool NT>
__global__ void __launch_bounds__(BLOCK_THREADS) k_step(StepParams p) {
  step_tile<Q, ST, BC, CO, NT>(p, NT ? stream_tile((int)blockIdx.x, p.gx) : xcd_tile((int)blockIdx.x, p.ntiles));
}

// Persistent batched step: every member of one configuration group, one launch.  NT: non-temporal
// population stores (groups whose buffers cannot stay in the caches, which take their tiles in
// stream order).
template <int Q, int ST, int BC, int CO, bool ONE, bool NT = false>
__global__ void __launch_bounds__(BLOCK_THREADS) k_step_batch(BatchArgs b) {
  for_batch_tiles<false, ONE>(b, [&](const auto &p, int tile) { step_tile<Q, ST, BC, CO, NT>(p, tile); });
}

// ---- dispatch ----------------------------------------------------------------------------
template <class F>
inline hipError_t with_storage(const Config &c, F &&f) {
  auto byQ = [&](auto Qc) -> hipError_t {
    switch (c.storage) {
      case ST_F64: return f(Qc, IC<ST_F64>{});
      case ST_F32: return f(Qc, IC<ST_F32>{});
      case ST_F16S: return f(Qc, IC<ST_F16S>{});
      case ST_BF16S: return f(Qc, IC<ST_BF16S>{});
    }
    return hipErrorInvalidValue;
  };
  if (c.Q == 9) return byQ(IC<9>{});
  if (c.Q == 19) return byQ(IC<19>{});
  if (c.Q == 27) return byQ(IC<27>{});
  return hipErrorInvalidValue;
}

template <class F>
inline hipError_t with_full(const Config &c, F &&f) {
  return with_storage(c, [&](auto Qc, auto STc) -> hipError_t {
    constexpr int Q = decltype(Qc)::value;
    auto byBC = [&](auto BCc) -> hipError_t {
      if (c.collision == CO_BGK) return f(Qc, STc, BCc, IC<CO_BGK>{});
      if constexpr (Q == 27) {
        if (c.collision == CO_MRT_RAW) return f(Qc, STc, BCc, IC<CO_MRT_RAW>{});
        if (c.collision == CO_MRT_GS) return f(Qc, STc, BCc, IC<CO_MRT_GS>{});
      }
      if constexpr (Q == 19) {  // the orthogonal basis only
        if (c.collision == CO_MRT_GS) return f(Qc, STc, BCc, IC<CO_MRT_GS>{});
      }
      return hipErrorInvalidValue;
    };
    return c.bc == BC_REF ? byBC(IC<BC_REF>{}) : byBC(IC<BC_PHYS>{});
  });
}

}  // namespace lbm
