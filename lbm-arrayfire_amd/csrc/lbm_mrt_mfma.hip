// lbm_mrt_mfma.hip — D3Q27 MRT step with the moment transforms on MFMA (gfx950).
//
// The reference evaluates the MRT collision as three N x 27 x 27 GEMMs through ArrayFire's BLAS
// (rust/examples/d3q27_mrt_lid.rs:336 m = f M^T, :385 (m - meq) S^T, :387 f = m* M^-T).  Here one
// wavefront owns 64 nodes:
//   phase 1  node per lane: pull + boundary rules + moments + overrides (same code as the scalar
//            kernel), then the 27 pulled populations and the 13 equilibrium monomials
//            phi = (rho, rho u_a, rho u_a^2, rho u_a u_b, u_a^2) are staged in LDS, one row of
//            MT_STRIDE floats per node;
//   phase 2  per 16-node tile, on v_mfma_f32_16x16x4_f32 (rows padded 27 -> 32):
//              m  = M f                    2 row tiles x 8 K-steps   (16 MFMA)
//              m *= (1 - s)                per-row lane constants
//              m += (s . E) phi            2 x 4 K-steps             (8 MFMA)  meq = E phi
//              f' = M^-1 m                 2 x 8 K-steps             (16 MFMA)
//            The K order of the inverse is permuted so that the accumulator of the forward
//            transform (lane l holds rows 4(l/16)+r of node l%16) is directly the B operand of
//            the inverse: no data movement between the two transforms;
//   phase 3  f' goes back through LDS to node-per-lane order and is stored like the scalar path
//            (full-way bounce-back at solid nodes in REFERENCE mode).
// Per-lane A operands (M, s.E, M^-1 entries and the 1 - s scales, 48 floats) come from a table
// the host builds for the lattice's basis and rates (build_mfma_table).
//
// The fused engine folds the three stages into one operator on the host,
//   f' = M^-1 (I - S) M f + M^-1 S E phi = [C | D] [f; phi],
// and runs it as a single K = 48 contraction per tile (24 MFMA instead of 40).
#include "lbm_device.hpp"

namespace lbm {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// LDS row of one node.  Two-stage engine: f[0..26], 0, phi[0..12] at 28..40, pad to 44.
// Fused engine: f[0..26], phi[0..12] at 27..39, zeros to 52 (K = 48 slots, 12 per lane group).
template <bool FUSED> constexpr int mt_stride() { return FUSED ? 52 : 44; }
template <bool FUSED> constexpr int phi_at() { return FUSED ? 27 : 28; }
constexpr int MT_STRIDE_MAX = 52;

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int ST, int BC, bool FUSED, class Params>
__device__ __forceinline__ void mrt_mfma_tile(const Params &p, int tile, float (*lds)[64 * MT_STRIDE_MAX]) {
  constexpr int MT_STRIDE = mt_stride<FUSED>();
  using CD = Codec<ST, 27>;
  using C = typename CD::C;
  static_assert(sizeof(C) == 4, "MFMA path computes in FP32");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  float *L = lds[wv];

  // ---- phase 1: node per lane -----------------------------------------------------------
  const NodePos q = node_of_tile(p, tile);
  const uint8_t t = q.valid ? p.type[q.n] : (uint8_t)NODE_SOLID;
  const bool active = q.valid && !(BC == BC_PHYS && (t & NODE_SOLID));
  {
    C f[27], rho = 0, ux = 0, uy = 0, uz = 0;
    if (q.valid) {
      if constexpr (BC == BC_REF) pull_periodic<27, ST>(p, q, f);
      else pull_physical<27, ST>(p, q, t, f);
    }
    if (active) {
      moments<27>(f, rho, ux, uy, uz);
      if constexpr (BC == BC_REF) overrides<27>(p, t, q.y, q.z, rho, ux, uy, uz);
    } else {
      static_for<27>([&](auto I) { f[decltype(I)::value] = 0; });
    }
    float row[MT_STRIDE];
    static_for<27>([&](auto I) { row[decltype(I)::value] = f[decltype(I)::value]; });
    constexpr int P = phi_at<FUSED>();
    const float phi[13] = {rho, rho * ux, rho * uy, rho * uz, rho * ux * ux, rho * uy * uy, rho * uz * uz,
                           rho * ux * uy, rho * ux * uz, rho * uy * uz, ux * ux, uy * uy, uz * uz};
#pragma unroll
    for (int k = 27; k < MT_STRIDE; ++k) row[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 13; ++k) row[P + k] = phi[k];
    f32x4 *dst = (f32x4 *)(L + lane * MT_STRIDE);
#pragma unroll
    for (int k = 0; k < MT_STRIDE / 4; ++k)
      dst[k] = f32x4{row[4 * k], row[4 * k + 1], row[4 * k + 2], row[4 * k + 3]};
  }
  // lane constants: two-stage AF[2][8], AP[2][4], SC[2][4], AI[2][8]; fused AC[2][12]
  f32x4 tab[12];
  {
    const f32x4 *src = (const f32x4 *)(p.mrt_tab + lane * MFMA_TAB);
#pragma unroll
    for (int k = 0; k < 12; ++k) tab[k] = src[k];
  }
  __syncthreads();

  // ---- phase 2: 4 tiles of 16 nodes on MFMA ----------------------------------------------
  if constexpr (FUSED) {
    // f' = [M^-1 (I - S) M | M^-1 S E] [f; phi]: one K = 48 contraction, 2 row tiles x 12 steps
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float *row = L + (16 * j + c) * MT_STRIDE;
      const f32x4 b0 = *(const f32x4 *)(row + 12 * g);
      const f32x4 b1 = *(const f32x4 *)(row + 12 * g + 4);
      const f32x4 b2 = *(const f32x4 *)(row + 12 * g + 8);
      f32x4 o0 = {0, 0, 0, 0}, o1 = {0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < 12; ++s) {
        const float b = s < 4 ? b0[s & 3] : (s < 8 ? b1[s & 3] : b2[s & 3]);
        o0 = mfma4(tab[s >> 2][s & 3], b, o0);        // AC[0][s]
        o1 = mfma4(tab[3 + (s >> 2)][s & 3], b, o1);  // AC[1][s]
      }
      *(f32x4 *)(row + 4 * g) = o0;
      *(f32x4 *)(row + 16 + 4 * g) = o1;
    }
  } else {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float *row = L + (16 * j + c) * MT_STRIDE;
    const f32x4 g0 = *(const f32x4 *)(row + 8 * g);
    const f32x4 g1 = *(const f32x4 *)(row + 8 * g + 4);
    const f32x4 ph = *(const f32x4 *)(row + 28 + 4 * g);
    f32x4 a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const float b = s < 4 ? g0[s & 3] : g1[s & 3];
      a0 = mfma4(tab[s >> 2][s & 3], b, a0);          // AF[0][s]
      a1 = mfma4(tab[2 + (s >> 2)][s & 3], b, a1);    // AF[1][s]
    }
    a0 *= tab[6];  // SC[0][r] = 1 - s_row
    a1 *= tab[7];  // SC[1][r]
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      a0 = mfma4(tab[4][s], ph[s], a0);  // AP[0][s] = s_row E[row][4g+s]
      a1 = mfma4(tab[5][s], ph[s], a1);  // AP[1][s]
    }
    f32x4 o0 = {0, 0, 0, 0}, o1 = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const float b = s < 4 ? a0[s & 3] : a1[s & 3];  // m*[16(s/4) + 4g + s%4][node]
      o0 = mfma4(tab[8 + (s >> 2)][s & 3], b, o0);    // AI[0][s]
      o1 = mfma4(tab[10 + (s >> 2)][s & 3], b, o1);   // AI[1][s]
    }
    *(f32x4 *)(row + 4 * g) = o0;       // f'[4g + r]
    *(f32x4 *)(row + 16 + 4 * g) = o1;  // f'[16 + 4g + r]
  }
  }
  __syncthreads();

  // ---- phase 3: node per lane again, store ----------------------------------------------
  C out[27];
  {
    const float *src = L + lane * MT_STRIDE;
    static_for<27>([&](auto I) { out[decltype(I)::value] = src[decltype(I)::value]; });
  }
  __syncthreads();  // the LDS rows are free for the next tile of a persistent workgroup
  if (!active) return;
  if constexpr (BC == BC_REF) {
    if (t & NODE_SOLID) {  // F(REFLECTED) = BOUNCEDBACK, as the scalar kernel
      C f[27];
      pull_periodic<27, ST>(p, q, f);
      static_for<26>([&](auto J) {
        constexpr int jj = decltype(J)::value + 1;
        out[Lat<27>::opp[jj]] = f[jj];
      });
    }
  }
  store_all<27, ST>(p, q.n, out);
}

template <int ST, int BC, bool FUSED>
__global__ void __launch_bounds__(BLOCK_THREADS) k_step_mrt_mfma(StepParams p) {
  __shared__ __attribute__((aligned(16))) float lds[BLOCK_THREADS / 64][64 * MT_STRIDE_MAX];
  mrt_mfma_tile<ST, BC, FUSED>(p, xcd_tile((int)blockIdx.x, p.ntiles), lds);
}

template <int ST, int BC, bool FUSED>
__global__ void __launch_bounds__(BLOCK_THREADS) k_step_batch_mrt_mfma(BatchArgs b) {
  __shared__ __attribute__((aligned(16))) float lds[BLOCK_THREADS / 64][64 * MT_STRIDE_MAX];
  for_batch_tiles(b, [&](const auto &p, int tile) { mrt_mfma_tile<ST, BC, FUSED>(p, tile, lds); });
}

template <class F>
static hipError_t with_mfma_variant(const Config &c, F &&f) {
  auto byST = [&](auto STc) -> hipError_t {
    auto byBC = [&](auto BCc) -> hipError_t {
      return c.engine == MRT_ENGINE_FUSED ? f(STc, BCc, std::true_type{}) : f(STc, BCc, std::false_type{});
    };
    return c.bc == BC_REF ? byBC(IC<BC_REF>{}) : byBC(IC<BC_PHYS>{});
  };
  switch (c.storage) {
    case ST_F32: return byST(IC<ST_F32>{});
    case ST_F16S: return byST(IC<ST_F16S>{});
    case ST_BF16S: return byST(IC<ST_BF16S>{});
  }
  return hipErrorInvalidValue;
}

hipError_t launch_step_batch_mrt_mfma(const Config &c, const BatchArgs &b, int grid, hipStream_t s) {
  return with_mfma_variant(c, [&](auto STc, auto BCc, auto Fc) -> hipError_t {
    hipLaunchKernelGGL((k_step_batch_mrt_mfma<decltype(STc)::value, decltype(BCc)::value, decltype(Fc)::value>),
                       dim3(grid), dim3(BLOCK_THREADS), 0, s, b);
    return hipGetLastError();
  });
}

hipError_t launch_step_mrt_mfma(const Config &c, const StepParams &p, hipStream_t s) {
  return with_mfma_variant(c, [&](auto STc, auto BCc, auto Fc) -> hipError_t {
    hipLaunchKernelGGL((k_step_mrt_mfma<decltype(STc)::value, decltype(BCc)::value, decltype(Fc)::value>),
                       dim3(p.ntiles), dim3(BLOCK_THREADS), 0, s, p);
    return hipGetLastError();
  });
}

}  // namespace lbm
