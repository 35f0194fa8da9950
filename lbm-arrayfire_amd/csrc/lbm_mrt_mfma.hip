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
#include "lbm_device.hpp"

namespace lbm {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MT_STRIDE = 44;  // floats per node row: f[0..26], 0, phi[0..12] at 28..40, pad

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int ST, int BC>
__device__ __forceinline__ void mrt_mfma_tile(const StepParams &p, int tile, float (*lds)[64 * MT_STRIDE]) {
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
    const float row[MT_STRIDE] = {
        f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], f[9], f[10], f[11], f[12], f[13],
        f[14], f[15], f[16], f[17], f[18], f[19], f[20], f[21], f[22], f[23], f[24], f[25], f[26], 0.f,
        rho, rho * ux, rho * uy, rho * uz, rho * ux * ux, rho * uy * uy, rho * uz * uz,
        rho * ux * uy, rho * ux * uz, rho * uy * uz, ux * ux, uy * uy, uz * uz, 0.f, 0.f, 0.f};
    f32x4 *dst = (f32x4 *)(L + lane * MT_STRIDE);
#pragma unroll
    for (int k = 0; k < MT_STRIDE / 4; ++k)
      dst[k] = f32x4{row[4 * k], row[4 * k + 1], row[4 * k + 2], row[4 * k + 3]};
  }
  // lane constants: AF[2][8], AP[2][4], SC[2][4], AI[2][8]
  f32x4 tab[12];
  {
    const f32x4 *src = (const f32x4 *)(p.mrt_tab + lane * MFMA_TAB);
#pragma unroll
    for (int k = 0; k < 12; ++k) tab[k] = src[k];
  }
  __syncthreads();

  // ---- phase 2: 4 tiles of 16 nodes on MFMA ----------------------------------------------
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

template <int ST, int BC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_step_mrt_mfma(StepParams p) {
  __shared__ __attribute__((aligned(16))) float lds[BLOCK_THREADS / 64][64 * MT_STRIDE];
  mrt_mfma_tile<ST, BC>(p, xcd_tile((int)blockIdx.x, p.ntiles), lds);
}

template <int ST, int BC>
__global__ void __launch_bounds__(BLOCK_THREADS) k_step_batch_mrt_mfma(BatchArgs b) {
  __shared__ __attribute__((aligned(16))) float lds[BLOCK_THREADS / 64][64 * MT_STRIDE];
  for_batch_tiles(b, [&](const StepParams &p, int tile) { mrt_mfma_tile<ST, BC>(p, tile, lds); });
}

hipError_t launch_step_batch_mrt_mfma(const Config &c, const BatchArgs &b, int grid, hipStream_t s) {
  auto go = [&](auto STc) -> hipError_t {
    constexpr int ST = decltype(STc)::value;
    if (c.bc == BC_REF)
      hipLaunchKernelGGL((k_step_batch_mrt_mfma<ST, BC_REF>), dim3(grid), dim3(BLOCK_THREADS), 0, s, b);
    else
      hipLaunchKernelGGL((k_step_batch_mrt_mfma<ST, BC_PHYS>), dim3(grid), dim3(BLOCK_THREADS), 0, s, b);
    return hipGetLastError();
  };
  switch (c.storage) {
    case ST_F32: return go(IC<ST_F32>{});
    case ST_F16S: return go(IC<ST_F16S>{});
    case ST_BF16S: return go(IC<ST_BF16S>{});
  }
  return hipErrorInvalidValue;
}

hipError_t launch_step_mrt_mfma(const Config &c, const StepParams &p, hipStream_t s) {
  auto go = [&](auto STc) -> hipError_t {
    constexpr int ST = decltype(STc)::value;
    if (c.bc == BC_REF)
      hipLaunchKernelGGL((k_step_mrt_mfma<ST, BC_REF>), dim3(p.ntiles), dim3(BLOCK_THREADS), 0, s, p);
    else
      hipLaunchKernelGGL((k_step_mrt_mfma<ST, BC_PHYS>), dim3(p.ntiles), dim3(BLOCK_THREADS), 0, s, p);
    return hipGetLastError();
  };
  switch (c.storage) {
    case ST_F32: return go(IC<ST_F32>{});
    case ST_F16S: return go(IC<ST_F16S>{});
    case ST_BF16S: return go(IC<ST_BF16S>{});
  }
  return hipErrorInvalidValue;
}

}  // namespace lbm
