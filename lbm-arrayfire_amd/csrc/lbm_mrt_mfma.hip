// lbm_mrt_mfma.hip — D3Q27 MRT step with the moment transforms on MFMA (gfx950).
//
// The reference evaluates the MRT collision as three N x 27 x 27 GEMMs through ArrayFire's BLAS
// (rust/examples/d3q27_mrt_lid.rs:336 m = f M^T, :385 (m - meq) S^T, :387 f = m* M^-T).  The
// engines here evaluate the same map in correction form,
//   f' = f - M^-1 S (M f - meq) = f - M^-1 S M g,   g = f - f_eq(rho, u),
// which holds because meq = M f_eq (projected equilibrium moments); for the reference's own table
// (raw basis, meq_mode REFERENCE) row 26 differs and its difference delta26 = (M f_eq)_26 -
// meq_ref_26 is added to that row.  The matrix units see only g, the small non-equilibrium part,
// so their accumulation errors scale with |g| rather than with |f|.  The conserved rows (rho, j)
// do not come from M g: the FP32 f_eq values do not sum to rho exactly, and relaxing that offset
// every step drifts the mass (1.4e-4 over 20,000 steps of the 128^3 lid).  Their deviations
// (rho_f - rho, j_f - rho u, with rho_f, j_f the moments of f and rho, u after the overrides:
// exactly 0 where no override applies) are computed on the vector unit and enter through the four
// spare K slots 28..31 (unit columns of the forward operator, or the fused operator's columns).  This matters: FP32 MFMA
// accumulation does not round to nearest (measured on MI355X: 57 % of 16x16x4 results equal the
// correctly rounded value, mean error -0.2 ulp), and an engine that forms f' = M^-1 m* from the
// full moments drifts by that bias every step (1.4e-4 in rho over 20,000 steps of the 128^3 lid).
//
// One wavefront owns 64 nodes:
//   phase 1  node per lane: pull + boundary rules + moments + overrides (the scalar kernel's
//            code), then g, delta26 and the conserved deviations are staged in LDS, one row of ROW
//            floats per node (slots 0..26 g, 27 delta26, 28..31 the deviations of rho, jx, jy,
//            jz); f stays in registers;
//   phase 2  per 16-node tile, the lane group gq = lane / 16 supplies the populations (K slots)
//            8 gq .. 8 gq + 7 of node lane % 16; the accumulator of the first transform (lane l
//            holds rows 4 (l / 16) + r and 16 + 4 (l / 16) + r of node l % 16) is directly the
//            B operand of the second (its K order is permuted to match), so nothing moves
//            between the transforms:
//              MFMA_F32   d = M g on v_mfma_f32_16x16x4_f32 (2 row tiles x 8 K-steps),
//                         r = (s / n) d (+ delta26), c = (n M^-1) r (16 MFMA);
//              FUSED      c = [M^-1 S' M | M^-1 s_26 e_26 | M^-1 S_c] [g; delta26; dev] (16 MFMA),
//                         S' the rates with the conserved rows' set to 0, S_c those rows';
//              BF16X3     as MFMA_F32 on v_mfma_f32_16x16x32_bf16: g and r are split into three
//                         bfloat16 terms each (hi + mid + lo carry 24 significant bits), M and
//                         n M^-1 (integers up to 24: the raw basis' 24 M^-1, the Gram-Schmidt
//                         M^T) are exact in bfloat16, products are exact in FP32;
//            n is 24 (raw basis) or the squared row norms (Gram-Schmidt, M^-1 = M^T / n);
//   phase 3  c goes back through LDS to node-per-lane order, f' = f - c is stored like the scalar
//            path (full-way bounce-back at solid nodes in REFERENCE mode).
// Per-lane A operands and rates come from a table the host builds for the lattice's basis and
// rates (build_mfma_table).
#include "lbm_device.hpp"

namespace lbm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// LDS row of one node: g[0..26], delta26, the conserved rows' deviations; 36 floats so that the 16
// nodes a b128 read of a lane group touches spread over the banks.
constexpr int ROW = 36;

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// x = hi + mid + lo, each term a bfloat16 rounded to nearest from what the previous ones leave
__device__ __forceinline__ void split3(const float *x, bf16x8 &hi, bf16x8 &mid, bf16x8 &lo) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const __bf16 h = (__bf16)x[j];
    const float r1 = x[j] - (float)h;
    const __bf16 m = (__bf16)r1;
    const float r2 = r1 - (float)m;
    hi[j] = h;
    mid[j] = m;
    lo[j] = (__bf16)r2;
  }
}

template <int ST, int BC, int ENG, class Params>
__device__ __forceinline__ void mrt_mfma_tile(const Params &p, int tile, float (*lds)[64 * ROW]) {
  using CD = Codec<ST, 27>;
  using C = typename CD::C;
  static_assert(sizeof(C) == 4, "MFMA path computes in FP32");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, c = lane & 15;
  float *L = lds[wv];

  // ---- phase 1: node per lane -----------------------------------------------------------
  constexpr bool SH = CD::shifted;  // f holds g = f - w_i (2-byte storages)
  const uint32_t key = step_begin<SH>(p, tile, 1);
  const NodePos q = node_of_tile(p, tile);
  const uint8_t t = q.valid ? p.type[q.n] : (uint8_t)NODE_SOLID;
  const bool active = q.valid && !(BC == BC_PHYS && (t & NODE_SOLID));
  C f[27];
  bool any_fluid;
  {
    C rho = 0, ux = 0, uy = 0, uz = 0;
    if (q.valid) {
      if constexpr (BC == BC_REF) pull_periodic<27, ST>(p, q, f);
      else pull_physical<27, ST>(p, q, t, f);
    }
    any_fluid = __builtin_amdgcn_ballot_w64(!(t & NODE_SOLID)) != 0;  // in this wavefront
    float row[ROW];
    float dev[4] = {0.f, 0.f, 0.f, 0.f};
    if (active) {
      C drho = 0, dr0 = 0;
      if constexpr (SH) moments_s<27>(f, drho, rho, ux, uy, uz, CD::S);
      else moments<27>(f, rho, ux, uy, uz);
      const C r0 = rho, x0 = ux, y0 = uy, z0 = uz;
      dr0 = drho;
      if constexpr (BC == BC_REF) {
        if constexpr (SH) overrides_s<27>(p, t, q.y, q.z, drho, rho, ux, uy, uz, CD::S);
        else overrides<27>(p, t, q.y, q.z, rho, ux, uy, uz);
      }
      // conserved rows: rho_f - rho and rho_f u_f - rho u, exactly 0 without an override (in the
      // compute units of the populations: S rho for the shifted storages)
      const C rs0 = SH ? r0 * CD::S : r0, rs = SH ? rho * CD::S : rho;
      dev[0] = SH ? dr0 - drho : r0 - rho;
      dev[1] = rs0 * x0 - rs * ux;
      dev[2] = rs0 * y0 - rs * uy;
      dev[3] = rs0 * z0 - rs * uz;
      if constexpr (SH) {
        const float m15 = m15_of<27>(ux, uy, uz);
        static_for<27>([&](auto I) {
          constexpr int i = decltype(I)::value;
          row[i] = f[i] - geq_i<27, i>(drho, rs, ux, uy, uz, m15);
        });
      } else {
        const C usq = usq_of<27>(ux, uy, uz);
        static_for<27>([&](auto I) {
          constexpr int i = decltype(I)::value;
          row[i] = f[i] - feq_i<27, i>(rho, ux, uy, uz, usq);
        });
      }
      float d26 = 0.f;
      if (p.meq_fix26) {  // the reference table's row 26 against M f_eq
        const float phi[10] = {1.0f, ux, uy, uz, ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz};
        float acc = 0.0f;
        static_for<10>([&](auto J) {
          constexpr int j = decltype(J)::value;
          constexpr float cf = (float)MEQ_RAW[26][j];
          if constexpr (cf != 0.0f && !(SH && j == 0)) acc = __builtin_fmaf(cf, phi[j], acc);
        });
        if constexpr (SH) {
          // both rows less their common value at rest (rho / 27): rho (u terms) - u.u
          d26 = rs * acc - ((ux * ux + uy * uy) + uz * uz) * CD::S;
        } else {
          float ref[27];
          meq_reference<float>(rho, ux, uy, uz, ref);
          d26 = rho * acc - ref[26];
        }
      }
      row[27] = d26;
    } else {
      static_for<27>([&](auto I) {
        f[decltype(I)::value] = 0;
        row[decltype(I)::value] = 0.f;
      });
      row[27] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) row[28 + k] = dev[k];
#pragma unroll
    for (int k = 32; k < ROW; ++k) row[k] = 0.f;
    f32x4 *dst = (f32x4 *)(L + lane * ROW);
#pragma unroll
    for (int k = 0; k < ROW / 4; ++k) dst[k] = f32x4{row[4 * k], row[4 * k + 1], row[4 * k + 2], row[4 * k + 3]};
  }
  // per-lane constants (see build_mfma_table)
  f32x4 tab[10];
  {
    const f32x4 *src = (const f32x4 *)(p.mrt_tab + lane * MFMA_TAB);
#pragma unroll
    for (int k = 0; k < 10; ++k) tab[k] = src[k];
  }
  __syncthreads();

  // ---- phase 2: 4 tiles of 16 nodes on MFMA ----------------------------------------------
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float *row = L + (16 * j + c) * ROW;
    const f32x4 b0 = *(const f32x4 *)(row + 8 * g);      // g of populations 8 gq .. 8 gq + 3
    const f32x4 b1 = *(const f32x4 *)(row + 8 * g + 4);  // and 8 gq + 4 .. 8 gq + 7
    f32x4 o0 = {0, 0, 0, 0}, o1 = {0, 0, 0, 0};
    if constexpr (ENG == MRT_ENGINE_FUSED) {
      // c = C [g; delta26; dev]: A = the fused operator, K slot (t, gq) = population 8 gq + t
      // (slot 27 delta26, slots 28..31 the conserved deviations)
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const float b = s < 4 ? b0[s & 3] : b1[s & 3];
        o0 = mfma4(tab[s >> 2][s & 3], b, o0);        // AC[0][s]
        o1 = mfma4(tab[2 + (s >> 2)][s & 3], b, o1);  // AC[1][s]
      }
    } else {
      const float d26 = row[27];
      f32x4 a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
      if constexpr (ENG == MRT_ENGINE_BF16X3) {
        // tab[0..1] forward A (R = 0, 1), tab[2..3] inverse A, as packed bfloat16 x 8
        const float x[8] = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
        bf16x8 hi, mid, lo;
        split3(x, hi, mid, lo);
        const bf16x8 af0 = __builtin_bit_cast(bf16x8, tab[0]), af1 = __builtin_bit_cast(bf16x8, tab[1]);
        a0 = mfma_bf16(af0, lo, a0);  // smallest terms first
        a1 = mfma_bf16(af1, lo, a1);
        a0 = mfma_bf16(af0, mid, a0);
        a1 = mfma_bf16(af1, mid, a1);
        a0 = mfma_bf16(af0, hi, a0);
        a1 = mfma_bf16(af1, hi, a1);
      } else {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const float b = s < 4 ? b0[s & 3] : b1[s & 3];
          a0 = mfma4(tab[s >> 2][s & 3], b, a0);        // AF[0][s]
          a1 = mfma4(tab[2 + (s >> 2)][s & 3], b, a1);  // AF[1][s]
        }
      }
      // rows 4 gq + r and 16 + 4 gq + r of node c: row 26 is (R = 1, gq = 2, r = 2)
      if (g == 2) a1[2] += d26;
      const f32x4 rate0 = tab[ENG == MRT_ENGINE_BF16X3 ? 4 : 8], rate1 = tab[ENG == MRT_ENGINE_BF16X3 ? 5 : 9];
      a0 *= rate0;
      a1 *= rate1;
      if constexpr (ENG == MRT_ENGINE_BF16X3) {
        const float x[8] = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
        bf16x8 hi, mid, lo;
        split3(x, hi, mid, lo);
        const bf16x8 ai0 = __builtin_bit_cast(bf16x8, tab[2]), ai1 = __builtin_bit_cast(bf16x8, tab[3]);
        o0 = mfma_bf16(ai0, lo, o0);
        o1 = mfma_bf16(ai1, lo, o1);
        o0 = mfma_bf16(ai0, mid, o0);
        o1 = mfma_bf16(ai1, mid, o1);
        o0 = mfma_bf16(ai0, hi, o0);
        o1 = mfma_bf16(ai1, hi, o1);
      } else {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const float b = s < 4 ? a0[s & 3] : a1[s & 3];  // r[16 (s / 4) + 4 gq + s % 4][node]
          o0 = mfma4(tab[4 + (s >> 2)][s & 3], b, o0);   // AI[0][s]
          o1 = mfma4(tab[6 + (s >> 2)][s & 3], b, o1);   // AI[1][s]
        }
      }
    }
    *(f32x4 *)(row + 4 * g) = o0;       // c[4 gq + r]
    *(f32x4 *)(row + 16 + 4 * g) = o1;  // c[16 + 4 gq + r]
  }
  __syncthreads();

  // ---- phase 3: node per lane again, f' = f - c, store ------------------------------------
  C out[27];
  {
    const float *src = L + lane * ROW;
    static_for<27>([&](auto I) { out[decltype(I)::value] = f[decltype(I)::value] - src[decltype(I)::value]; });
  }
  __syncthreads();  // the LDS rows are free for the next tile of a persistent workgroup
  if (!q.valid) return;
  if (t & NODE_SOLID) {
    if constexpr (BC == BC_REF) {  // F(REFLECTED) = BOUNCEDBACK, as the scalar kernel
      static_for<26>([&](auto J) {
        constexpr int jj = decltype(J)::value + 1;
        out[Lat<27>::opp[jj]] = f[jj];
      });
    } else {  // the rest equilibrium, stored with the fluid lanes, as the scalar kernel
      if (!any_fluid) return;
      rest_populations<27, ST>(out);
    }
  }
  store_all<27, ST>(p, q.n, out, SH ? node_bits(q.n, key) : 0u);
}

template <int ST, int BC, int ENG>
__global__ void __launch_bounds__(BLOCK_THREADS) k_step_mrt_mfma(StepParams p) {
  __shared__ __attribute__((aligned(16))) float lds[BLOCK_THREADS / 64][64 * ROW];
  mrt_mfma_tile<ST, BC, ENG>(p, p.stream_order ? stream_tile((int)blockIdx.x, p.gx) : xcd_tile((int)blockIdx.x, p.ntiles),
                             lds);
}

template <int ST, int BC, int ENG>
__global__ void __launch_bounds__(BLOCK_THREADS) k_step_batch_mrt_mfma(BatchArgs b) {
  __shared__ __attribute__((aligned(16))) float lds[BLOCK_THREADS / 64][64 * ROW];
  for_batch_tiles(b, [&](const auto &p, int tile) { mrt_mfma_tile<ST, BC, ENG>(p, tile, lds); });
}

template <class F>
static hipError_t with_mfma_variant(const Config &c, F &&f) {
  auto byST = [&](auto STc) -> hipError_t {
    auto byBC = [&](auto BCc) -> hipError_t {
      switch (c.engine) {
        case MRT_ENGINE_MFMA: return f(STc, BCc, IC<MRT_ENGINE_MFMA>{});
        case MRT_ENGINE_FUSED: return f(STc, BCc, IC<MRT_ENGINE_FUSED>{});
        case MRT_ENGINE_BF16X3: return f(STc, BCc, IC<MRT_ENGINE_BF16X3>{});
      }
      return hipErrorInvalidValue;
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
  return with_mfma_variant(c, [&](auto STc, auto BCc, auto Ec) -> hipError_t {
    hipLaunchKernelGGL((k_step_batch_mrt_mfma<decltype(STc)::value, decltype(BCc)::value, decltype(Ec)::value>),
                       dim3(grid), dim3(BLOCK_THREADS), 0, s, b);
    return hipGetLastError();
  });
}

hipError_t launch_step_mrt_mfma(const Config &c, const StepParams &p, hipStream_t s) {
  return with_mfma_variant(c, [&](auto STc, auto BCc, auto Ec) -> hipError_t {
    hipLaunchKernelGGL((k_step_mrt_mfma<decltype(STc)::value, decltype(BCc)::value, decltype(Ec)::value>),
                       dim3(p.ntiles), dim3(BLOCK_THREADS), 0, s, p);
    return hipGetLastError();
  });
}

}  // namespace lbm
