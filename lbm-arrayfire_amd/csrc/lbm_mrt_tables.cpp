// lbm_mrt_tables.cpp — per-lane operand tables of the MFMA MRT kernels.
//
// The collision is evaluated in correction form, f' = f - M^-1 S M g with g = f - f_eq
// (rust/examples/d3q27_mrt_lid.rs:336,385-387 with meq = M f_eq; the reference table's row-26
// difference enters as delta26, lbm_mrt_mfma.hip).  M^-1 = (n M^-1) / n with n M^-1 integral:
// the raw basis' 24 M^-1 (n = 24 for every row) or the Gram-Schmidt basis' M^T (n = the squared
// row norms), so the rates become s_k / n_k.
//
// Lane l = 16 g + c of a wavefront.  v_mfma_f32_16x16x4_f32 takes A[row][k] from lane row + 16 k,
// B[k][col] from lane col + 16 k and leaves D[4 (l / 16) + r][l % 16] in lane l, register r;
// v_mfma_f32_16x16x32_bf16 takes A[row][8 (l / 16) + j] and B[8 (l / 16) + j][col] (j = 0..7, one
// bfloat16 x 8 per lane) and leaves D in the same layout.  In both the lane group g supplies the
// K slots 8 g .. 8 g + 7 (the 16x16x4 form over 8 K-steps t, slot 8 g + t in step t).
//
// K slots 0..26 carry g, slot 27 delta26 and slots 28..31 the conserved rows' deviations (rho,
// jx, jy, jz) computed on the vector unit; the forward operator's rows 0..3 pick those up through
// unit columns (M' below), so the conserved rows never see the MFMA sums of g.
// MRT_ENGINE_MFMA (floats 0..39):
//   AF[R][t] = M'[16R + c][8g + t]                   forward, K slot = population 8g + t
//   AI[R][t] = (n M^-1)[16R + c][16(t/4) + 4g + t%4]  inverse, K slot = accumulator register t%4 of
//                                                   row tile t/4, i.e. moment 16(t/4) + 4g + t%4
//   RT[R][r] = s_k / n_k, k = 16R + 4g + r           rates of the rows this lane's accumulators hold
// MRT_ENGINE_FUSED (floats 0..15):
//   AC[R][t] = C[16R + c][8g + t],  C = [M^-1 S' M | M^-1[:, 26] s_26 | M^-1[:, 0..3] S_c]
//   (K slot 27: delta26; 28..31: the conserved deviations; S' = S with the rows 0..3 set to 0)
// MRT_ENGINE_BF16X3 (floats 0..23; A operands as packed bfloat16, exact: integers up to 24):
//   dwords 0..3   forward  A, R = 0: M'[c][8g + j]
//   dwords 4..7   forward  A, R = 1: M'[16 + c][8g + j]
//   dwords 8..11  inverse  A, R = 0: (n M^-1)[c][k(j)]       k(j) = j < 4 ? 4g + j : 16 + 4g + j - 4
//   dwords 12..15 inverse  A, R = 1: (n M^-1)[16 + c][k(j)]
//   floats 16..23 RT as above
// M' = M with its rows 0..3 (1, cx, cy, cz) replaced by the unit rows e_{28 + k}.  Entries outside
// these matrices are zero.
#include <cstdint>
#include <cstring>

#include "lbm_internal.hpp"
#include "mrt_tables.h"

namespace lbm {

namespace {
uint16_t bf16_exact(double v) {  // v is a small integer: its float's upper half is exact
  const float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (uint16_t)(u >> 16);
}
}  // namespace

void build_mfma_table(int collision, const double *s, int engine, float *tab) {
  const bool raw = collision == CO_MRT_RAW;
  double M[27][27], nMi[27][27], n[27];
  for (int k = 0; k < 27; ++k) {
    n[k] = raw ? 24.0 : (double)MGS_NORM2[k];
    for (int i = 0; i < 27; ++i) {
      M[k][i] = raw ? MRAW[k][i] : MGS[k][i];
      nMi[i][k] = raw ? MRAW_INV24[i][k] : MGS[k][i];
    }
  }
  auto m_at = [&](int row, int col) {  // M'
    if (row < 4) return col == 28 + row ? 1.0 : 0.0;
    return (row < 27 && col < 27) ? M[row][col] : 0.0;
  };
  auto nmi_at = [&](int row, int k) { return (row < 27 && k < 27) ? nMi[row][k] : 0.0; };
  auto rate = [&](int k) { return k < 27 ? s[k] / n[k] : 0.0; };
  for (int lane = 0; lane < 64; ++lane) {
    const int g = lane >> 4, c = lane & 15;
    float *T = tab + lane * MFMA_TAB;
    for (int k = 0; k < MFMA_TAB; ++k) T[k] = 0.f;
    if (engine == MRT_ENGINE_FUSED) {
      for (int R = 0; R < 2; ++R)
        for (int t = 0; t < 8; ++t) {
          const int i = 16 * R + c, j = 8 * g + t;
          double acc = 0;
          if (i < 27 && j < 27)
            for (int k = 4; k < 27; ++k) acc += nMi[i][k] / n[k] * s[k] * M[k][j];
          else if (i < 27 && j == 27)
            acc = nMi[i][26] / n[26] * s[26];
          else if (i < 27 && j >= 28)
            acc = nMi[i][j - 28] / n[j - 28] * s[j - 28];
          T[R * 8 + t] = (float)acc;
        }
      continue;
    }
    float *RT = T + (engine == MRT_ENGINE_BF16X3 ? 16 : 32);
    for (int R = 0; R < 2; ++R)
      for (int r = 0; r < 4; ++r) RT[R * 4 + r] = (float)rate(16 * R + 4 * g + r);
    if (engine == MRT_ENGINE_BF16X3) {
      uint16_t *B = reinterpret_cast<uint16_t *>(T);
      for (int R = 0; R < 2; ++R)
        for (int j = 0; j < 8; ++j) {
          const int k = j < 4 ? 4 * g + j : 16 + 4 * g + j - 4;
          B[R * 8 + j] = bf16_exact(m_at(16 * R + c, 8 * g + j));
          B[16 + R * 8 + j] = bf16_exact(nmi_at(16 * R + c, k));
        }
      continue;
    }
    for (int R = 0; R < 2; ++R)
      for (int t = 0; t < 8; ++t) {
        T[R * 8 + t] = (float)m_at(16 * R + c, 8 * g + t);
        T[16 + R * 8 + t] = (float)nmi_at(16 * R + c, 16 * (t / 4) + 4 * g + t % 4);
      }
  }
}

}  // namespace lbm
