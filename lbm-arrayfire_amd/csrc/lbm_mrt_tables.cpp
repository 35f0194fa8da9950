// lbm_mrt_tables.cpp — per-lane operand table of the MFMA MRT kernel.
//
// The collision is m = M f, m* = m - s (m - meq), f' = M^-1 m* (rust/examples/d3q27_mrt_lid.rs:
// 336,385-387) with meq = E phi linear in the 13 monomials
// phi = (rho, rho ux, rho uy, rho uz, rho ux^2, rho uy^2, rho uz^2, rho ux uy, rho ux uz, rho uy uz,
//        ux^2, uy^2, uz^2).
// E is either the reference's equilibrium-moment table (rust/examples/d3q27_mrt_lid.rs:357-383,
// raw basis) or M W with f_eq = W phi (projected, any basis).
//
// Lane l = 16 g + c of a wavefront supplies, for v_mfma_f32_16x16x4_f32 (A[row][k] from lane
// row + 16 k, B[k][col] from lane col + 16 k, D[4 (l/16) + r][l % 16] in lane l, register r):
//   AF[R][t] = M[16R + c][8g + t]                 forward, K slot (t, g) = population 8g + t
//   AP[R][t] = s_row E[16R + c][4g + t]           equilibrium term, K slot = monomial 4g + t
//   SC[R][r] = 1 - s[16R + 4g + r]                scale of the accumulator rows this lane holds
//   AI[R][t] = M^-1[16R + c][16(t/4) + 4g + t%4]  inverse, K slot = accumulator register t%4 of
//                                                 row tile t/4, i.e. moment 16(t/4) + 4g + t%4
// Entries outside the 27 x 27 (x 13) matrices are zero.
//
// Fused engine: lane l supplies AC[R][t] = [C | D][16R + c][12g + t] with
// C = M^-1 diag(1 - s) M and D = M^-1 diag(s) E (K slots 0..26 populations, 27..39 monomials).
#include <vector>

#include "lbm_internal.hpp"
#include "mrt_tables.h"

namespace lbm {

void build_mfma_table(int collision, int meq_mode, const double *s, bool fused, float *tab) {
  double M[27][27], Mi[27][27], E[27][13] = {};
  const bool raw = collision == CO_MRT_RAW;
  for (int k = 0; k < 27; ++k)
    for (int i = 0; i < 27; ++i) {
      M[k][i] = raw ? MRAW[k][i] : MGS[k][i];
      Mi[i][k] = raw ? MRAW_INV24[i][k] / 24.0 : (double)MGS[k][i] / (double)MGS_NORM2[k];
    }
  if (raw && meq_mode == 0) {
    const double c2 = 1.0 / 3.0;
    E[0][0] = 1;
    E[1][1] = E[2][2] = E[3][3] = 1;
    E[4][7] = E[5][8] = E[6][9] = 1;
    E[7][0] = E[7][4] = E[7][5] = E[7][6] = 1;
    E[8][4] = 1; E[8][5] = -1;
    E[9][4] = 1; E[9][6] = -1;
    E[10][1] = E[11][1] = E[12][2] = E[13][3] = E[14][2] = E[15][3] = c2;
    E[17][0] = c2 * c2; E[17][4] = E[17][5] = c2;
    E[18][0] = c2 * c2; E[18][4] = E[18][6] = c2;
    E[19][0] = c2 * c2; E[19][5] = E[19][6] = c2;
    E[20][9] = E[21][8] = E[22][7] = c2;
    E[23][1] = E[24][2] = E[25][3] = c2 * c2;
    E[26][10] = E[26][11] = E[26][12] = 1;
    E[26][0] = c2 * c2 * c2;
  } else {
    // f_eq,i = w_i (rho + 3 c.(rho u) + 4.5 (c.u)^2 rho - 1.5 rho u^2) = sum_t W[i][t] phi_t
    double W[27][13] = {};
    for (int i = 0; i < 27; ++i) {
      const double w = Lat<27>::w[i], c[3] = {(double)Lat<27>::cx[i], (double)Lat<27>::cy[i],
                                              (double)Lat<27>::cz[i]};
      W[i][0] = w;
      for (int a = 0; a < 3; ++a) {
        W[i][1 + a] = 3 * w * c[a];
        W[i][4 + a] = w * (4.5 * c[a] * c[a] - 1.5);
      }
      W[i][7] = 9 * w * c[0] * c[1];
      W[i][8] = 9 * w * c[0] * c[2];
      W[i][9] = 9 * w * c[1] * c[2];
    }
    for (int k = 0; k < 27; ++k)
      for (int t = 0; t < 13; ++t) {
        double acc = 0;
        for (int i = 0; i < 27; ++i) acc += M[k][i] * W[i][t];
        E[k][t] = acc;
      }
  }
  if (fused) {
    // K slot k of the fused operator: population k (k < 27), monomial k - 27 (27 <= k < 40), zero.
    double Cf[27][48] = {};
    for (int i = 0; i < 27; ++i) {
      for (int j = 0; j < 27; ++j) {
        double acc = 0;
        for (int k = 0; k < 27; ++k) acc += Mi[i][k] * (1.0 - s[k]) * M[k][j];
        Cf[i][j] = acc;
      }
      for (int t = 0; t < 13; ++t) {
        double acc = 0;
        for (int k = 0; k < 27; ++k) acc += Mi[i][k] * s[k] * E[k][t];
        Cf[i][27 + t] = acc;
      }
    }
    for (int lane = 0; lane < 64; ++lane) {
      const int g = lane >> 4, c = lane & 15;
      float *T = tab + lane * MFMA_TAB;
      for (int k = 0; k < MFMA_TAB; ++k) T[k] = 0.f;
      for (int R = 0; R < 2; ++R)
        for (int t = 0; t < 12; ++t) {
          const int row = 16 * R + c;
          T[R * 12 + t] = row < 27 ? (float)Cf[row][12 * g + t] : 0.f;
        }
    }
    return;
  }
  for (int lane = 0; lane < 64; ++lane) {
    const int g = lane >> 4, c = lane & 15;
    float *T = tab + lane * MFMA_TAB;
    for (int R = 0; R < 2; ++R) {
      const int row = 16 * R + c;
      for (int t = 0; t < 8; ++t) {
        const int col = 8 * g + t;
        T[R * 8 + t] = (row < 27 && col < 27) ? (float)M[row][col] : 0.f;
        const int k = 16 * (t / 4) + 4 * g + t % 4;
        T[32 + R * 8 + t] = (row < 27 && k < 27) ? (float)Mi[row][k] : 0.f;
      }
      for (int t = 0; t < 4; ++t) {
        const int term = 4 * g + t;
        T[16 + R * 4 + t] = (row < 27 && term < 13) ? (float)(s[row] * E[row][term]) : 0.f;
        const int rr = 16 * R + 4 * g + t;
        T[24 + R * 4 + t] = rr < 27 ? (float)(1.0 - s[rr]) : 0.f;
      }
    }
  }
}

}  // namespace lbm
