// lbm_lattice.hpp — velocity sets and storage formats shared by the host code and the kernels.
//
// Velocity order, weights and opposite maps follow the reference tables:
//   D2Q9  cpp/d2q9_bgk_lid.cpp:72-84 (cx, cy, weights, nb_index_arr)
//   D3Q27 rust/examples/d3q27_mrt_lid.rs:98-105 (ex, ey, ez, nbidx), weights
//         cpp/d3q27_bgk_lid.cpp:86-89
//   D3Q19 no reference table (cpp/d3q19_mrt_channel.cpp still uses the 27 velocities): the
//         subset of D3Q27 without the corners, in D3Q27's order
#pragma once
#include <cstdint>

namespace lbm {

// Node-type bits (public ones mirror LBM_NODE_* of lbm_capi.h).
constexpr uint8_t NODE_SOLID = 0x01, NODE_MOVING = 0x02, NODE_INLET = 0x04, NODE_OUTLET = 0x08,
                  NODE_RHO_FIX = 0x10;
// Internal: a non-solid node whose pull needs the boundary path (some pull source is solid, or
// the node is an inlet/outlet), of those the ones with a solid pull source, and of those the
// ones with a moving solid pull source.  Set by the library at create time, PHYSICAL mode only.
constexpr uint8_t NODE_SLOW = 0x80, NODE_WALLNB = 0x40, NODE_WALLMOV = 0x20;

enum Storage { ST_F64 = 0, ST_F32 = 1, ST_F16S = 2, ST_BF16S = 3 };
enum Collision { CO_BGK = 0, CO_MRT_RAW = 1, CO_MRT_GS = 2 };
enum BcMode { BC_REF = 0, BC_PHYS = 1 };

template <int Q> struct Lat;

template <> struct Lat<9> {
  static constexpr int D = 2;
  static constexpr int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr double w[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9.,
                                  1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

template <> struct Lat<27> {
  static constexpr int D = 3;
  static constexpr int cx[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
  static constexpr int cy[27] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
  static constexpr int cz[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
  static constexpr int opp[27] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15, 26, 25, 24, 23, 22, 21, 20, 19};
  static constexpr double w[27] = {8. / 27.,
                                   2. / 27., 2. / 27., 2. / 27., 2. / 27., 2. / 27., 2. / 27.,
                                   1. / 54., 1. / 54., 1. / 54., 1. / 54., 1. / 54., 1. / 54.,
                                   1. / 54., 1. / 54., 1. / 54., 1. / 54., 1. / 54., 1. / 54.,
                                   1. / 216., 1. / 216., 1. / 216., 1. / 216.,
                                   1. / 216., 1. / 216., 1. / 216., 1. / 216.};
};

// D3Q19: the rest, face and edge directions of D3Q27, in D3Q27's order (its first 19 entries;
// opposites stay within them); weights 1/3, 1/18, 1/36.
template <> struct Lat<19> {
  static constexpr int D = 3;
  static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  static constexpr double w[19] = {1. / 3.,
                                   1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                                   1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                                   1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

// Bytes of one stored population.
inline int storage_bytes(int st) { return st == ST_F64 ? 8 : (st == ST_F32 ? 4 : 2); }

}  // namespace lbm
