// lbm_oracle.cpp — CPU restatement of the reference lattice Boltzmann time step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP kernels are checked against and
// the CPU baseline timed beside them.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it; the product library never does.
//
// What it restates (paths relative to the reference tree):
//  * lattice tables     cpp/d2q9_bgk_lid.cpp:72-84, rust/examples/d3q27_mrt_lid.rs:98-105,
//                       cpp/d3q27_bgk_lid.cpp:86-89 (D3Q27 weights)
//  * pull map           stream() cpp/d2q9_bgk_lid.cpp:18-28,86-87 (af::shift = circular shift,
//                       out[x] = in[x - c]); rust/examples/d3q27_mrt_lid.rs:12-48,107-108
//  * REFERENCE step     cpp/d2q9_bgk_lid.cpp:124-155 in the same pass order: gather through a
//                       materialised index array, bounce-back capture (TO_REFLECT), rho/u
//                       reductions, Dirichlet overrides (lid, inlet, solid, density columns:
//                       rust/examples/d2q9_bgk_channel.rs:190-201), equilibrium + BGK relaxation
//                       at every node, bounce-back scatter (REFLECTED)
//  * MRT                rust/examples/d3q27_mrt_lid.rs:167-199 (raw monomial M, inverse),
//                       :111-150 (rates), :336-387 (m = M f, meq table, relax, f = M^-1 m*),
//                       :228-287 (initial state rho0/27 + one collision)
//  * f32 index rounding the FP32 builds index through float arrays (cpp/d2q9_bgk_lid.cpp:86,
//                       rust/examples/d2q9_bgk_lid.rs:97); f32_index_compat reproduces it.
//  * D3Q19              no reference counterpart (cpp/d3q19_mrt_channel.cpp is a work in progress
//                       on 27 velocities): D3Q27's rest, face and edge directions, weights
//                       1/3, 1/18, 1/36; MRT in the orthogonal basis of d'Humieres et al. (2002)
//                       with projected equilibrium moments; initial state f_eq(rho0, 0)
//  * PHYSICAL step      no reference counterpart: half-way / moving-wall bounce-back, velocity
//                       inlet by Zou-He (non-equilibrium bounce-back) or by non-equilibrium
//                       extrapolation, zero-gradient outlet at a reference density
//                       (non-equilibrium extrapolation of every population).  In 3D lattices a
//                       solid node takes the rest equilibrium w_i each step when its aligned
//                       block of 64 node indices holds a non-solid node (the kernels' wavefronts
//                       store whole blocks) and keeps its values otherwise; in 2D lattices solid
//                       nodes keep their values.  No step reads them.
//  * Shan-Chen          cpp/d2q9_shanchen.cpp:39-181 with its work-in-progress defects corrected
//                       (oracle_sc_run); pinned in tests/test_shanchen.py by a numpy restatement,
//                       the single-fluid reduction at G = 0, conservation, a point-perturbation
//                       closed form and the sign of the coupling.
//
// Shifted storage: F16S holds (f - w_i) * 2^15 in IEEE half, BF16S holds f - w_i in bfloat16.
// Neither has a reference counterpart; this restates the kernels' own algorithm for them: the
// state is held as g = f - w_i, every step computes on g (density deviation drho = sum g, rho =
// 1 + drho, equilibria as deviations w_i (drho + rho P_i)) and rounds g to the storage with
// stochastic rounding, restated bit for bit from the gfx950 conversions (f2h_sr, f2bf_sr) with
// the kernels' random bits (mix32 of the node index and the step count).  Initial states are
// rounded to nearest.  Test switch: desc.reserved bit 0 makes the steps round to nearest too
// (the rounding whose frozen errors the stochastic rounding removes), for tests that show the
// difference; the library has no such mode.
//
// FP32 mode contracts the equilibrium and the BGK relaxation into FMAs (see feq()) and takes
// u = j * (1/rho), matching the kernels; FP64 mode keeps every operation separately rounded and
// divides as the reference does.
//
// Summation order inside a node is ascending population index; ArrayFire's reduction order is
// not observable without ArrayFire, so FP32 bitwise parity with ArrayFire is unpinnable (FP64
// differences stay at round-off).  Pinned in tests/test_oracle.py by closed forms, brute force
// on tiny grids against an independent numpy restatement, and exact-algebra identities.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <chrono>
#include <type_traits>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {
typedef struct {
  int32_t lattice, nx, ny, nz, collision, storage, bc_mode, mrt_engine, meq_mode, inlet_profile, inlet_closure,
      reserved;
  double omega;
  double s[27];
  double u_wall[3];
  double u_in;
  double rho_bc;
} oracle_desc;  // field-for-field the lbm_desc of the library's C ABI
}

namespace {

enum { F64 = 0, F32 = 1, F16S = 2, BF16S = 3 };
enum { BGK = 0, MRT_RAW = 1, MRT_GS = 2 };
enum { SOLID = 1, MOVING = 2, INLET = 4, OUTLET = 8, RHO_FIX = 16 };

// ---- lattice tables (transcribed from the reference arrays) ------------------------------
const int CX9[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
const int CY9[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
const int NBIDX9[8] = {2, 3, 0, 1, 6, 7, 4, 5};
const int CX27[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
const int CY27[27] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
const int CZ27[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
const int NBIDX27[26] = {1, 0, 3, 2, 5, 4, 9, 8, 7, 6, 13, 12, 11, 10, 17, 16, 15, 14, 25, 24, 23, 22, 21, 20, 19, 18};

struct Lattice {
  int Q, D;
  int cx[27], cy[27], cz[27], opp[27];
  double w[27];  // exact value as double; cast to T where the reference computes in T
};

Lattice make_lattice(int Q) {
  Lattice L{};
  L.Q = Q;
  if (Q == 9) {
    L.D = 2;
    for (int i = 0; i < 9; ++i) { L.cx[i] = CX9[i]; L.cy[i] = CY9[i]; L.cz[i] = 0; }
    L.opp[0] = 0;
    for (int j = 0; j < 8; ++j) L.opp[j + 1] = NBIDX9[j] + 1;
    const double t1 = 4. / 9., t2 = 1. / 9., t3 = 1. / 36.;
    const double w[9] = {t1, t2, t2, t2, t2, t3, t3, t3, t3};
    for (int i = 0; i < 9; ++i) L.w[i] = w[i];
  } else if (Q == 19) {  // D3Q27 without its corners, in D3Q27's order
    L.D = 3;
    for (int i = 0; i < 19; ++i) { L.cx[i] = CX27[i]; L.cy[i] = CY27[i]; L.cz[i] = CZ27[i]; }
    L.opp[0] = 0;
    for (int j = 0; j < 18; ++j) L.opp[j + 1] = NBIDX27[j] + 1;
    for (int i = 0; i < 19; ++i) L.w[i] = (i == 0) ? 1. / 3. : (i <= 6 ? 1. / 18. : 1. / 36.);
  } else {
    L.D = 3;
    for (int i = 0; i < 27; ++i) { L.cx[i] = CX27[i]; L.cy[i] = CY27[i]; L.cz[i] = CZ27[i]; }
    L.opp[0] = 0;
    for (int j = 0; j < 26; ++j) L.opp[j + 1] = NBIDX27[j] + 1;
    const double t1 = 8. / 27., t2 = 2. / 27., t3 = 1. / 54., t4 = 1. / 216.;
    for (int i = 0; i < 27; ++i) L.w[i] = (i == 0) ? t1 : (i <= 6 ? t2 : (i <= 18 ? t3 : t4));
  }
  return L;
}

// ---- storage codecs (round to nearest even) ----------------------------------------------
uint16_t f2h(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // >= 65520 rounds to inf
  if (ax < 0x38800000u) {                                    // half subnormal range
    float v;
    std::memcpy(&v, &ax, 4);
    uint32_t m = (uint32_t)std::nearbyint((double)v * 16777216.0);  // exact scaling, RNE
    return (uint16_t)(sign | m);
  }
  uint32_t e = (ax >> 23) - 127 + 15, mant = ax & 0x7fffffu;
  uint32_t h = (e << 10) | (mant >> 13), rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)(sign | h);
}
float h2f(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  float v;
  if (e == 0) {
    v = (float)std::ldexp((double)m, -24);
    return sign ? -v : v;
  }
  uint32_t x;
  if (e == 31) x = sign | 0x7f800000u | (m << 13);
  else x = sign | ((e - 15 + 127) << 23) | (m << 13);
  std::memcpy(&v, &x, 4);
  return v;
}
uint16_t f2bf(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  uint32_t lsb = (x >> 16) & 1u;
  x += 0x7fffu + lsb;
  return (uint16_t)(x >> 16);
}
float bf2f(uint16_t b) {
  uint32_t x = (uint32_t)b << 16;
  float v;
  std::memcpy(&v, &x, 4);
  return v;
}

// Stochastic rounding as v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32 do it (pinned against MI355X
// outputs by tests/golden/conversions/sr_conversions.npz): |x| is cut to 13 bits below the
// half-precision ulp of its binade (the subnormal ulp 2^-24 below 2^-14), r >> 19 is added, and the sum is cut to
// the ulp; values from 65520 up give inf.  bfloat16: r >> 16 added to the magnitude bits, cut to
// their upper half.  The sign is kept.
uint16_t f2h_sr(float f, uint32_t r) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
  int e = (int)(ax >> 23) - 127;
  uint64_t m = ax & 0x7fffffu;
  if (ax >> 23) m |= 0x800000u;
  else e = -126;
  const int ulp_e = e >= -14 ? e - 10 : -24;           // log2 of the half ulp of |x|
  const int sh = (ulp_e - 13) - (e - 23);              // m in units of 2^(ulp_e - 13)
  uint64_t t = sh >= 0 ? (sh >= 64 ? 0 : m >> sh) : m << (-sh);
  t += r >> 19;
  const uint64_t k = t >> 13;                           // |result| / 2^ulp_e
  const uint64_t bits = ulp_e == -24 ? k : ((uint64_t)(e + 14) << 10) + k;
  return (uint16_t)(sign | (bits >= 0x7c00u ? 0x7c00u : (uint32_t)bits));
}
uint16_t f2bf_sr(float f, uint32_t r) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = x & 0x80000000u, ax = x & 0x7fffffffu;
  if (ax > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
  return (uint16_t)((sign | (ax + (r >> 16))) >> 16);
}

// Random bits of the shifted storages' stochastic rounding, as the kernels draw them: node n in
// step t takes mix32(n ^ mix32(t + 0x9E3779B9)), population i those bits rotated left by 7 i.
uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
uint32_t rotl32(uint32_t h, int k) { return k == 0 ? h : (h << k) | (h >> (32 - k)); }

// ---- MRT matrices ------------------------------------------------------------------------
// Raw monomial rows in the reference's order (rust/examples/d3q27_mrt_lid.rs:167-193).
void raw_moment_matrix(const Lattice &L, double M[27][27]) {
  for (int i = 0; i < 27; ++i) {
    const double x = L.cx[i], y = L.cy[i], z = L.cz[i];
    const double r[27] = {1, x, y, z, x * y, x * z, y * z, x * x + y * y + z * z, x * x - y * y,
                          x * x - z * z, x * y * y, x * z * z, y * x * x, z * x * x, y * z * z,
                          z * y * y, x * y * z, x * x * y * y, x * x * z * z, y * y * z * z,
                          x * x * y * z, x * y * y * z, x * y * z * z, x * y * y * z * z,
                          x * x * y * z * z, x * x * y * y * z, x * x * y * y * z * z};
    for (int k = 0; k < 27; ++k) M[k][i] = r[k];
  }
}
// Classical Gram-Schmidt of the raw rows, in row order, unweighted inner product over the 27
// velocities.  Rows are left unnormalised (any row scaling leaves the collision unchanged).
void gs_moment_matrix(const Lattice &L, double M[27][27]) {
  double R[27][27];
  raw_moment_matrix(L, R);
  long double G[27][27];
  for (int k = 0; k < 27; ++k) {
    for (int i = 0; i < 27; ++i) G[k][i] = R[k][i];
    for (int j = 0; j < k; ++j) {
      long double num = 0, den = 0;
      for (int i = 0; i < 27; ++i) { num += (long double)R[k][i] * G[j][i]; den += G[j][i] * G[j][i]; }
      const long double c = num / den;
      for (int i = 0; i < 27; ++i) G[k][i] -= c * G[j][i];
    }
  }
  for (int k = 0; k < 27; ++k)
    for (int i = 0; i < 27; ++i) M[k][i] = (double)G[k][i];
}
// Gauss-Jordan inverse with partial pivoting in long double (stands in for the reference's
// LAPACK inverse, rust/examples/d3q27_mrt_lid.rs:196).
// D3Q19 moment rows of d'Humieres, Ginzburg, Krafczyk, Lallemand & Luo, Phil. Trans. R. Soc. A
// 360 (2002) 437: rho, e = 19c^2 - 30, eps = (21c^4 - 53c^2 + 24)/2, j_a = c_a,
// q_a = (5c^2 - 9)c_a, 3p_xx = 3c_x^2 - c^2, 3pi_xx = (3c^2 - 5)(3c_x^2 - c^2), p_ww = c_y^2 - c_z^2,
// pi_ww = (3c^2 - 5)(c_y^2 - c_z^2), p_xy, p_yz, p_xz = c_a c_b, m_x = (c_y^2 - c_z^2)c_x,
// m_y = (c_z^2 - c_x^2)c_y, m_z = (c_x^2 - c_y^2)c_z (rows in that order, j and q interleaved).
void d3q19_moment_matrix(const Lattice &L, double M[27][27]) {
  for (int i = 0; i < 19; ++i) {
    const double x = L.cx[i], y = L.cy[i], z = L.cz[i], c2 = x * x + y * y + z * z;
    const double r[19] = {1, 19 * c2 - 30, (21 * c2 * c2 - 53 * c2 + 24) / 2, x, (5 * c2 - 9) * x, y,
                          (5 * c2 - 9) * y, z, (5 * c2 - 9) * z, 3 * x * x - c2, (3 * c2 - 5) * (3 * x * x - c2),
                          y * y - z * z, (3 * c2 - 5) * (y * y - z * z), x * y, y * z, x * z,
                          (y * y - z * z) * x, (z * z - x * x) * y, (x * x - y * y) * z};
    for (int k = 0; k < 19; ++k) M[k][i] = r[k];
  }
}

// Gauss-Jordan inverse of the leading n x n block, partial pivoting, in long double.
bool invert27(const double A[27][27], double Ainv[27][27], int n = 27) {
  long double a[27][54];
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < 2 * n; ++c) a[r][c] = (c < n) ? A[r][c] : (c - n == r ? 1.0L : 0.0L);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs((double)a[r][c]) > std::fabs((double)a[p][c])) p = r;
    if (std::fabs((double)a[p][c]) < 1e-12) return false;
    if (p != c)
      for (int k = 0; k < 2 * n; ++k) std::swap(a[p][k], a[c][k]);
    const long double piv = a[c][c];
    for (int k = 0; k < 2 * n; ++k) a[c][k] /= piv;
    for (int r = 0; r < n; ++r) {
      if (r == c || a[r][c] == 0) continue;
      const long double f = a[r][c];
      for (int k = 0; k < 2 * n; ++k) a[r][k] -= f * a[c][k];
    }
  }
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) Ainv[r][c] = (double)a[r][c + n];
  return true;
}

// ---- simulation state ---------------------------------------------------------------------
template <class T>
struct Sim {
  oracle_desc d;
  Lattice L;
  int64_t N;
  int nx, ny, nz;
  const uint8_t *type;
  std::vector<T> F;          // current post-collision populations (decoded for shifted storage)
  std::vector<int64_t> nb;   // materialised pull map (REFERENCE mode)
  std::vector<int64_t> on;   // linear indices of solid nodes, ascending (af::where)
  // MRT
  T M[27][27], Minv[27][27], s[27];
  double Md[27][27];
  bool mrt = false;
  T w[27];
  // shifted storage (F16S, BF16S): F holds g = f - w_i and the arithmetic is on deviations
  bool sh = false;
  uint32_t tstep = 0;  // step count of the stochastic rounding

  Sim(const oracle_desc &dd, const uint8_t *types) : d(dd), L(make_lattice(dd.lattice)), type(types) {
    nx = d.nx; ny = d.ny; nz = d.nz;
    N = (int64_t)nx * ny * nz;
    sh = std::is_same<T, float>::value && (d.storage == F16S || d.storage == BF16S);
    for (int i = 0; i < L.Q; ++i) w[i] = (T)L.w[i];
    if (d.collision != BGK) {
      mrt = true;
      double Mi[27][27];
      if (L.Q == 19) d3q19_moment_matrix(L, Md);
      else if (d.collision == MRT_RAW) raw_moment_matrix(L, Md);
      else gs_moment_matrix(L, Md);
      invert27(Md, Mi, L.Q);
      for (int k = 0; k < L.Q; ++k)
        for (int i = 0; i < L.Q; ++i) { M[k][i] = (T)Md[k][i]; Minv[k][i] = (T)Mi[k][i]; }
      for (int k = 0; k < L.Q; ++k) s[k] = (T)d.s[k];
    }
  }

  int64_t idx(int x, int y, int z) const { return (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z); }
  static int wrap(int a, int n) { return ((a % n) + n) % n; }

  // nb_index[p(x,i)] = p((x - c_i) mod dims, i); f32 builds round the index through float.
  // Returns false when a rounded index falls outside the N*Q populations (above 2^24 consecutive
  // floats are 2, 4, ... apart, and the last index can round up to N*Q): the reference's FP32
  // build would then read past its array, which this restatement refuses to do.
  bool build_pull_map(bool f32_compat) {
    nb.resize((size_t)(N * L.Q));
    for (int i = 0; i < L.Q; ++i)
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < nx; ++x) {
            const int64_t src = idx(wrap(x - L.cx[i], nx), wrap(y - L.cy[i], ny), wrap(z - L.cz[i], nz)) + N * i;
            nb[(size_t)(idx(x, y, z) + N * i)] = f32_compat ? (int64_t)(float)src : src;
          }
    for (const int64_t v : nb)
      if (v < 0 || v >= N * L.Q) return false;
    on.clear();
    for (int64_t n = 0; n < N; ++n)
      if (type[n] & SOLID) on.push_back(n);
    return true;
  }

  T inlet_u(int y, int z) const {
    if (d.inlet_profile == 0) return (T)d.u_in;
    const T H = (T)(ny - 2), eta = (T)y - (T)0.5;
    T u = (((T)4 * (T)d.u_in) * eta * (H - eta)) / (H * H);
    if (L.D == 3) {
      const T Hz = (T)(nz - 2), ez = (T)z - (T)0.5;
      u = u * (((T)4 * ez * (Hz - ez)) / (Hz * Hz));
    }
    return u;
  }

  // Reference overrides in the reference's order (cpp/d2q9_bgk_lid.cpp:141-146,
  // rust/examples/d2q9_bgk_channel.rs:190-201, rust/examples/d3q27_mrt_lid.rs:323-330).
  void overrides(int64_t n, int y, int z, T &rho, T &ux, T &uy, T &uz) const {
    const uint8_t t = type[n];
    if (t & MOVING) { ux = (T)d.u_wall[0]; uy = (T)d.u_wall[1]; uz = (T)d.u_wall[2]; }
    if (t & INLET) ux = inlet_u(y, z);
    if (t & SOLID) { ux = 0; uy = 0; uz = 0; rho = 0; }
    if (t & RHO_FIX) rho = (T)d.rho_bc;
  }
  // the same in the shifted form: drho follows rho
  void overrides_s(int64_t n, int y, int z, T &drho, T &rho, T &ux, T &uy, T &uz) const {
    overrides(n, y, z, rho, ux, uy, uz);
    if (type[n] & SOLID) drho = (T)-1;
    if (type[n] & RHO_FIX) drho = (T)(d.rho_bc - 1.0);
  }

  // Equilibrium.  FP64: the reference's expression term by term (cpp/d2q9_bgk_lid.cpp:149-151).
  // FP32: the contracted form w rho (A + eu (3 + 4.5 eu)), A = 1 - 1.5 u.u, as the reference's
  // FP32 GPU builds get from their JIT compilers' multiply-add contraction.
  T usq_term(T ux, T uy, T uz) const {
    if constexpr (std::is_same<T, float>::value) {
      T u2 = std::fma(uy, uy, ux * ux);
      if (L.D == 3) u2 = std::fma(uz, uz, u2);
      return std::fma(-1.5f, u2, 1.0f);
    } else {
      T u2 = ux * ux + uy * uy;
      if (L.D == 3) u2 = u2 + uz * uz;
      return u2;
    }
  }
  // the bracket of f_eq,i = (w_i rho) qeq
  T qeq(int i, T ux, T uy, T uz, T usq) const {
    T eu = (T)L.cx[i] * ux + (T)L.cy[i] * uy;
    if (L.D == 3) eu = eu + (T)L.cz[i] * uz;
    if constexpr (std::is_same<T, float>::value) {
      const T t = std::fma(4.5f, eu, 3.0f);
      return std::fma(eu, t, usq);
    } else {
      return (((T)1 + (T)3 * eu) + (T)4.5 * (eu * eu)) - (T)1.5 * usq;
    }
  }
  T feq(int i, T rho, T ux, T uy, T uz, T usq) const { return (w[i] * rho) * qeq(i, ux, uy, uz, usq); }

  // Shifted equilibrium (FP32 only): f_eq,i - w_i = w_i (drho + rho P_i), P_i = eu (3 + 4.5 eu)
  // - 1.5 u.u; m15 = -1.5 u.u.
  T m15_term(T ux, T uy, T uz) const {
    T u2 = std::fma(uy, uy, ux * ux);
    if (L.D == 3) u2 = std::fma(uz, uz, u2);
    return (T)-1.5 * u2;
  }
  T peq(int i, T ux, T uy, T uz, T m15) const {  // P_i
    T P = m15;
    if (L.cx[i] != 0 || L.cy[i] != 0 || L.cz[i] != 0) {
      T eu = (T)L.cx[i] * ux + (T)L.cy[i] * uy;
      if (L.D == 3) eu = eu + (T)L.cz[i] * uz;
      P = std::fma(eu, std::fma((T)4.5, eu, (T)3), m15);
    }
    return P;
  }
  T geq(int i, T drho, T rho, T ux, T uy, T uz, T m15) const {
    return w[i] * std::fma(rho, peq(i, ux, uy, uz, m15), drho);
  }
  void bgk_s(const T *S, T drho, T rho, T ux, T uy, T uz, T *out) const {
    const T omega = (T)d.omega, m = m15_term(ux, uy, uz);
    for (int i = 0; i < L.Q; ++i) out[i] = std::fma(omega, geq(i, drho, rho, ux, uy, uz, m) - S[i], S[i]);
  }

  // BGK relaxation (cpp/d2q9_bgk_lid.cpp:153): omega feq + (1 - omega) f; FP32 f + omega (feq - f).
  void bgk(const T *S, T rho, T ux, T uy, T uz, T *out) const {
    const T omega = (T)d.omega, om1 = (T)1 - omega;
    const T usq = usq_term(ux, uy, uz);
    for (int i = 0; i < L.Q; ++i) {
      const T fe = feq(i, rho, ux, uy, uz, usq);
      if constexpr (std::is_same<T, float>::value) out[i] = std::fma(omega, fe - S[i], S[i]);
      else out[i] = omega * fe + om1 * S[i];
    }
  }

  // Equilibrium moments (rust/examples/d3q27_mrt_lid.rs:341-383 operation order).
  void meq_reference(T r, T ux, T uy, T uz, T *meq) const {
    const T cs2 = (T)1 / (T)3, cs4 = cs2 * cs2, cs6 = cs2 * cs2 * cs2;
    const T rvx = r * ux, rvy = r * uy, rvz = r * uz, rc2 = cs2 * r, rc4 = cs4 * r, rc6 = cs6 * r;
    const T vx2 = ux * ux, vy2 = uy * uy, vz2 = uz * uz, v2 = vx2 + vy2 + vz2;
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

  // m = M S; m* = m - s (m - meq); out = M^-1 m*  (rust/examples/d3q27_mrt_lid.rs:336,385-387).
  void mrt_collide(const T *S, T rho, T ux, T uy, T uz, T *out) const {
    const int Q = L.Q;
    T m[27], meq[27], mc[27];
    for (int k = 0; k < Q; ++k) {
      T acc = 0;
      for (int i = 0; i < Q; ++i) acc += S[i] * M[k][i];
      m[k] = acc;
    }
    if (Q == 27 && d.collision == MRT_RAW && d.meq_mode == 0) {
      meq_reference(rho, ux, uy, uz, meq);
    } else {  // projected: meq = M f_eq(rho, u)
      T feq[27];
      const T usq = usq_term(ux, uy, uz);
      for (int i = 0; i < Q; ++i) feq[i] = this->feq(i, rho, ux, uy, uz, usq);
      for (int k = 0; k < Q; ++k) {
        T acc = 0;
        for (int i = 0; i < Q; ++i) acc += feq[i] * M[k][i];
        meq[k] = acc;
      }
    }
    for (int k = 0; k < Q; ++k) mc[k] = m[k] - (m[k] - meq[k]) * s[k];
    for (int i = 0; i < Q; ++i) {
      T acc = 0;
      for (int k = 0; k < Q; ++k) acc += mc[k] * Minv[i][k];
      out[i] = acc;
    }
  }

  // The shifted forms: S holds g = f - w_i, drho = rho - 1; the reference table less its value at
  // rest, and the projected equilibrium moments as M g_eq.
  void meq_reference_s(T dr, T r, T ux, T uy, T uz, T *meq) const {
    const T cs2 = (T)1 / (T)3, cs4 = cs2 * cs2, cs6 = cs2 * cs2 * cs2;
    const T rvx = r * ux, rvy = r * uy, rvz = r * uz, rc2 = cs2 * r, rc4 = cs4 * r;
    const T vx2 = ux * ux, vy2 = uy * uy, vz2 = uz * uz, v2 = vx2 + vy2 + vz2;
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
    meq[26] = v2 + cs6 * dr;
  }
  void mrt_collide_s(const T *S, T drho, T rho, T ux, T uy, T uz, T *out) const {
    const int Q = L.Q;
    T m[27], meq[27], mc[27];
    for (int k = 0; k < Q; ++k) {
      T acc = 0;
      for (int i = 0; i < Q; ++i) acc += S[i] * M[k][i];
      m[k] = acc;
    }
    if (Q == 27 && d.collision == MRT_RAW && d.meq_mode == 0) {
      meq_reference_s(drho, rho, ux, uy, uz, meq);
    } else {
      T ge[27];
      const T m15 = m15_term(ux, uy, uz);
      for (int i = 0; i < Q; ++i) ge[i] = geq(i, drho, rho, ux, uy, uz, m15);
      for (int k = 0; k < Q; ++k) {
        T acc = 0;
        for (int i = 0; i < Q; ++i) acc += ge[i] * M[k][i];
        meq[k] = acc;
      }
    }
    for (int k = 0; k < Q; ++k) mc[k] = m[k] - (m[k] - meq[k]) * s[k];
    for (int i = 0; i < Q; ++i) {
      T acc = 0;
      for (int k = 0; k < Q; ++k) acc += mc[k] * Minv[i][k];
      out[i] = acc;
    }
  }
  void collide_s(const T *S, T drho, T rho, T ux, T uy, T uz, T *out) const {
    if (mrt) mrt_collide_s(S, drho, rho, ux, uy, uz, out);
    else bgk_s(S, drho, rho, ux, uy, uz, out);
  }

  void moment_sums(const T *S, T &r, T &jx, T &jy, T &jz) const {
    r = 0, jx = 0, jy = 0, jz = 0;
    for (int i = 0; i < L.Q; ++i) r += S[i];
    for (int i = 0; i < L.Q; ++i) jx += (T)L.cx[i] * S[i];
    for (int i = 0; i < L.Q; ++i) jy += (T)L.cy[i] * S[i];
    if (L.D == 3)
      for (int i = 0; i < L.Q; ++i) jz += (T)L.cz[i] * S[i];
  }
  // shifted: drho = sum g, rho = 1 + drho, u = (sum c g) * (1 / rho) (FP32 only)
  void moments_s(const T *S, T &drho, T &rho, T &ux, T &uy, T &uz) const {
    T r, jx, jy, jz;
    moment_sums(S, r, jx, jy, jz);
    drho = r;
    rho = (T)1 + r;
    const T inv = (T)1 / rho;
    ux = jx * inv; uy = jy * inv; uz = (L.D == 3) ? jz * inv : (T)0;
  }

  void moments(const T *S, T &rho, T &ux, T &uy, T &uz) const {
    T r, jx, jy, jz;
    moment_sums(S, r, jx, jy, jz);
    rho = r;
    if (sizeof(T) == 4) {  // FP32 as the kernels: the correctly rounded 1/rho, then products
      const T inv = (T)1 / r;
      ux = jx * inv; uy = jy * inv; uz = (L.D == 3) ? jz * inv : (T)0;
    } else {
      ux = jx / r; uy = jy / r; uz = (L.D == 3) ? jz / r : (T)0;
    }
  }

  // ---- REFERENCE step, multi-pass as the reference's ArrayFire program ----
  std::vector<T> S, B, RHO, DRHO, UX, UY, UZ;
  void ref_step() {
    const int Q = L.Q;
    const int64_t NQ = N * Q, nON = (int64_t)on.size();
    S.resize((size_t)NQ);
    B.resize((size_t)(nON * (Q - 1)));
    RHO.resize((size_t)N); UX.resize((size_t)N); UY.resize((size_t)N); UZ.resize((size_t)N);
    if (sh) DRHO.resize((size_t)N);
    // K1 gather: F_streamed = F(nb_index)
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < NQ; ++p) S[(size_t)p] = F[(size_t)nb[(size_t)p]];
    // K2 capture: BOUNCEDBACK = F_streamed(TO_REFLECT), k = s + nON*(j-1)
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nON * (Q - 1); ++k) {
      const int64_t sidx = k % nON, j = k / nON + 1;
      B[(size_t)k] = S[(size_t)(on[(size_t)sidx] + N * j)];
    }
    // K3 reductions + K5 overrides
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
      T Sn[27];
      for (int i = 0; i < Q; ++i) Sn[i] = S[(size_t)(n + N * i)];
      T rho, ux, uy, uz, drho = 0;
      const int y = (int)((n / nx) % ny), z = (int)(n / ((int64_t)nx * ny));
      if (sh) {
        moments_s(Sn, drho, rho, ux, uy, uz);
        overrides_s(n, y, z, drho, rho, ux, uy, uz);
        DRHO[(size_t)n] = drho;
      } else {
        moments(Sn, rho, ux, uy, uz);
        overrides(n, y, z, rho, ux, uy, uz);
      }
      RHO[(size_t)n] = rho; UX[(size_t)n] = ux; UY[(size_t)n] = uy; UZ[(size_t)n] = uz;
    }
    // K4/K7 collision at every node
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
      T Sn[27], out[27];
      for (int i = 0; i < Q; ++i) Sn[i] = S[(size_t)(n + N * i)];
      if (sh) collide_s(Sn, DRHO[(size_t)n], RHO[(size_t)n], UX[(size_t)n], UY[(size_t)n], UZ[(size_t)n], out);
      else if (mrt) mrt_collide(Sn, RHO[(size_t)n], UX[(size_t)n], UY[(size_t)n], UZ[(size_t)n], out);
      else bgk(Sn, RHO[(size_t)n], UX[(size_t)n], UY[(size_t)n], UZ[(size_t)n], out);
      for (int i = 0; i < Q; ++i) F[(size_t)(n + N * i)] = out[i];
    }
    // K6 scatter: F(REFLECTED) = BOUNCEDBACK
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nON * (Q - 1); ++k) {
      const int64_t sidx = k % nON, j = k / nON + 1;
      F[(size_t)(on[(size_t)sidx] + N * L.opp[j])] = B[(size_t)k];
    }
  }

  // ---- PHYSICAL step: node loop ----
  // Non-equilibrium bounce-back closure of the unknown populations (c_x == sgn) of an x-face
  // node with prescribed rho, u; transverse momentum corrected through
  // N_a = sum_{c_x=0} f c_a - (2/3) rho u_a (Zou-He in 2D).
  void nebb(T *f, int sgn, T rho, T ux, T uy, T uz) const {
    T ny_ = 0, nz_ = 0;
    for (int i = 0; i < L.Q; ++i)
      if (L.cx[i] == 0) { ny_ += (T)L.cy[i] * f[i]; nz_ += (T)L.cz[i] * f[i]; }
    ny_ = ny_ - (((T)2 / (T)3) * rho) * uy;
    nz_ = nz_ - (((T)2 / (T)3) * rho) * uz;
    for (int i = 0; i < L.Q; ++i)
      if (L.cx[i] == sgn) {
        T cu = (T)L.cx[i] * ux + (T)L.cy[i] * uy;
        if (L.D == 3) cu = cu + (T)L.cz[i] * uz;
        f[i] = (f[L.opp[i]] + (((T)6 * w[i]) * rho) * cu) -
               ((T)18 * w[i]) * ((T)L.cy[i] * ny_ + (T)L.cz[i] * nz_);
      }
  }

  // Pulled populations of fluid node (x,y,z) with the physical boundary rules applied.
  void phys_pull(const std::vector<T> &Fin, int x, int y, int z, T *f) const {
    const int64_t n = idx(x, y, z);
    const uint8_t t = type[n];
    const T rw = (T)d.rho_bc;
    for (int i = 0; i < L.Q; ++i) {
      const int64_t src = idx(wrap(x - L.cx[i], nx), wrap(y - L.cy[i], ny), wrap(z - L.cz[i], nz));
      if (type[src] & SOLID) {
        T v = Fin[(size_t)(n + N * L.opp[i])];
        if (type[src] & MOVING) {
          T cu = (T)L.cx[i] * (T)d.u_wall[0] + (T)L.cy[i] * (T)d.u_wall[1];
          if (L.D == 3) cu = cu + (T)L.cz[i] * (T)d.u_wall[2];
          v = v + (((T)6 * w[i]) * rw) * cu;
        }
        f[i] = v;
      } else {
        f[i] = Fin[(size_t)(src + N * i)];
      }
    }
    if ((t & INLET) && d.inlet_closure == 1) {
      // Extrapolated velocity inlet (Guo, Zheng & Shi, Chinese Physics 11 (2002) 366): every
      // population f_eq(rho_e, u_in) + (g - f_eq(rho_e, u_e)), g = the downstream (x + 1)
      // neighbour's populations as the pull reads them, rho_e and u_e their moments; f_eq(rho_bc,
      // u_in) when that neighbour is solid or absent.  The node then holds rho_e and u_in.  The
      // two equilibria share w_i rho_e, so their difference is formed inside the bracket:
      // g_i + w_i rho_e (qeq_i(u_in) - qeq_i(u_e)) (shifted: P_i in place of qeq_i).
      const T uin = inlet_u(y, z);
      T g[27], rn = (T)1, drn = 0, vx = 0, vy = 0, vz = 0;
      const bool nb = x + 1 < nx && !(type[n + 1] & SOLID);
      if (nb) {
        for (int i = 0; i < L.Q; ++i) g[i] = Fin[(size_t)(n + 1 + N * i)];
        if (sh) moments_s(g, drn, rn, vx, vy, vz);
        else moments(g, rn, vx, vy, vz);
      }
      if (sh) {  // the same with the equilibria as deviations from w_i
        const T mb = m15_term(uin, (T)0, (T)0), mg = m15_term(vx, vy, vz);
        for (int i = 0; i < L.Q; ++i)
          f[i] = nb ? std::fma(w[i] * rn, peq(i, uin, (T)0, (T)0, mb) - peq(i, vx, vy, vz, mg), g[i])
                    : geq(i, (T)(d.rho_bc - 1.0), rw, uin, (T)0, (T)0, mb);
        return;
      }
      const T us_b = usq_term(uin, (T)0, (T)0), us_g = usq_term(vx, vy, vz);
      for (int i = 0; i < L.Q; ++i) {
        if (!nb) {
          f[i] = feq(i, rw, uin, (T)0, (T)0, us_b);
          continue;
        }
        const T dq = qeq(i, uin, (T)0, (T)0, us_b) - qeq(i, vx, vy, vz, us_g);
        if constexpr (std::is_same<T, float>::value) f[i] = std::fma(w[i] * rn, dq, g[i]);
        else f[i] = g[i] + (w[i] * rn) * dq;
      }
    } else if (t & INLET) {  // velocity inlet (Zou-He), unknowns c_x = +1
      const T uin = inlet_u(y, z);
      T s0 = 0, sm = 0;
      for (int i = 0; i < L.Q; ++i) {
        if (L.cx[i] == 0) s0 += f[i];
        if (L.cx[i] == -1) sm += f[i];
      }
      // shifted: the weights of the c_x = 0 directions plus twice those of c_x = -1 sum to 1
      const T rho = sh ? (((T)1 + s0) + (T)2 * sm) / ((T)1 - uin) : (s0 + (T)2 * sm) / ((T)1 - uin);
      nebb(f, 1, rho, uin, (T)0, (T)0);
    } else if (t & OUTLET) {
      // Zero-gradient outlet by non-equilibrium extrapolation (Guo, Zheng & Shi, Chinese Physics
      // 11 (2002) 366): every population of the node is f_eq(rho_bc, u_nb) + (g - f_eq(rho_nb,
      // u_nb)), g = the upstream neighbour's populations as the pull reads them (its
      // post-collision state of the previous step).  The node then holds rho_bc and u_nb exactly.
      // f_eq is linear in rho: evaluated as g + (rho_bc - rho_nb) f_eq(1, u_nb), FP32 as one FMA.
      T g[27], rn = (T)1, drn = 0, vx = 0, vy = 0, vz = 0;
      const bool nb = x > 0 && !(type[n - 1] & SOLID);
      if (nb) {
        for (int i = 0; i < L.Q; ++i) g[i] = Fin[(size_t)(n - 1 + N * i)];
        if (sh) moments_s(g, drn, rn, vx, vy, vz);
        else moments(g, rn, vx, vy, vz);
      }
      const T usq = usq_term(vx, vy, vz);
      if (sh) {  // rho_bc - rho_nb between deviations; f_eq(1, u_nb) unshifted
        const T drbc = (T)(d.rho_bc - 1.0), dr = drbc - drn, mg = m15_term(vx, vy, vz);
        for (int i = 0; i < L.Q; ++i)
          f[i] = nb ? std::fma(dr, feq(i, (T)1, vx, vy, vz, usq), g[i]) : geq(i, drbc, rw, vx, vy, vz, mg);
        return;
      }
      const T dr = rw - rn;
      for (int i = 0; i < L.Q; ++i) {
        if (nb) {
          const T e = feq(i, (T)1, vx, vy, vz, usq);
          if constexpr (std::is_same<T, float>::value) f[i] = std::fma(dr, e, g[i]);
          else f[i] = g[i] + dr * e;
        } else {
          f[i] = feq(i, rw, vx, vy, vz, usq);
        }
      }
    }
  }

  std::vector<T> Fnext;
  std::vector<char> block_fluid;  // per aligned block of 64 node indices: holds a non-solid node
  void phys_step() {
    const int Q = L.Q;
    Fnext.resize(F.size());
    if (block_fluid.empty()) {
      block_fluid.assign((size_t)((N + 63) / 64), 0);
      for (int64_t n = 0; n < N; ++n)
        if (!(type[n] & SOLID)) block_fluid[(size_t)(n / 64)] = 1;
    }
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < N; ++n) {
      const int x = (int)(n % nx), y = (int)((n / nx) % ny), z = (int)(n / ((int64_t)nx * ny));
      if (type[n] & SOLID) {
        for (int i = 0; i < Q; ++i)
          Fnext[(size_t)(n + N * i)] =
              Q >= 19 && block_fluid[(size_t)(n / 64)] ? (sh ? (T)0 : w[i]) : F[(size_t)(n + N * i)];
        continue;
      }
      T f[27], out[27], rho, ux, uy, uz, drho;
      phys_pull(F, x, y, z, f);
      if (sh) {
        moments_s(f, drho, rho, ux, uy, uz);
        collide_s(f, drho, rho, ux, uy, uz, out);
      } else {
        moments(f, rho, ux, uy, uz);
        if (mrt) mrt_collide(f, rho, ux, uy, uz, out);
        else bgk(f, rho, ux, uy, uz, out);
      }
      for (int i = 0; i < Q; ++i) Fnext[(size_t)(n + N * i)] = out[i];
    }
    F.swap(Fnext);
  }

  // Shifted storages: round g to the storage and back, to nearest (initial states) or
  // stochastically with random bits r (steps).
  float round_g(int i, float g) const {
    (void)i;
    return d.storage == F16S ? h2f(f2h(g * 32768.0f)) * (1.0f / 32768.0f) : bf2f(f2bf(g));
  }
  float round_g_sr(float g, uint32_t r) const {
    return d.storage == F16S ? h2f(f2h_sr(g * 32768.0f, r)) * (1.0f / 32768.0f) : bf2f(f2bf_sr(g, r));
  }
  // after a step: every population through the storage with this step's random bits (values
  // a step did not change are representable and come back unchanged)
  void codec_step() {
    if constexpr (std::is_same<T, float>::value) {
      const int Q = L.Q;
      const uint32_t key = mix32(tstep + 0x9E3779B9u);
#pragma omp parallel for schedule(static)
      for (int64_t n = 0; n < N; ++n) {
        const uint32_t h = mix32((uint32_t)n ^ key);
        for (int i = 0; i < Q; ++i) {
          float &v = F[(size_t)(n + N * i)];
          v = (d.reserved & 1) ? round_g(i, v) : round_g_sr(v, rotl32(h, (7 * i) & 31));
        }
      }
    }
    ++tstep;
  }
  // an initial state given unshifted (F holds f): F becomes g = f - w_i rounded to nearest
  void encode_initial() {
    if constexpr (std::is_same<T, float>::value) {
      if (!sh) return;
      const int Q = L.Q;
      for (int64_t n = 0; n < N; ++n)
        for (int i = 0; i < Q; ++i) {
          float &v = F[(size_t)(n + N * i)];
          v = round_g(i, v - w[i]);
        }
    }
  }

  void step() {
    if (d.bc_mode == 0) ref_step();
    else phys_step();
    if (sh) codec_step();
  }

  // state in/out as doubles; shifted storages hold g and decode to g + w_i in double (exact)
  void load(const double *f) {
    F.resize((size_t)(N * L.Q));
    for (int i = 0; i < L.Q; ++i)
      for (int64_t n = 0; n < N; ++n) {
        const double v = f[n + N * i];
        if (sh) F[(size_t)(n + N * i)] = (T)round_g(i, (float)(v - (double)(float)L.w[i]));
        else F[(size_t)(n + N * i)] = (T)v;
      }
  }
  void store(double *f) const {
    for (int i = 0; i < L.Q; ++i)
      for (int64_t n = 0; n < N; ++n) {
        const T v = F[(size_t)(n + N * i)];
        f[n + N * i] = sh ? (double)v + (double)(float)L.w[i] : (double)v;
      }
  }
};

bool desc_ok(const oracle_desc *d) {
  if (!d) return false;
  if (d->lattice != 9 && d->lattice != 19 && d->lattice != 27) return false;
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return false;
  if (d->lattice == 9 && d->nz != 1) return false;
  if (d->collision != BGK && d->lattice == 9) return false;
  if (d->collision == MRT_RAW && d->lattice == 19) return false;  // orthogonal basis only
  return true;
}

template <class T>
int run_impl(const oracle_desc *d, const uint8_t *types, double *f, int64_t nsteps, int f32c, uint32_t step0) {
  Sim<T> sim(*d, types);
  sim.tstep = step0;
  sim.load(f);
  if (d->bc_mode == 0 && !sim.build_pull_map(f32c != 0)) return 3;  // rounded index out of range
  for (int64_t t = 0; t < nsteps; ++t) sim.step();
  sim.store(f);
  return 0;
}

template <class T>
int init_fields_impl(const oracle_desc *d, const uint8_t *types, const double *rho, const double *ux,
                     const double *uy, const double *uz, double *f) {
  Sim<T> sim(*d, types);
  const int Q = sim.L.Q;
  sim.F.assign((size_t)(sim.N * Q), (T)0);
  for (int64_t n = 0; n < sim.N; ++n) {
    const T r = rho ? (T)rho[n] : (T)1, vx = ux ? (T)ux[n] : (T)0, vy = uy ? (T)uy[n] : (T)0,
            vz = (uz && sim.L.D == 3) ? (T)uz[n] : (T)0;
    T out[27];
    const T usq = sim.usq_term(vx, vy, vz);
    for (int i = 0; i < Q; ++i) out[i] = sim.feq(i, r, vx, vy, vz, usq);
    for (int i = 0; i < Q; ++i) sim.F[(size_t)(n + sim.N * i)] = out[i];
  }
  sim.encode_initial();
  sim.store(f);
  return 0;
}

template <class T>
int init_reference_impl(const oracle_desc *d, const uint8_t *types, double rho0, double *f) {
  Sim<T> sim(*d, types);
  const int Q = sim.L.Q;
  sim.F.assign((size_t)(sim.N * Q), (T)0);
  for (int64_t n = 0; n < sim.N; ++n) {
    T out[27];
    if (sim.mrt && Q == 27) {  // D3Q19 MRT starts from f_eq(rho0, 0), as BGK
      T S[27];
      for (int i = 0; i < 27; ++i) S[i] = (T)rho0 / (T)27;
      sim.mrt_collide(S, (T)rho0, 0, 0, 0, out);
    } else {
      const T usq = sim.usq_term((T)0, (T)0, (T)0);
      for (int i = 0; i < Q; ++i) out[i] = sim.feq(i, (T)rho0, (T)0, (T)0, (T)0, usq);
    }
    for (int i = 0; i < Q; ++i) sim.F[(size_t)(n + sim.N * i)] = out[i];
  }
  sim.encode_initial();
  sim.store(f);
  return 0;
}

template <class T>
int fields_impl(const oracle_desc *d, const uint8_t *types, const double *f, double *rho, double *ux,
                double *uy, double *uz) {
  Sim<T> sim(*d, types);
  sim.load(f);
  const int Q = sim.L.Q;
  for (int z = 0; z < sim.nz; ++z)
    for (int y = 0; y < sim.ny; ++y)
      for (int x = 0; x < sim.nx; ++x) {
        const int64_t n = sim.idx(x, y, z);
        T S[27], r, vx, vy, vz, dr = -1;
        if (d->bc_mode == 0) {
          for (int i = 0; i < Q; ++i) {
            const int64_t src = sim.idx(Sim<T>::wrap(x - sim.L.cx[i], sim.nx), Sim<T>::wrap(y - sim.L.cy[i], sim.ny),
                                        Sim<T>::wrap(z - sim.L.cz[i], sim.nz));
            S[i] = sim.F[(size_t)(src + sim.N * i)];
          }
          if (sim.sh) {
            sim.moments_s(S, dr, r, vx, vy, vz);
            sim.overrides_s(n, y, z, dr, r, vx, vy, vz);
          } else {
            sim.moments(S, r, vx, vy, vz);
            sim.overrides(n, y, z, r, vx, vy, vz);
          }
        } else if (types[n] & SOLID) {
          r = 0; vx = 0; vy = 0; vz = 0;
        } else {
          sim.phys_pull(sim.F, x, y, z, S);
          if (sim.sh) sim.moments_s(S, dr, r, vx, vy, vz);
          else sim.moments(S, r, vx, vy, vz);
        }
        // shifted storages report rho = 1 + drho in FP64, as the kernels do
        if (rho) rho[n] = sim.sh ? 1.0 + (double)dr : (double)r;
        if (ux) ux[n] = (double)vx;
        if (uy) uy[n] = (double)vy;
        if (uz) uz[n] = (double)vz;
      }
  return 0;
}

template <class T>
double time_impl(const oracle_desc *d, const uint8_t *types, int64_t nsteps, int64_t warmup) {
  Sim<T> sim(*d, types);
  const int Q = sim.L.Q;
  sim.F.assign((size_t)(sim.N * Q), (T)0);
  for (int i = 0; i < Q; ++i)
    for (int64_t n = 0; n < sim.N; ++n) sim.F[(size_t)(n + sim.N * i)] = sim.sh ? (T)0 : sim.w[i];
  if (d->bc_mode == 0) sim.build_pull_map(false);
  for (int64_t t = 0; t < warmup; ++t) sim.step();  // warm-up (page in temporaries)
  const auto t0 = std::chrono::steady_clock::now();
  for (int64_t t = 0; t < nsteps; ++t) sim.step();
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

bool is_f64(const oracle_desc *d) { return d->storage == F64; }

// ---- two-component Shan-Chen D2Q9 ----------------------------------------------------------
// cpp/d2q9_shanchen.cpp:39-181 with its work-in-progress defects corrected: both fluids are pulled
// (streamed) before the collision (the reference collides its unstreamed f and never applies the
// streamed copy, :117-118,175-176), each fluid's interaction sum uses the other fluid's density
// (the reference builds both from rho1, :145-149), the sums restart from zero every step (the
// reference accumulates them across iterations, :84-89,145), and the common velocity weights the
// momenta j_s = sum_i c_i f_i (the reference divides them by rho_s first, :128-131):
//   u' = (omega1 j1 + omega2 j2) / (omega1 rho1 + omega2 rho2)                       (:133-135)
//   u_1 = u' - (G / omega1) sum_{i>=1} w_i rho2(x - c_i) c_i, u_2 likewise with rho1  (:141-157)
//   f_out = f - omega (f - f_eq(rho, u_s)) per fluid                                   (:161-172)
// Evaluated node by node in the kernels' operation order (FP32: contracted equilibrium, relaxation
// fma(omega, f_eq - f, f), u' through the correctly rounded reciprocal of its denominator).
typedef struct {
  int32_t nx, ny, storage, pad_;
  double omega1, omega2, g;
} oracle_sc_desc;

template <class T>
T sc_usq(T ux, T uy) {
  if constexpr (std::is_same<T, float>::value) return std::fma(-1.5f, std::fma(uy, uy, ux * ux), 1.0f);
  else return ux * ux + uy * uy;
}

template <class T>
T sc_feq(const Lattice &L, int i, T rho, T ux, T uy, T usq) {
  const T eu = (T)L.cx[i] * ux + (T)L.cy[i] * uy;
  if constexpr (std::is_same<T, float>::value) {
    const T t = std::fma(4.5f, eu, 3.0f);
    return ((T)L.w[i] * rho) * std::fma(eu, t, usq);
  } else {
    return ((T)L.w[i] * rho) * ((((T)1 + (T)3 * eu) + (T)4.5 * (eu * eu)) - (T)1.5 * usq);
  }
}

template <class T>
void sc_run_impl(const oracle_sc_desc *d, double *f, int64_t nsteps) {
  const Lattice L = make_lattice(9);
  const int nx = d->nx, ny = d->ny;
  const int64_t N = (int64_t)nx * ny;
  std::vector<T> F((size_t)(18 * N)), S((size_t)(18 * N)), rho((size_t)(2 * N));
  for (size_t k = 0; k < F.size(); ++k) F[k] = (T)f[k];
  const T w1 = (T)d->omega1, w2 = (T)d->omega2, gw1 = (T)(d->g / d->omega1), gw2 = (T)(d->g / d->omega2);
  auto wrap = [](int a, int n) { return ((a % n) + n) % n; };
  for (int64_t t = 0; t < nsteps; ++t) {
    // stream both fluids (pull), then the densities of the streamed populations
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t n = x + (int64_t)nx * y;
        for (int i = 0; i < 9; ++i) {
          const int64_t src = wrap(x - L.cx[i], nx) + (int64_t)nx * wrap(y - L.cy[i], ny);
          S[(size_t)(i * N + n)] = F[(size_t)(i * N + src)];
          S[(size_t)((9 + i) * N + n)] = F[(size_t)((9 + i) * N + src)];
        }
        T r1 = S[(size_t)n], r2 = S[(size_t)(9 * N + n)];
        for (int i = 1; i < 9; ++i) { r1 = r1 + S[(size_t)(i * N + n)]; r2 = r2 + S[(size_t)((9 + i) * N + n)]; }
        rho[(size_t)n] = r1;
        rho[(size_t)(N + n)] = r2;
      }
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t n = x + (int64_t)nx * y;
        const T *sf = &S[0], *sg = &S[(size_t)(9 * N)];
        const T r1 = rho[(size_t)n], r2 = rho[(size_t)(N + n)];
        T j1x = -0.0, j1y = -0.0, j2x = -0.0, j2y = -0.0;
        for (int i = 0; i < 9; ++i) {
          const T a = sf[(size_t)(i * N + n)], b = sg[(size_t)(i * N + n)];
          if (L.cx[i] == 1) { j1x = j1x + a; j2x = j2x + b; }
          if (L.cx[i] == -1) { j1x = j1x - a; j2x = j2x - b; }
          if (L.cy[i] == 1) { j1y = j1y + a; j2y = j2y + b; }
          if (L.cy[i] == -1) { j1y = j1y - a; j2y = j2y - b; }
        }
        const T den = w1 * r1 + w2 * r2;
        T upx, upy;
        if constexpr (std::is_same<T, float>::value) {
          const T inv = (T)1 / den;
          upx = (w1 * j1x + w2 * j2x) * inv;
          upy = (w1 * j1y + w2 * j2y) * inv;
        } else {
          upx = (w1 * j1x + w2 * j2x) / den;
          upy = (w1 * j1y + w2 * j2y) / den;
        }
        T c[2][2];  // [fluid][axis]: sum_{i>=1} w_i rho_fluid(x - c_i) c_i
        for (int s = 0; s < 2; ++s) {
          T ax = -0.0, ay = -0.0;
          for (int i = 1; i < 9; ++i) {
            const int64_t src = wrap(x - L.cx[i], nx) + (int64_t)nx * wrap(y - L.cy[i], ny);
            const T v = (T)L.w[i] * rho[(size_t)(s * N + src)];
            if (L.cx[i] == 1) ax = ax + v;
            if (L.cx[i] == -1) ax = ax - v;
            if (L.cy[i] == 1) ay = ay + v;
            if (L.cy[i] == -1) ay = ay - v;
          }
          c[s][0] = ax;
          c[s][1] = ay;
        }
        const T u1x = upx - gw1 * c[1][0], u1y = upy - gw1 * c[1][1];
        const T u2x = upx - gw2 * c[0][0], u2y = upy - gw2 * c[0][1];
        const T a1 = sc_usq(u1x, u1y), a2 = sc_usq(u2x, u2y);
        for (int i = 0; i < 9; ++i) {
          const T fi = sf[(size_t)(i * N + n)], gi = sg[(size_t)(i * N + n)];
          const T e1 = sc_feq(L, i, r1, u1x, u1y, a1), e2 = sc_feq(L, i, r2, u2x, u2y, a2);
          if constexpr (std::is_same<T, float>::value) {
            F[(size_t)(i * N + n)] = std::fma(w1, e1 - fi, fi);
            F[(size_t)((9 + i) * N + n)] = std::fma(w2, e2 - gi, gi);
          } else {
            F[(size_t)(i * N + n)] = w1 * e1 + ((T)1 - w1) * fi;
            F[(size_t)((9 + i) * N + n)] = w2 * e2 + ((T)1 - w2) * gi;
          }
        }
      }
  }
  for (size_t k = 0; k < F.size(); ++k) f[k] = (double)F[k];
}

}  // namespace

extern "C" {

// 18 N populations (fluid 1 then fluid 2, population i of node n at n + N i), advanced nsteps.
int oracle_sc_run(const oracle_sc_desc *d, double *f, int64_t nsteps) {
  if (!d || d->nx < 1 || d->ny < 1) return 1;
  if (d->storage == 0) sc_run_impl<double>(d, f, nsteps);
  else if (d->storage == 1) sc_run_impl<float>(d, f, nsteps);
  else return 1;
  return 0;
}


int oracle_abi(void) { return 1; }

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Lattice tables: c (3*Q ints, x then y then z rows), w (Q), opp (Q).
int oracle_tables(int Q, int32_t *c, double *w, int32_t *opp) {
  if (Q != 9 && Q != 19 && Q != 27) return 1;
  Lattice L = make_lattice(Q);
  for (int i = 0; i < Q; ++i) {
    c[i] = L.cx[i]; c[Q + i] = L.cy[i]; c[2 * Q + i] = L.cz[i];
    w[i] = L.w[i]; opp[i] = L.opp[i];
  }
  return 0;
}

// basis 1 = raw monomial (reference), 2 = Gram-Schmidt: row-major 27x27.  basis 3 = the D3Q19
// orthogonal basis: row-major 19x19.
int oracle_mrt_matrices(int basis, double *M, double *Minv) {
  const int n = basis == 3 ? 19 : 27;
  Lattice L = make_lattice(n);
  double A[27][27], Ai[27][27];
  if (basis == 1) raw_moment_matrix(L, A);
  else if (basis == 2) gs_moment_matrix(L, A);
  else if (basis == 3) d3q19_moment_matrix(L, A);
  else return 1;
  if (!invert27(A, Ai, n)) return 2;
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < n; ++i) { M[k * n + i] = A[k][i]; Minv[k * n + i] = Ai[k][i]; }
  return 0;
}

int oracle_init_fields(const oracle_desc *d, const uint8_t *types, const double *rho, const double *ux,
                       const double *uy, const double *uz, double *f) {
  if (!desc_ok(d)) return 1;
  return is_f64(d) ? init_fields_impl<double>(d, types, rho, ux, uy, uz, f)
                   : init_fields_impl<float>(d, types, rho, ux, uy, uz, f);
}

int oracle_init_reference(const oracle_desc *d, const uint8_t *types, double rho0, double *f) {
  if (!desc_ok(d)) return 1;
  return is_f64(d) ? init_reference_impl<double>(d, types, rho0, f)
                   : init_reference_impl<float>(d, types, rho0, f);
}

// step0: the step count the run starts from (the shifted storages' random bits depend on it; a
// lattice's count restarts at 0 with every initial state).
int oracle_run_from(const oracle_desc *d, const uint8_t *types, double *f, int64_t nsteps, int f32_index_compat,
                    uint32_t step0) {
  if (!desc_ok(d)) return 1;
  return is_f64(d) ? run_impl<double>(d, types, f, nsteps, f32_index_compat, step0)
                   : run_impl<float>(d, types, f, nsteps, f32_index_compat, step0);
}
int oracle_run(const oracle_desc *d, const uint8_t *types, double *f, int64_t nsteps, int f32_index_compat) {
  return oracle_run_from(d, types, f, nsteps, f32_index_compat, 0);
}

int oracle_fields(const oracle_desc *d, const uint8_t *types, const double *f, double *rho, double *ux,
                  double *uy, double *uz) {
  if (!desc_ok(d)) return 1;
  return is_f64(d) ? fields_impl<double>(d, types, f, rho, ux, uy, uz)
                   : fields_impl<float>(d, types, f, rho, ux, uy, uz);
}

// Seconds for nsteps steps from a rest state (CPU baseline timing).
// Seconds for nsteps steps after `warmup` untimed ones (populations start at w_i).
double oracle_time_steps(const oracle_desc *d, const uint8_t *types, int64_t nsteps, int64_t warmup) {
  if (!desc_ok(d) || warmup < 0) return -1.0;
  return is_f64(d) ? time_impl<double>(d, types, nsteps, warmup) : time_impl<float>(d, types, nsteps, warmup);
}

// Codecs exposed for tests (checked against numpy's float16 conversion).
uint16_t oracle_f2h(float f) { return f2h(f); }
uint16_t oracle_f2h_sr(float f, uint32_t r) { return f2h_sr(f, r); }
uint16_t oracle_f2bf_sr(float f, uint32_t r) { return f2bf_sr(f, r); }
uint32_t oracle_mix32(uint32_t x) { return mix32(x); }
float oracle_h2f(uint16_t h) { return h2f(h); }
uint16_t oracle_f2bf(float f) { return f2bf(f); }
float oracle_bf2f(uint16_t b) { return bf2f(b); }

}  // extern "C"
