// lbm_internal.hpp — interface between the C ABI host code and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "lbm_lattice.hpp"

namespace lbm {

// Everything one fused step launch needs, passed by value as the kernel argument.
struct StepParams {
  const void *fin;          // populations read (post-collision of the previous step)
  void *fout;               // populations written
  const uint8_t *type;      // node types (with NODE_SLOW in PHYSICAL mode)
  int32_t nx, ny, nz;
  int32_t gx, gy, gz;       // tile grid: 64 x 4 x 1 nodes per tile
  int32_t ntiles;
  int32_t inlet_profile;
  int32_t meq_mode;
  int32_t pad_;
  int64_t N;
  double omega;
  double u_wall[3];
  double u_in;
  double rho_bc;
  double s[27];
};

struct Config {
  int Q, storage, bc, collision, engine;
};

// Tile shape of the node-per-thread kernels: one wavefront per 64-node row segment,
// four rows per 256-thread workgroup.
constexpr int TILE_X = 64, TILE_Y = 4, BLOCK_THREADS = 256;

inline void tile_grid(int nx, int ny, int nz, int32_t &gx, int32_t &gy, int32_t &gz, int32_t &nt) {
  gx = (nx + TILE_X - 1) / TILE_X;
  gy = (ny + TILE_Y - 1) / TILE_Y;
  gz = nz;
  nt = gx * gy * gz;
}

hipError_t launch_step(const Config &c, const StepParams &p, hipStream_t s);
// Reference-style fields (rho, u after the boundary overrides) of the state p.fin.
hipError_t launch_fields(const Config &c, const StepParams &p, double *rho, double *ux, double *uy,
                         double *uz, hipStream_t s);
// f = f_eq(rho, u) into both buffers (fin and fout); field pointers may be null.
hipError_t launch_init_fields(const Config &c, const StepParams &p, const double *rho, const double *ux,
                              const double *uy, const double *uz, hipStream_t s);
// MRT: f = rho0/27 then one collision at rho0, u = 0; BGK: f_eq(rho0, 0). Both buffers.
hipError_t launch_init_reference(const Config &c, const StepParams &p, double rho0, hipStream_t s);
// Per-tile partial sums (mass, kinetic energy, nonfinite) then a final single-block reduction
// into out[3].  partials must hold 3 * p.ntiles doubles.
hipError_t launch_monitors(const Config &c, const StepParams &p, double *partials, double *out,
                           hipStream_t s);

}  // namespace lbm
