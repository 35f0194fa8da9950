// lbm_internal.hpp — interface between the C ABI host code and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "lbm_fastdiv.hpp"
#include "lbm_lattice.hpp"

namespace lbm {

// Everything one fused step launch needs, passed by value as the kernel argument.
struct StepParams {
  const void *fin;          // populations read (post-collision of the previous step)
  void *fout;               // populations written
  const uint8_t *type;      // node types (with NODE_SLOW in PHYSICAL mode)
  // PHYSICAL lattices: per node, bit i - 1 set when population i (1 <= i < Q) is pulled from a
  // solid node (the wall link mask); null otherwise.  Read only by wall-neighbour nodes, which
  // reload those links from their own populations.
  const uint32_t *links;
  int32_t nx, ny, nz;
  int32_t gx, gy, gz;       // tile grid: 64 x 4 x 1 nodes per tile (flat lattices: ntiles x 1 x 1)
  int32_t flat;             // tiles of 256 consecutive node indices (rows not a multiple of 64 wide)
  uint32_t div_nx[3], div_ny[3];  // exact division by nx and by ny (FastDiv: multiplier, shifts)
  int32_t ntiles;
  int32_t inlet_profile;
  int32_t inlet_closure;    // 0 Zou-He (non-equilibrium bounce-back), 1 extrapolated (Guo et al.)
  int32_t meq_fix26;        // MFMA MRT: raw basis with the reference's equilibrium table (row 26)
  int32_t stream_order;     // MFMA MRT: tiles in stream_tile order (lattices too large for the caches)
  int32_t meq_mode;
  int32_t wrap_cut;         // bit 0 / 1: nothing crosses the periodic x / y wrap (PHYSICAL, 2D)
  int64_t N;                // nodes
  int64_t stride;           // elements between consecutive population arrays (>= N)
  const float *mrt_tab;     // MFMA MRT: per-lane A operands, 64 x MFMA_TAB floats
  // Step counter of the stochastic rounding (shifted storages): a pair of device words, one per
  // buffer orientation.  A launch reads *step_in and writes the count after it to *step_out, the
  // word the next launch (the other orientation) reads.
  const uint32_t *step_in;
  uint32_t *step_out;
  double omega;
  double u_wall[3];
  double u_in;
  double rho_bc;
  double s[27];
  float s32[27];            // FP32 MRT: the rates s rounded to FP32
};

struct Config {
  int Q, storage, bc, collision, engine;
  int nt = 0;  // 3D lattices too large for the caches: non-temporal population stores (single
               // lattices only; batches use ordinary stores)
};

// Tile shape of the node-per-thread kernels: one wavefront per 64-node row segment,
// four rows per 256-thread workgroup.
constexpr int TILE_X = 64, TILE_Y = 4, BLOCK_THREADS = 256;
// Per-lane constants of the MFMA MRT kernels (at most 40 floats used; see build_mfma_table).
constexpr int MFMA_TAB = 48;

// Population arrays are laid out with a stride of round_up(N * s, 4 KiB) + 4 KiB + 256 B bytes, so
// that arrays of power-of-two sized lattices do not start at power-of-two distances (which would
// send the 2Q streams of a node update to the same memory channels).  Arrays under 1 MiB are not
// padded.  LBM_POP_PAD=0 in the environment disables the padding.
int64_t population_stride(int64_t N, int esize);

// Lattices whose rows are not a multiple of 64 nodes wide are cut into tiles of 256 consecutive
// node indices ("flat" tiles): every lane of a wavefront then holds a node, where 64 x 4 tiles
// leave the lanes past each row's end idle (a 193-wide row takes 4 wavefronts for 193 nodes; the
// kernels cost per wavefront, so D3Q27 BGK F16S ran 32.4k MLUPS at 193 x 192 x 192 against 43.4k
// at 192).
inline bool flat_tiles(int nx) { return nx % TILE_X != 0; }
inline void tile_grid(int nx, int ny, int nz, int32_t &gx, int32_t &gy, int32_t &gz, int32_t &nt) {
  if (flat_tiles(nx)) {
    const int64_t N = (int64_t)nx * ny * nz;
    nt = (int32_t)((N + BLOCK_THREADS - 1) / BLOCK_THREADS);
    gx = nt;
    gy = gz = 1;
    return;
  }
  gx = (nx + TILE_X - 1) / TILE_X;
  gy = (ny + TILE_Y - 1) / TILE_Y;
  gz = nz;
  nt = gx * gy * gz;
}

hipError_t launch_step(const Config &c, const StepParams &p, hipStream_t s);
hipError_t launch_step_scalar(const Config &c, const StepParams &p, hipStream_t s);
hipError_t launch_step_scalar_nt(const Config &c, const StepParams &p, hipStream_t s);  // Q = 19, 27
hipError_t launch_step_mrt_mfma(const Config &c, const StepParams &p, hipStream_t s);
// Batched launches: member m of a batch owns tiles [prefix[m], prefix[m+1]) of the batch's tile
// range.  Everything a launch needs about the members is read from a device-resident table whose
// header locates the rest, so a launch captured in a graph stays valid when the host rewrites the
// table (members joining or leaving) between replays.
struct BatchTable {
  int32_t nmembers;
  int32_t total_tiles;
  int32_t total_multi;           // multi-step tiles (D2Q9 BGK groups), 0 otherwise
  int32_t stream;                // members too large for the caches together: tiles in stream order
  const int32_t *prefix;         // nmembers + 1 tile offsets
  const int32_t *prefix_multi;   // nmembers + 1 multi-step tile offsets (or null)
  const StepParams *members[2];  // parameter blocks of both buffer orientations
};
struct BatchArgs {
  const BatchTable *tab;
  int32_t phase;     // which orientation this launch steps
  int32_t one_tile;  // the grid has one workgroup per tile (selects the loop-free kernel)
  int32_t nt;        // 3D scalar groups too large for the caches: non-temporal population stores
};
hipError_t launch_step_batch(const Config &c, const BatchArgs &b, int grid, hipStream_t s);
hipError_t launch_step_batch_scalar(const Config &c, const BatchArgs &b, int grid, hipStream_t s);
hipError_t launch_step_batch_scalar_nt(const Config &c, const BatchArgs &b, int grid, hipStream_t s);  // Q = 19, 27
hipError_t launch_step_batch_mrt_mfma(const Config &c, const BatchArgs &b, int grid, hipStream_t s);
// Several D2Q9 BGK steps per launch from an LDS tile (2D lattices, FP32-arithmetic storages).
// multi_steps_per_launch = the steps one launch of the given shape advances, 0 if c or the shape
// does not qualify; the launch reads p.fin and leaves the state K steps later in p.fout.
int multi_variant_count();
constexpr int MULTI_DEFAULT = 9;  // the default launch shape (64 x 16 tiles, 5 steps), also batched
constexpr int MULTI_DEFAULT_F64 = 3;  // FP64: 64 x 16 tiles, 4 steps, one LDS copy (133 KB)
int multi_steps_per_launch(const Config &c, int variant);
hipError_t launch_step_multi(const Config &c, int variant, const StepParams &p, hipStream_t s);
// Multi-step tiles of a nx x ny lattice in the given shape; batched multi-step launches (shape
// MULTI_DEFAULT, one workgroup per multi-step tile of the group by default).
int multi_tiles(int variant, int nx, int ny);
hipError_t launch_step_multi_batch(const Config &c, const BatchArgs &b, int grid, hipStream_t s);
// True when the step of this configuration runs on the MFMA MRT kernel.
// Engine AUTO resolves to MRT_ENGINE_AUTO_FP32 for FP32-arithmetic storages (the sparse integer
// scalar transform: faster on gfx950 than either MFMA engine), scalar for FP64.
enum { MRT_ENGINE_AUTO = 0, MRT_ENGINE_SCALAR = 1, MRT_ENGINE_MFMA = 2, MRT_ENGINE_BF16X3 = 3, MRT_ENGINE_FUSED = 4,
       MRT_ENGINE_AUTO_FP32 = MRT_ENGINE_SCALAR };
inline bool uses_mfma(const Config &c) {
  return c.Q == 27 && c.collision != CO_BGK && c.storage != ST_F64 &&
         (c.engine == MRT_ENGINE_MFMA || c.engine == MRT_ENGINE_FUSED || c.engine == MRT_ENGINE_BF16X3);
}
// Fills 64 * MFMA_TAB floats of per-lane MFMA operands for a D3Q27 MRT lattice and the given
// engine (MRT_ENGINE_MFMA, MRT_ENGINE_FUSED or MRT_ENGINE_BF16X3).
void build_mfma_table(int collision, const double *s, int engine, float *tab);
// Reference-style fields (rho, u after the boundary overrides) of the state p.fin.
hipError_t launch_fields(const Config &c, const StepParams &p, double *rho, double *ux, double *uy,
                         double *uz, hipStream_t s);
// out = in & 0x1f, plus NODE_SLOW on non-solid nodes that are inlet/outlet nodes or pull from a
// solid node, NODE_WALLNB on the latter, NODE_WALLMOV on those with a moving solid source (PHYSICAL
// mode); links (if not null): the wall link masks (StepParams::links).
hipError_t launch_prepare_types(int Q, int bc_mode, int nx, int ny, int nz, const uint8_t *in, uint8_t *out,
                                uint32_t *links, hipStream_t s);
// A wavefront's 64 consecutive stores of one population array fill whole 128-byte lines: its
// nodes are 64 consecutive indices (flat tiles, or rows a multiple of 64 nodes wide) and a row of
// whole lines starts each 64-node run (non-temporal stores only then).
inline bool rows_fill_lines(int nx, int esize) {
  return nx % 64 != 0 /* flat tiles: 64 consecutive indices, 64 * esize >= 128 bytes */ ||
         ((int64_t)nx * esize) % 128 == 0;
}
// The node-type block of a lattice: N type bytes, then (wall_links_kept) the 32-bit wall link masks
// from byte offset wall_links_offset(N).
inline bool wall_links_kept(int Q, int bc_mode) { return (void)Q, bc_mode == BC_PHYS; }
inline size_t wall_links_offset(int64_t N) { return ((size_t)N + 255) & ~(size_t)255; }
inline size_t types_block_bytes(int Q, int bc_mode, int64_t N) {
  return wall_links_kept(Q, bc_mode) ? wall_links_offset(N) + 4 * (size_t)N : (size_t)N;
}
// out[a * n + k] = fields[a * N + nodes[k]] for a = rho, ux, uy, uz.
hipError_t launch_gather_fields(const double *fields, int64_t N, const int64_t *nodes, int n, double *out,
                                hipStream_t s);
// f = f_eq(rho, u) into both buffers (fin and fout); field pointers may be null.
hipError_t launch_init_fields(const Config &c, const StepParams &p, const double *rho, const double *ux,
                              const double *uy, const double *uz, hipStream_t s);
// f = f_eq(rho0, u0 + SplitMix64 noise at non-solid nodes), generated on the device.  Both buffers.
hipError_t launch_init_equilibrium(const Config &c, const StepParams &p, double rho0, const double *u0,
                                   uint64_t seed, double amp, hipStream_t s);
// MRT: f = rho0/27 then one collision at rho0, u = 0; BGK: f_eq(rho0, 0). Both buffers.
hipError_t launch_init_reference(const Config &c, const StepParams &p, double rho0, hipStream_t s);
// Momentum-exchange force on the solid nodes inside box[6] = {x0, x1, y0, y1, z0, z1} (half-open)
// into out[3]; partials as for launch_monitors.
hipError_t launch_force(const Config &c, const StepParams &p, const int *box, double *partials, double *out,
                        hipStream_t s);
// Adds to *count the x in [lo, hi) (float bit patterns) whose recip_rn(x) differs from 1.0f / x.
hipError_t launch_check_reciprocal(uint32_t lo, uint32_t hi, unsigned long long *count, hipStream_t s);
// Per-tile partial sums (mass, kinetic energy, nonfinite) then a final single-block reduction
// into out[3].  partials must hold 3 * p.ntiles doubles.
hipError_t launch_monitors(const Config &c, const StepParams &p, double *partials, double *out,
                           hipStream_t s);

}  // namespace lbm
