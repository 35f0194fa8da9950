// d3q27_mrt_lid.cpp — D3Q27 MRT lid-driven cavity through the C ABI.
//
// The reference benchmark program rust/examples/d3q27_mrt_lid.rs:62-465: raw monomial moment
// matrix, the reference relaxation rates (:111-150), lid velocity 0.1, f = rho0/27 plus one
// collision as the initial state, sizes looped as the reference's main (:455-465).  Writes the
// reference's CSV format (<gpu>_d3q27_mrt_lid_mlups_<n>.csv, Iterations,MLUPS;
// rust/examples/d3q27_mrt_lid.rs:50-60; <gpu> from LBM_GPU_NAME, default MI355X) with the 1e-6
// MLUPS factor and a column with the reference's 10e-6 factor, one row per timed chunk, and the
// final fields as a VTK file (d3q27_mrt_lid_<n>.vtk) when viz is given.
//
//   d3q27_mrt_lid [device] [steps] [engine: auto|scalar|mfma] [storage: f32|f16s|bf16s|f64] [sizes...] [viz]
//
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_capi.h"
#include "lbm_output.hpp"

static int die(const char *what) {
  std::fprintf(stderr, "%s: %s\n", what, lbm_last_error());
  return 1;
}

int main(int argc, char **argv) {
  const int device = argc > 1 ? std::atoi(argv[1]) : 0;
  const long steps = argc > 2 ? std::atol(argv[2]) : 5000;
  const std::string eng = argc > 3 ? argv[3] : "auto";
  const std::string sto = argc > 4 ? argv[4] : "f32";
  std::vector<int> sizes;
  bool viz = false;
  for (int k = 5; k < argc; ++k) {
    if (std::strcmp(argv[k], "viz") == 0) viz = true;
    else sizes.push_back(std::atoi(argv[k]));
  }
  const char *gpu = std::getenv("LBM_GPU_NAME") ? std::getenv("LBM_GPU_NAME") : "MI355X";
  if (sizes.empty()) sizes = {16, 32, 64, 128, 256};
  if (lbm_set_device(device) != LBM_OK) return die("lbm_set_device");

  const double rates[27] = {1.3, 1.3, 1.3, 1.3, 1.1, 1.1, 1.1, 0.9, 1.1, 1.1, 1.2, 1.2, 1.2, 1.2,
                            1.2, 1.2, 0.9, 1.2, 1.2, 1.2, 1.2, 1.2, 1.2, 1.1, 1.1, 1.1, 1.1};
  for (int n : sizes) {
    std::vector<uint8_t> type((size_t)n * n * n);
    for (int z = 0; z < n; ++z)
      for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {  // rust/examples/d3q27_mrt_lid.rs:202-212,323-330
          const bool fluid = x >= 1 && x <= n - 2 && y >= 1 && y <= n - 2 && z >= 1;
          type[(size_t)x + (size_t)n * (y + (size_t)n * z)] =
              fluid ? (z == n - 1 ? LBM_NODE_MOVING : LBM_NODE_FLUID) : LBM_NODE_SOLID;
        }
    lbm_desc d;
    std::memset(&d, 0, sizeof d);
    d.lattice = LBM_D3Q27;
    d.nx = d.ny = d.nz = n;
    d.collision = LBM_MRT_RAW;
    d.storage = sto == "f16s" ? LBM_F16S : sto == "bf16s" ? LBM_BF16S : sto == "f64" ? LBM_F64 : LBM_F32;
    d.bc_mode = LBM_BC_REFERENCE;
    d.mrt_engine = eng == "scalar" ? LBM_MRT_SCALAR : eng == "mfma" ? LBM_MRT_MFMA_F32 : LBM_MRT_AUTO;
    d.meq_mode = LBM_MEQ_REFERENCE;
    for (int k = 0; k < 27; ++k) d.s[k] = rates[k];
    d.u_wall[0] = 0.1;
    d.rho_bc = 1.0;
    lbm_handle h;
    if (lbm_create(&d, type.data(), &h) != LBM_OK) return die("lbm_create");
    if (lbm_init_reference(h, 1.0) != LBM_OK) return die("lbm_init_reference");
    lbm_out::MlupsCsv csv;
    csv.open("", gpu, "d3q27_mrt_lid", std::to_string(n));
    const long chunk = steps >= 50 ? steps / 50 : 1;
    double ms_total = 0;
    const double N = (double)n * n * n;
    for (long done = 0; done < steps;) {
      const long k = steps - done < chunk ? steps - done : chunk;
      double ms = 0;
      if (lbm_step_timed(h, k, &ms) != LBM_OK) return die("lbm_step_timed");
      ms_total += ms;
      done += k;
      csv.row(done, N, ms_total * 1e-3);
    }
    double mass = 0, ke = 0;
    int64_t bad = 0;
    lbm_monitors(h, &mass, &ke, &bad);
    std::printf("size %d: %ld steps, %.3f s, %.1f MLUPS, mass %.6f, kinetic energy %.4e%s -> %s\n", n,
                steps, ms_total * 1e-3, N * steps / (ms_total * 1e3), mass, ke, bad ? " NON-FINITE" : "",
                csv.path.c_str());
    if (viz) {
      const size_t nn = (size_t)n * n * n;
      std::vector<double> fld(4 * nn);
      if (lbm_get_fields(h, fld.data(), fld.data() + nn, fld.data() + 2 * nn, fld.data() + 3 * nn) != LBM_OK)
        return die("lbm_get_fields");
      const std::string vtk = "d3q27_mrt_lid_" + std::to_string(n) + ".vtk";
      if (!lbm_out::write_vtk(vtk.c_str(), n, n, n, fld.data(), fld.data() + nn, fld.data() + 2 * nn, fld.data() + 3 * nn))
        return die("write_vtk");
    }
    lbm_destroy(h);
  }
  return 0;
}
