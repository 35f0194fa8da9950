// d2q9_bgk_lid.cpp — D2Q9 BGK lid-driven cavity through the C ABI.
//
// Same inputs and outputs as the reference program cpp/d2q9_bgk_lid.cpp:30-207 (lid velocity
// 0.05, rho0 = 1, nu = u_lid * nx / Re, omega = 1/(3 nu + 1/2), BOUND mask with the open lid,
// equilibrium start; rho/ux/uy out), without the Forge window: every viz_every steps it writes
// the window's content as files (lid_<iter>.ppm: normalised |u| image; lid_<iter>_vectors.csv:
// the velocity sampled every n/30 nodes; lid_<iter>.vtk: rho and u), and the MLUPS history in the
// reference's CSV format (<gpu>_d2q9_bgk_lid_mlups_<n>.csv, rust/examples/d2q9_bgk_lid.rs:25-35;
// <gpu> from LBM_GPU_NAME, default MI355X), one row per 100 steps.  The final fields can also be
// written as raw little-endian doubles.  MLUPS uses the 1e-6 factor (the reference's 10e-6 reports
// 10x; the CSV carries both).
//
//   d2q9_bgk_lid [device] [n] [steps] [re] [out.bin|-] [mode: reference|physical] [viz_every]
#include <chrono>
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
  const int n = argc > 2 ? std::atoi(argv[2]) : 128;
  const long steps = argc > 3 ? std::atol(argv[3]) : 10000;
  const double re = argc > 4 ? std::atof(argv[4]) : 100.0;
  const char *out = argc > 5 && std::strcmp(argv[5], "-") != 0 ? argv[5] : nullptr;
  const bool physical = argc > 6 && std::strcmp(argv[6], "physical") == 0;
  const long viz_every = argc > 7 ? std::atol(argv[7]) : 0;
  const char *gpu = std::getenv("LBM_GPU_NAME") ? std::getenv("LBM_GPU_NAME") : "MI355X";
  if (lbm_set_device(device) != LBM_OK) return die("lbm_set_device");

  const double u_lid = 0.05, nu = u_lid * n / re, omega = 1.0 / (3.0 * nu + 0.5);
  std::printf("D2Q9 BGK lid cavity %dx%d, Re %.1f, nu %.6f, omega %.7f, %s boundaries\n", n, n, re, nu,
              omega, physical ? "physical" : "reference");
  std::vector<uint8_t> type((size_t)n * n, LBM_NODE_FLUID);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      uint8_t &t = type[(size_t)x + (size_t)n * y];
      if (physical) {  // walls all round, moving top row
        if (x == 0 || x == n - 1 || y == 0 || y == n - 1) t = LBM_NODE_SOLID;
        if (y == n - 1) t |= LBM_NODE_MOVING;
      } else {  // cpp/d2q9_bgk_lid.cpp:90-91,141-142
        const bool fluid = x >= 1 && x <= n - 2 && y >= 1;
        t = fluid ? (y == n - 1 ? LBM_NODE_MOVING : LBM_NODE_FLUID) : LBM_NODE_SOLID;
      }
    }
  lbm_desc d;
  std::memset(&d, 0, sizeof d);
  d.lattice = LBM_D2Q9;
  d.nx = n;
  d.ny = n;
  d.nz = 1;
  d.collision = LBM_BGK;
  d.storage = LBM_F32;
  d.bc_mode = physical ? LBM_BC_PHYSICAL : LBM_BC_REFERENCE;
  d.omega = omega;
  d.u_wall[0] = u_lid;
  d.rho_bc = 1.0;
  lbm_handle h;
  if (lbm_create(&d, type.data(), &h) != LBM_OK) return die("lbm_create");
  if (lbm_init_reference(h, 1.0) != LBM_OK) return die("lbm_init_reference");

  const size_t N = (size_t)n * n;
  lbm_out::MlupsCsv csv;
  csv.open("", gpu, "d2q9_bgk_lid", std::to_string(n));
  std::vector<double> fld(4 * N);
  auto snapshot = [&](long iter) -> bool {  // the reference's window, as files
    if (lbm_get_fields(h, fld.data(), fld.data() + N, fld.data() + 2 * N, nullptr) != LBM_OK) return false;
    const std::string tag = "lid_" + std::to_string(iter);
    return lbm_out::write_speed_ppm((tag + ".ppm").c_str(), n, n, fld.data() + N, fld.data() + 2 * N, type.data()) &&
           lbm_out::write_vectors_csv((tag + "_vectors.csv").c_str(), n, n, fld.data() + N, fld.data() + 2 * N) &&
           lbm_out::write_vtk((tag + ".vtk").c_str(), n, n, 1, fld.data(), fld.data() + N, fld.data() + 2 * N, nullptr);
  };
  const long chunk = 100;
  double ms_total = 0;
  for (long done = 0; done < steps;) {
    const long k = steps - done < chunk ? steps - done : chunk;
    double ms = 0;
    if (lbm_step_timed(h, k, &ms) != LBM_OK) return die("lbm_step_timed");
    ms_total += ms;
    done += k;
    csv.row(done, (double)N, ms_total * 1e-3);
    if (viz_every > 0 && done % viz_every == 0 && !snapshot(done)) return die("snapshot");
    if (done % 1000 != 0 && done != steps) continue;
    double mass = 0, ke = 0;
    int64_t bad = 0;
    lbm_monitors(h, &mass, &ke, &bad);
    std::printf("%ld iterations, %.3f s device time (%.1f MLUPS), mass %.9f, kinetic energy %.4e\n", done,
                ms_total * 1e-3, (double)N * done / (ms_total * 1e3), mass, ke);
    if (bad) {
      std::fprintf(stderr, "non-finite populations after %ld steps\n", done);
      return 2;
    }
  }
  if (out) {
    std::vector<double> f((size_t)4 * n * n);
    if (lbm_get_fields(h, f.data(), f.data() + N, f.data() + 2 * N, nullptr) != LBM_OK)
      return die("lbm_get_fields");
    FILE *fp = std::fopen(out, "wb");
    if (!fp) return 1;
    std::fwrite(f.data(), sizeof(double), 3 * N, fp);
    std::fclose(fp);
    std::printf("rho, ux, uy written to %s (%zu doubles each, x fastest)\n", out, N);
  }
  lbm_destroy(h);
  return 0;
}
