// d3q19_mrt_channel.cpp — D3Q19 MRT duct flow past a sphere through the C ABI.
//
// What the reference's work-in-progress program cpp/d3q19_mrt_channel.cpp sets out to run (an MRT
// channel with a zero-gradient outlet, :9-10,299-301), on the library's D3Q19 lattice:
// - geometry: walls on y = 0, ny - 1 and z = 0, nz - 1, a sphere of radius r at
//   (nx/4, ny/2, nz/2), a uniform velocity inlet (non-equilibrium extrapolation) on x = 0 and a
//   zero-gradient outlet at rho = 1 on x = nx - 1;
// - collision: MRT in d'Humieres et al.'s (2002) orthogonal basis, s_nu = omega on the stress
//   rows, their values on the others (1.19, 1.4, 1.2, 1.4) except s_m = 1.2 on the ghost rows
//   mx, my, mz (their 1.98 lets the inlet/outlet closures drive those modes unstable in a duct),
//   conserved rows 1;
// - Re = u_in * 2r / nu.
// Every report_every steps it prints the drag on the sphere (momentum exchange, lbm_force) as
// C_d = 2 F_x / (rho u_in^2 pi r^2) and the cumulative MLUPS, writes a row of
// <gpu>_d3q19_mrt_channel_mlups_<nx>.csv, and with viz a VTK file of rho and u
// (channel_<iter>.vtk).  The final fields can be written as raw little-endian doubles
// (rho, ux, uy, uz).
//
//   d3q19_mrt_channel [device] [nx] [ny] [nz] [r] [steps] [re] [f32|f16s|bf16s|f64] [report_every] [out.bin|-] [viz]
#include <cmath>
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
  const int nx = argc > 2 ? std::atoi(argv[2]) : 256;
  const int ny = argc > 3 ? std::atoi(argv[3]) : 96;
  const int nz = argc > 4 ? std::atoi(argv[4]) : 96;
  const int r = argc > 5 ? std::atoi(argv[5]) : 8;
  const long steps = argc > 6 ? std::atol(argv[6]) : 4000;
  const double re = argc > 7 ? std::atof(argv[7]) : 30.0;
  const std::string sto = argc > 8 ? argv[8] : "f32";
  const long report_every = argc > 9 ? std::atol(argv[9]) : 1000;
  const char *out = argc > 10 && std::strcmp(argv[10], "-") != 0 ? argv[10] : nullptr;
  const bool viz = argc > 11 && std::strcmp(argv[11], "viz") == 0;
  const char *gpu = std::getenv("LBM_GPU_NAME") ? std::getenv("LBM_GPU_NAME") : "MI355X";
  if (lbm_set_device(device) != LBM_OK) return die("lbm_set_device");

  const double u_in = 0.05, nu = u_in * 2.0 * r / re, omega = 1.0 / (3.0 * nu + 0.5);
  const int cx = nx / 4, cy = ny / 2, cz = nz / 2;
  const size_t N = (size_t)nx * ny * nz;
  std::vector<uint8_t> type(N, LBM_NODE_FLUID);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const long dx = x - cx, dy = y - cy, dz = z - cz;
        const bool solid = dx * dx + dy * dy + dz * dz <= (long)r * r || y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
        uint8_t t = solid ? LBM_NODE_SOLID : LBM_NODE_FLUID;
        if (!solid && x == 0) t |= LBM_NODE_INLET;
        if (!solid && x == nx - 1) t |= LBM_NODE_OUTLET;
        type[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)] = t;
      }
  lbm_desc d;
  std::memset(&d, 0, sizeof d);
  d.lattice = LBM_D3Q19;
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.collision = LBM_MRT_GS;
  d.storage = sto == "f16s" ? LBM_F16S : sto == "bf16s" ? LBM_BF16S : sto == "f64" ? LBM_F64 : LBM_F32;
  d.bc_mode = LBM_BC_PHYSICAL;
  d.omega = omega;
  // rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pi_xx, pww, pi_ww, pxy, pyz, pxz, mx, my, mz
  const double s19[19] = {1.0, 1.19, 1.4, 1.0, 1.2, 1.0, 1.2, 1.0, 1.2, omega, 1.4, omega, 1.4,
                          omega, omega, omega, 1.2, 1.2, 1.2};
  for (int k = 0; k < 27; ++k) d.s[k] = k < 19 ? s19[k] : 1.0;
  d.u_in = u_in;
  d.inlet_profile = LBM_INLET_UNIFORM;
  d.inlet_closure = LBM_INLET_EXTRAPOLATED;
  d.rho_bc = 1.0;
  std::printf("D3Q19 MRT channel %dx%dx%d, sphere r %d at (%d, %d, %d), Re %.1f, omega %.6f, %s\n", nx, ny, nz, r, cx,
              cy, cz, re, omega, sto.c_str());
  lbm_handle h;
  if (lbm_create(&d, type.data(), &h) != LBM_OK) return die("lbm_create");
  const double u0[3] = {u_in, 0.0, 0.0};
  if (lbm_init_equilibrium(h, 1.0, u0, 1, 0.0) != LBM_OK) return die("lbm_init_equilibrium");

  lbm_out::MlupsCsv csv;
  csv.open("", gpu, "d3q19_mrt_channel", std::to_string(nx));
  const double area = M_PI * r * r;
  const int32_t box[6] = {cx - r - 1, cx + r + 2, cy - r - 1, cy + r + 2, cz - r - 1, cz + r + 2};
  std::vector<double> fld;
  double ms_total = 0;
  for (long done = 0; done < steps;) {
    const long k = steps - done < report_every ? steps - done : report_every;
    double ms = 0;
    if (lbm_step_timed(h, k, &ms) != LBM_OK) return die("lbm_step_timed");
    ms_total += ms;
    done += k;
    csv.row(done, (double)N, ms_total * 1e-3);
    double F[3] = {0, 0, 0};
    if (lbm_force(h, box, F) != LBM_OK) return die("lbm_force");
    double mass = 0, ke = 0;
    int64_t bad = 0;
    lbm_monitors(h, &mass, &ke, &bad);
    std::printf("%ld steps, %.3f s device time (%.1f MLUPS), C_d %.4f, mass %.6f\n", done, ms_total * 1e-3,
                (double)N * done / (ms_total * 1e3), 2.0 * F[0] / (u_in * u_in * area), mass);
    if (bad) {
      std::fprintf(stderr, "non-finite populations after %ld steps\n", done);
      return 2;
    }
    if (viz || (out && done == steps)) {
      fld.resize(4 * N);
      if (lbm_get_fields(h, fld.data(), fld.data() + N, fld.data() + 2 * N, fld.data() + 3 * N) != LBM_OK)
        return die("lbm_get_fields");
      if (viz && !lbm_out::write_vtk(("channel_" + std::to_string(done) + ".vtk").c_str(), nx, ny, nz, fld.data(),
                                     fld.data() + N, fld.data() + 2 * N, fld.data() + 3 * N))
        return die("write_vtk");
    }
  }
  if (out) {
    FILE *fp = std::fopen(out, "wb");
    if (!fp) return 1;
    std::fwrite(fld.data(), sizeof(double), 4 * N, fp);
    std::fclose(fp);
    std::printf("rho, ux, uy, uz written to %s (%zu doubles each, x fastest)\n", out, N);
  }
  lbm_destroy(h);
  return 0;
}
