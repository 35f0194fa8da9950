// d2q9_shanchen.cpp — two-component Shan-Chen D2Q9 through the C ABI (lbm_sc_*).
//
// The reference program cpp/d2q9_shanchen.cpp:30-199 as a headless run: an n x n periodic box
// (default 100), G = -1.2, omega1 = omega2 = 1, rho1 = 1 + d, rho2 = 1 - d with d uniform in
// [-drho, drho], drho = 0.001, both fluids at rest; 15000 iterations (its maxiter).  Every
// viz_every iterations it writes the window's content as files (shanchen_<iter>.ppm: the density
// of fluid 1 scaled to its range; shanchen_<iter>.vtk: rho1 and the barycentric velocity), every
// 100 iterations a progress line with the cumulative MLUPS (:188-192, with the 1e-6 factor) and a
// row of <gpu>_d2q9_shanchen_mlups_<n>.csv.  The noise comes from splitmix64(seed + node) (the
// reference draws af::randu), so a run can be repeated exactly; the final fields can be written
// as raw little-endian doubles (rho1, rho2, ux, uy).
//
//   d2q9_shanchen [device] [n] [steps] [f32|f64] [viz_every] [out.bin|-] [seed]
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

// uniform double in [0, 1) of 53 random bits (splitmix64)
static double uniform01(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int main(int argc, char **argv) {
  const int device = argc > 1 ? std::atoi(argv[1]) : 0;
  const int n = argc > 2 ? std::atoi(argv[2]) : 100;
  const long steps = argc > 3 ? std::atol(argv[3]) : 15000;
  const bool f64 = argc > 4 && std::strcmp(argv[4], "f64") == 0;
  const long viz_every = argc > 5 ? std::atol(argv[5]) : 0;
  const char *out = argc > 6 && std::strcmp(argv[6], "-") != 0 ? argv[6] : nullptr;
  const uint64_t seed = argc > 7 ? std::strtoull(argv[7], nullptr, 10) : 1;
  const char *gpu = std::getenv("LBM_GPU_NAME") ? std::getenv("LBM_GPU_NAME") : "MI355X";
  if (lbm_set_device(device) != LBM_OK) return die("lbm_set_device");

  const double G = -1.2, omega1 = 1.0, omega2 = 1.0, drho = 0.001;
  std::printf("D2Q9 Shan-Chen %dx%d, G %.2f, omega %.2f / %.2f, %s\n", n, n, G, omega1, omega2,
              f64 ? "FP64" : "FP32");
  const size_t N = (size_t)n * n;
  std::vector<double> rho1(N), rho2(N);
  for (size_t i = 0; i < N; ++i) {
    const double d = -drho * (1.0 - 2.0 * uniform01(seed + i));  // cpp/d2q9_shanchen.cpp:44
    rho1[i] = 1.0 + d;
    rho2[i] = 1.0 - d;
  }
  lbm_sc_desc d;
  std::memset(&d, 0, sizeof d);
  d.nx = n;
  d.ny = n;
  d.storage = f64 ? LBM_F64 : LBM_F32;
  d.omega1 = omega1;
  d.omega2 = omega2;
  d.g = G;
  lbm_sc_handle h;
  if (lbm_sc_create(&d, &h) != LBM_OK) return die("lbm_sc_create");
  if (lbm_sc_init(h, rho1.data(), rho2.data()) != LBM_OK) return die("lbm_sc_init");

  lbm_out::MlupsCsv csv;
  csv.open("", gpu, "d2q9_shanchen", std::to_string(n));
  std::vector<double> fld(4 * N);
  auto fields = [&]() { return lbm_sc_get_fields(h, fld.data(), fld.data() + N, fld.data() + 2 * N, fld.data() + 3 * N); };
  auto snapshot = [&](long iter) -> bool {
    if (fields() != LBM_OK) return false;
    const std::string tag = "shanchen_" + std::to_string(iter);
    return lbm_out::write_scalar_ppm((tag + ".ppm").c_str(), n, n, fld.data()) &&
           lbm_out::write_vtk((tag + ".vtk").c_str(), n, n, 1, fld.data(), fld.data() + 2 * N, fld.data() + 3 * N, nullptr);
  };
  const long chunk = viz_every > 0 && viz_every < 100 ? viz_every : 100;
  double ms_total = 0;
  for (long done = 0; done < steps;) {
    const long k = steps - done < chunk ? steps - done : chunk;
    double ms = 0;
    if (lbm_sc_step_timed(h, k, &ms) != LBM_OK) return die("lbm_sc_step_timed");
    ms_total += ms;
    done += k;
    if (viz_every > 0 && done % viz_every == 0 && !snapshot(done)) return die("snapshot");
    if (done % 100 != 0 && done != steps) continue;
    csv.row(done, (double)N, ms_total * 1e-3);
    if (fields() != LBM_OK) return die("lbm_sc_get_fields");
    double m1 = 0, m2 = 0, lo = fld[0], hi = fld[0];
    for (size_t i = 0; i < N; ++i) {
      m1 += fld[i];
      m2 += fld[N + i];
      lo = fld[i] < lo ? fld[i] : lo;
      hi = fld[i] > hi ? fld[i] : hi;
    }
    if (!(m1 == m1) || !(m2 == m2)) {
      std::fprintf(stderr, "non-finite densities after %ld iterations\n", done);
      return 2;
    }
    if (done % 1000 == 0 || done == steps)
      std::printf("%ld iterations completed, %.3fs device time (%.1f MLUPS), mass %.6f / %.6f, rho1 in [%.4f, %.4f]\n",
                  done, ms_total * 1e-3, (double)N * done / (ms_total * 1e3), m1, m2, lo, hi);
  }
  if (out) {
    if (fields() != LBM_OK) return die("lbm_sc_get_fields");
    FILE *fp = std::fopen(out, "wb");
    if (!fp) return 1;
    std::fwrite(fld.data(), sizeof(double), 4 * N, fp);
    std::fclose(fp);
    std::printf("rho1, rho2, ux, uy written to %s (%zu doubles each, x fastest)\n", out, N);
  }
  lbm_sc_destroy(h);
  return 0;
}
