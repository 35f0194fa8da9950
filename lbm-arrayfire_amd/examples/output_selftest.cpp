// output_selftest.cpp — writes the files of lbm_output.hpp from fields given on stdin, so that a
// test can compare them byte for byte with the Python writers (lbm_amd/output.py).  No GPU.
//
//   output_selftest <dir> <nx> <ny> < fields (rho, ux, uy as 3 nx*ny little-endian doubles, then
//                                             nx*ny type bytes)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lbm_output.hpp"

int main(int argc, char **argv) {
  if (argc < 4) return 2;
  const std::string dir = argv[1];
  const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]);
  const size_t n = (size_t)nx * ny;
  std::vector<double> f(3 * n);
  std::vector<uint8_t> t(n);
  if (std::fread(f.data(), sizeof(double), 3 * n, stdin) != 3 * n) return 3;
  if (std::fread(t.data(), 1, n, stdin) != n) return 3;
  const double *rho = f.data(), *ux = f.data() + n, *uy = f.data() + 2 * n;
  bool ok = lbm_out::write_speed_ppm((dir + "/c.ppm").c_str(), nx, ny, ux, uy, t.data()) &&
            lbm_out::write_scalar_ppm((dir + "/c_rho.ppm").c_str(), nx, ny, rho) &&
            lbm_out::write_vectors_csv((dir + "/c_vectors.csv").c_str(), nx, ny, ux, uy) &&
            lbm_out::write_vtk((dir + "/c.vtk").c_str(), nx, ny, 1, rho, ux, uy, nullptr);
  lbm_out::MlupsCsv csv;
  ok = ok && csv.open(dir, "GPU", "c", std::to_string(nx));
  for (long it = 100; it <= 500; it += 100) csv.row(it, (double)n, 1e-3 * (double)it / 7.0);
  return ok ? 0 : 1;
}
