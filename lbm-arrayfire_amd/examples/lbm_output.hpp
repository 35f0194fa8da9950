// lbm_output.hpp — headless replacements for the reference's Forge window and its MLUPS CSV files.
//
// The reference shows, every 10 iterations, the normalised speed |u| as an image (solid nodes
// blank) and the velocity sampled every nx/30 nodes as a vector field (cpp/d2q9_bgk_lid.cpp:157-171,
// normalize :13-16), and writes one CSV row of cumulative MLUPS per iteration
// (rust/examples/d2q9_bgk_lid.rs:25-35).  Here:
//   write_speed_ppm     the image: binary PPM, rows top-down (y = ny - 1 first), x left to right
//                       (the reference displays flip(transpose(uu), 0)), value normalize(|u|) =
//                       |u| / (1.1 max|u|) + 0.1 over non-solid nodes, mapped through a spectrum
//                       colour map (hue 270 deg at 0 down to 0 deg at 1); solid nodes black
//   write_scalar_ppm    a scalar field (the two-fluid example's density of fluid 1,
//                       cpp/d2q9_shanchen.cpp:181-187) as the same image, values scaled to their
//                       range: (a - min) / (max - min), a constant field at 0.5
//   write_vectors_csv   the sampled vector field: x,y,ux,uy at x, y = 0, s, 2s, ... (s = max(1, nx/30))
//   write_vtk           rho and u of every node: legacy VTK, STRUCTURED_POINTS, big-endian doubles
//   MlupsCsv            Iterations,MLUPS,MLUPS_x10 rows in the reference's file naming
//                       (<gpu>_<example>_mlups_<size>.csv); MLUPS with the 1e-6 factor, MLUPS_x10
//                       with the reference's 10e-6, for plotting beside its published curves
// Host C++ only; the Python package (lbm_amd/output.py) writes byte-identical files.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace lbm_out {

// |u| per node (2D: uz may be null), NaN at solid nodes (type bit 0x01).
inline std::vector<double> speed(int64_t n, const double *ux, const double *uy, const double *uz,
                                 const uint8_t *type) {
  std::vector<double> s((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const double z = uz ? uz[i] : 0.0;
    s[(size_t)i] = (type && (type[i] & 0x01)) ? NAN : std::sqrt(ux[i] * ux[i] + uy[i] * uy[i] + z * z);
  }
  return s;
}

// The reference's normalize: a / (1.1 max |a|) + 0.1, NaN kept (max over non-NaN values).
inline void normalize(std::vector<double> &a) {
  double m = 0;
  for (double v : a)
    if (!std::isnan(v) && std::fabs(v) > m) m = std::fabs(v);
  const double d = m > 0 ? 1.1 * m : 1.0;
  for (double &v : a)
    if (!std::isnan(v)) v = v / d + 0.1;
}

// Spectrum colour of v in [0, 1]: hue from 270 deg (violet) down to 0 deg (red), full saturation
// and value; NaN is black.
inline void spectrum(double v, uint8_t rgb[3]) {
  if (std::isnan(v)) {
    rgb[0] = rgb[1] = rgb[2] = 0;
    return;
  }
  v = v < 0 ? 0 : (v > 1 ? 1 : v);
  const double h = (1.0 - v) * 270.0 / 60.0;  // sector 0..4.5
  const int k = (int)h;
  const double f = h - k;
  double r = 0, g = 0, b = 0;
  switch (k) {
    case 0: r = 1; g = f; break;
    case 1: r = 1 - f; g = 1; break;
    case 2: g = 1; b = f; break;
    case 3: g = 1 - f; b = 1; break;
    default: r = f; b = 1; break;  // 4: blue towards violet
  }
  rgb[0] = (uint8_t)std::floor(255.0 * r + 0.5);
  rgb[1] = (uint8_t)std::floor(255.0 * g + 0.5);
  rgb[2] = (uint8_t)std::floor(255.0 * b + 0.5);
}

inline bool write_speed_ppm(const char *path, int nx, int ny, const double *ux, const double *uy,
                            const uint8_t *type) {
  std::vector<double> s = speed((int64_t)nx * ny, ux, uy, nullptr, type);
  normalize(s);
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return false;
  std::fprintf(fp, "P6\n%d %d\n255\n", nx, ny);
  std::vector<uint8_t> row((size_t)nx * 3);
  for (int r = 0; r < ny; ++r) {
    const int y = ny - 1 - r;
    for (int x = 0; x < nx; ++x) spectrum(s[(size_t)x + (size_t)nx * y], &row[(size_t)3 * x]);
    std::fwrite(row.data(), 1, row.size(), fp);
  }
  return std::fclose(fp) == 0;
}

inline bool write_scalar_ppm(const char *path, int nx, int ny, const double *a) {
  const size_t n = (size_t)nx * ny;
  double lo = 0, hi = 0;
  for (size_t i = 0; i < n; ++i) {
    if (i == 0 || a[i] < lo) lo = a[i];
    if (i == 0 || a[i] > hi) hi = a[i];
  }
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return false;
  std::fprintf(fp, "P6\n%d %d\n255\n", nx, ny);
  std::vector<uint8_t> row((size_t)nx * 3);
  for (int r = 0; r < ny; ++r) {
    const int y = ny - 1 - r;
    for (int x = 0; x < nx; ++x) {
      const double v = a[(size_t)x + (size_t)nx * y];
      spectrum(hi > lo ? (v - lo) / (hi - lo) : 0.5, &row[(size_t)3 * x]);
    }
    std::fwrite(row.data(), 1, row.size(), fp);
  }
  return std::fclose(fp) == 0;
}

inline bool write_vectors_csv(const char *path, int nx, int ny, const double *ux, const double *uy) {
  const int s = nx / 30 > 1 ? nx / 30 : 1;
  FILE *fp = std::fopen(path, "w");
  if (!fp) return false;
  std::fprintf(fp, "x,y,ux,uy\n");
  for (int y = 0; y < ny; y += s)
    for (int x = 0; x < nx; x += s) {
      const size_t i = (size_t)x + (size_t)nx * y;
      std::fprintf(fp, "%d,%d,%.17g,%.17g\n", x, y, ux[i], uy[i]);
    }
  return std::fclose(fp) == 0;
}

inline void put_be64(FILE *fp, double v) {
  uint64_t u;
  static_assert(sizeof u == sizeof v, "double is 64-bit");
  __builtin_memcpy(&u, &v, sizeof u);
  uint8_t b[8];
  for (int k = 0; k < 8; ++k) b[k] = (uint8_t)(u >> (56 - 8 * k));
  std::fwrite(b, 1, 8, fp);
}

inline bool write_vtk(const char *path, int nx, int ny, int nz, const double *rho, const double *ux,
                      const double *uy, const double *uz) {
  const size_t n = (size_t)nx * ny * nz;
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return false;
  std::fprintf(fp,
               "# vtk DataFile Version 3.0\nlbm fields\nBINARY\nDATASET STRUCTURED_POINTS\n"
               "DIMENSIONS %d %d %d\nORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA %zu\n"
               "SCALARS rho double 1\nLOOKUP_TABLE default\n",
               nx, ny, nz, n);
  for (size_t i = 0; i < n; ++i) put_be64(fp, rho[i]);
  std::fprintf(fp, "\nVECTORS velocity double\n");
  for (size_t i = 0; i < n; ++i) {
    put_be64(fp, ux[i]);
    put_be64(fp, uy[i]);
    put_be64(fp, uz ? uz[i] : 0.0);
  }
  std::fprintf(fp, "\n");
  return std::fclose(fp) == 0;
}

// <dir>/<gpu>_<example>_mlups_<size>.csv with Iterations,MLUPS,MLUPS_x10 rows.
struct MlupsCsv {
  FILE *fp = nullptr;
  std::string path;
  bool open(const std::string &dir, const std::string &gpu, const std::string &example, const std::string &size) {
    path = (dir.empty() ? std::string() : dir + "/") + gpu + "_" + example + "_mlups_" + size + ".csv";
    fp = std::fopen(path.c_str(), "w");
    if (fp) std::fprintf(fp, "Iterations,MLUPS,MLUPS_x10\n");
    return fp != nullptr;
  }
  // cumulative rate after `iterations` steps of `nodes` nodes in `seconds`
  void row(long iterations, double nodes, double seconds) {
    if (!fp) return;
    const double m = seconds > 0 ? nodes * (double)iterations * 1e-6 / seconds : 0.0;
    std::fprintf(fp, "%ld,%.6f,%.6f\n", iterations, m, 10.0 * m);
  }
  ~MlupsCsv() {
    if (fp) std::fclose(fp);
  }
};

}  // namespace lbm_out
