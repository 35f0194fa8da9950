"""Headless replacements for the reference's Forge window and its MLUPS CSV files.

The reference shows, every 10 iterations, the normalised speed |u| as an image (solid nodes blank)
and the velocity sampled every nx/30 nodes as a vector field (cpp/d2q9_bgk_lid.cpp:157-171,
normalize :13-16), and writes one CSV row of cumulative MLUPS per iteration
(rust/examples/d2q9_bgk_lid.rs:25-35). These writers produce the same content as files, byte for
byte what the C++ examples' lbm-arrayfire_amd/examples/lbm_output.hpp writes:

- ``write_speed_ppm``: binary PPM, rows top-down (y = ny - 1 first), values normalize(|u|) =
  |u| / (1.1 max|u|) + 0.1 over non-solid nodes through a spectrum colour map, solid nodes black;
- ``write_scalar_ppm``: a scalar field (the two-fluid example's density of fluid 1,
  cpp/d2q9_shanchen.cpp:181-187) as the same image, values scaled to (a - min) / (max - min)
  (a constant field at 0.5);
- ``write_vectors_csv``: x,y,ux,uy sampled at x, y = 0, s, 2s, ... with s = max(1, nx // 30);
- ``write_vtk``: rho and u of every node, legacy VTK STRUCTURED_POINTS, big-endian doubles;
- ``MlupsCsv``: Iterations,MLUPS,MLUPS_x10 rows in the reference's file naming, MLUPS with the
  1e-6 factor and MLUPS_x10 with the reference's 10e-6.
"""
import math
import os

import numpy as np


def speed(ux, uy, uz=None, types=None):
    """|u| per node; NaN at solid nodes (type bit 0x01)."""
    ux = np.asarray(ux, dtype=np.float64).reshape(-1)
    uy = np.asarray(uy, dtype=np.float64).reshape(-1)
    uz = np.zeros_like(ux) if uz is None else np.asarray(uz, dtype=np.float64).reshape(-1)
    s = np.sqrt(ux * ux + uy * uy + uz * uz)
    if types is not None:
        s[(np.asarray(types).reshape(-1) & 0x01) != 0] = np.nan
    return s


def normalize(a):
    """The reference's normalize: a / (1.1 max |a|) + 0.1, NaN kept (max over the other values)."""
    a = np.array(a, dtype=np.float64, copy=True)
    ok = ~np.isnan(a)
    m = float(np.max(np.abs(a[ok]))) if ok.any() else 0.0
    d = 1.1 * m if m > 0 else 1.0
    a[ok] = a[ok] / d + 0.1
    return a


def spectrum(v):
    """RGB bytes of values in [0, 1]: hue 270 deg (violet) at 0 down to 0 deg (red) at 1; NaN
    black."""
    v = np.asarray(v, dtype=np.float64)
    out = np.zeros(v.shape + (3,), dtype=np.uint8)
    nan = np.isnan(v)
    w = np.clip(np.where(nan, 0.0, v), 0.0, 1.0)
    h = (1.0 - w) * 270.0 / 60.0
    k = h.astype(np.int64)
    f = h - k
    r = np.select([k == 0, k == 1, k == 2, k == 3], [np.ones_like(f), 1 - f, 0 * f, 0 * f], f)
    g = np.select([k == 0, k == 1, k == 2, k == 3], [f, np.ones_like(f), np.ones_like(f), 1 - f], 0 * f)
    b = np.select([k == 0, k == 1, k == 2, k == 3], [0 * f, 0 * f, f, np.ones_like(f)], np.ones_like(f))
    for c, x in enumerate((r, g, b)):
        out[..., c] = np.floor(255.0 * x + 0.5).astype(np.uint8)  # round half away from zero (x >= 0)
    out[nan] = 0
    return out


def speed_image(nx, ny, ux, uy, types=None):
    """ny x nx x 3 bytes: row 0 is y = ny - 1 (the reference displays flip(transpose(uu), 0))."""
    s = normalize(speed(ux, uy, None, types)).reshape(ny, nx)
    return spectrum(s[::-1])


def write_speed_ppm(path, nx, ny, ux, uy, types=None):
    img = speed_image(nx, ny, ux, uy, types)
    with open(path, "wb") as fp:
        fp.write(f"P6\n{nx} {ny}\n255\n".encode())
        fp.write(img.tobytes())


def scalar_image(nx, ny, a):
    """ny x nx x 3 bytes of a scalar field scaled to its range, row 0 at y = ny - 1."""
    a = np.asarray(a, dtype=np.float64).reshape(ny, nx)
    lo, hi = float(a.min()), float(a.max())
    v = (a - lo) / (hi - lo) if hi > lo else np.full_like(a, 0.5)
    return spectrum(v[::-1])


def write_scalar_ppm(path, nx, ny, a):
    with open(path, "wb") as fp:
        fp.write(f"P6\n{nx} {ny}\n255\n".encode())
        fp.write(scalar_image(nx, ny, a).tobytes())


def vector_samples(nx, ny, ux, uy):
    """(x, y, ux, uy) at x, y = 0, s, 2s, ... with s = max(1, nx // 30), rows of y."""
    s = max(1, nx // 30)
    ux = np.asarray(ux, dtype=np.float64).reshape(ny, nx)
    uy = np.asarray(uy, dtype=np.float64).reshape(ny, nx)
    ys, xs = np.meshgrid(np.arange(0, ny, s), np.arange(0, nx, s), indexing="ij")
    return xs.reshape(-1), ys.reshape(-1), ux[ys, xs].reshape(-1), uy[ys, xs].reshape(-1)


def write_vectors_csv(path, nx, ny, ux, uy):
    xs, ys, u, v = vector_samples(nx, ny, ux, uy)
    with open(path, "w") as fp:
        fp.write("x,y,ux,uy\n")
        for x, y, a, b in zip(xs, ys, u, v):
            fp.write("%d,%d,%.17g,%.17g\n" % (x, y, a, b))


def write_vtk(path, nx, ny, nz, rho, ux, uy, uz=None):
    n = nx * ny * nz
    rho = np.asarray(rho, dtype=">f8").reshape(-1)
    vel = np.zeros((n, 3), dtype=">f8")
    vel[:, 0] = np.asarray(ux).reshape(-1)
    vel[:, 1] = np.asarray(uy).reshape(-1)
    if uz is not None:
        vel[:, 2] = np.asarray(uz).reshape(-1)
    with open(path, "wb") as fp:
        fp.write(("# vtk DataFile Version 3.0\nlbm fields\nBINARY\nDATASET STRUCTURED_POINTS\n"
                  f"DIMENSIONS {nx} {ny} {nz}\nORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA {n}\n"
                  "SCALARS rho double 1\nLOOKUP_TABLE default\n").encode())
        fp.write(rho.tobytes())
        fp.write(b"\nVECTORS velocity double\n")
        fp.write(vel.tobytes())
        fp.write(b"\n")


def read_vtk(path):
    """(dims, rho, u) of a file written by write_vtk."""
    with open(path, "rb") as fp:
        data = fp.read()
    head_end = data.index(b"LOOKUP_TABLE default\n") + len(b"LOOKUP_TABLE default\n")
    head = data[:head_end].decode().split("\n")
    dims = tuple(int(v) for v in next(h for h in head if h.startswith("DIMENSIONS")).split()[1:])
    n = int(np.prod(dims))
    rho = np.frombuffer(data, dtype=">f8", count=n, offset=head_end).astype(np.float64)
    vstart = data.index(b"VECTORS velocity double\n", head_end + 8 * n) + len(b"VECTORS velocity double\n")
    u = np.frombuffer(data, dtype=">f8", count=3 * n, offset=vstart).astype(np.float64).reshape(n, 3)
    return dims, rho, u


class MlupsCsv:
    """<dir>/<gpu>_<example>_mlups_<size>.csv with Iterations,MLUPS,MLUPS_x10 rows."""

    def __init__(self, directory, gpu, example, size):
        self.path = os.path.join(directory, f"{gpu}_{example}_mlups_{size}.csv")
        self.fp = open(self.path, "w")
        self.fp.write("Iterations,MLUPS,MLUPS_x10\n")

    def row(self, iterations, nodes, seconds):
        m = nodes * iterations * 1e-6 / seconds if seconds > 0 else 0.0
        self.fp.write("%d,%.6f,%.6f\n" % (iterations, m, 10.0 * m))

    def close(self):
        self.fp.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def snapshot(lat, directory, tag, types=None):
    """Write <tag>.vtk of every node and, for 2D lattices, <tag>.ppm and <tag>_vectors.csv, from a
    Lattice's current fields (the headless form of the reference's per-iteration window)."""
    rho, ux, uy, uz = lat.fields()
    nx, ny, nz = lat.cfg["nx"], lat.cfg["ny"], lat.cfg.get("nz", 1)
    write_vtk(os.path.join(directory, f"{tag}.vtk"), nx, ny, nz, rho, ux, uy, uz if nz > 1 else None)
    if nz == 1:
        write_speed_ppm(os.path.join(directory, f"{tag}.ppm"), nx, ny, ux, uy, types)
        write_vectors_csv(os.path.join(directory, f"{tag}_vectors.csv"), nx, ny, ux, uy)
    return math.nan if rho.size == 0 else float(np.nanmax(speed(ux, uy, uz, types)))
