"""Headless field output and the MLUPS CSV (the reference's Forge window and CSV writer,
cpp/d2q9_bgk_lid.cpp:157-171, rust/examples/d2q9_bgk_lid.rs:25-35): the C++ writers of the host
examples (examples/lbm_output.hpp, through lib/output_selftest) and the Python writers
(lbm_amd/output.py) produce the same bytes, and the content has the reference's meaning:
normalize(a) = a / (1.1 max|a|) + 0.1 with solid nodes blank, rows with y pointing up, vectors
sampled every nx/30 nodes, VTK fields that read back exactly, MLUPS with the 1e-6 factor."""
import os
import subprocess

import numpy as np
import pytest

from lbm_amd import output

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SELFTEST = os.path.join(ROOT, "lbm-arrayfire_amd", "lib", "output_selftest")


def fields(nx, ny, seed=3):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.01 * rng.standard_normal(nx * ny)
    ux = 0.05 * rng.standard_normal(nx * ny)
    uy = 0.05 * rng.standard_normal(nx * ny)
    types = np.zeros(nx * ny, np.uint8)
    types[rng.random(nx * ny) < 0.1] = 1  # solid
    types[:nx] = 1
    ux[types == 1] = 0.0
    uy[types == 1] = 0.0
    return rho, ux, uy, types


@pytest.mark.skipif(not os.path.exists(SELFTEST), reason="lib/output_selftest not built (run build())")
@pytest.mark.parametrize("nx,ny", [(47, 23), (96, 64), (31, 1)])
def test_cpp_and_python_writers_are_byte_identical(tmp_path, nx, ny):
    rho, ux, uy, types = fields(nx, ny)
    blob = np.concatenate([rho, ux, uy]).astype("<f8").tobytes() + types.tobytes()
    subprocess.run([SELFTEST, str(tmp_path), str(nx), str(ny)], input=blob, check=True)
    output.write_speed_ppm(tmp_path / "p.ppm", nx, ny, ux, uy, types)
    output.write_vectors_csv(tmp_path / "p_vectors.csv", nx, ny, ux, uy)
    output.write_scalar_ppm(tmp_path / "p_rho.ppm", nx, ny, rho)
    output.write_vtk(tmp_path / "p.vtk", nx, ny, 1, rho, ux, uy)
    with output.MlupsCsv(str(tmp_path), "GPU", "p", nx) as csv:
        for it in range(100, 501, 100):
            csv.row(it, nx * ny, 1e-3 * it / 7.0)
    for a, b in (("c.ppm", "p.ppm"), ("c_rho.ppm", "p_rho.ppm"), ("c_vectors.csv", "p_vectors.csv"), ("c.vtk", "p.vtk"),
                 (f"GPU_c_mlups_{nx}.csv", f"GPU_p_mlups_{nx}.csv")):
        assert (tmp_path / a).read_bytes() == (tmp_path / b).read_bytes(), a


def test_speed_image_meaning(tmp_path):
    nx, ny = 40, 30
    rho, ux, uy, types = fields(nx, ny)
    types[:] = 0
    ux[:] = 0.0
    uy[:] = 0.0
    ux[5 + nx * (ny - 2)] = 0.2  # the fastest node, near the top
    ux[7 + nx * 3] = 0.1
    types[9 + nx * 4] = 1  # a solid node
    s = output.normalize(output.speed(ux, uy, None, types))
    assert np.isnan(s[9 + nx * 4])
    assert s[5 + nx * (ny - 2)] == pytest.approx(0.2 / (1.1 * 0.2) + 0.1)
    assert s[7 + nx * 3] == pytest.approx(0.1 / 0.22 + 0.1)
    assert s[0] == pytest.approx(0.1)
    output.write_speed_ppm(tmp_path / "i.ppm", nx, ny, ux, uy, types)
    data = (tmp_path / "i.ppm").read_bytes()
    head = f"P6\n{nx} {ny}\n255\n".encode()
    assert data.startswith(head)
    img = np.frombuffer(data[len(head):], np.uint8).reshape(ny, nx, 3)
    assert tuple(img[1, 5]) == (255, 0, 0)          # y = ny - 2 is row 1; fastest is red
    assert tuple(img[ny - 1 - 4, 9]) == (0, 0, 0)   # solid: blank
    r, g, b = img[ny - 1 - 10, 20]                   # at rest (0.1): violet-blue
    assert b == 255 and g == 0 and r > 0


def test_vectors_and_vtk_round_trip(tmp_path):
    nx, ny = 95, 20
    rho, ux, uy, _ = fields(nx, ny, seed=8)
    xs, ys, u, v = output.vector_samples(nx, ny, ux, uy)
    s = nx // 30
    assert set(np.unique(xs)) == set(range(0, nx, s)) and set(np.unique(ys)) == set(range(0, ny, s))
    assert np.array_equal(u, ux[xs + nx * ys]) and np.array_equal(v, uy[xs + nx * ys])
    output.write_vtk(tmp_path / "f.vtk", nx, ny, 1, rho, ux, uy)
    dims, r2, u2 = output.read_vtk(tmp_path / "f.vtk")
    assert dims == (nx, ny, 1)
    assert np.array_equal(r2, rho) and np.array_equal(u2[:, 0], ux) and np.array_equal(u2[:, 1], uy)
    assert not u2[:, 2].any()


def test_mlups_csv(tmp_path):
    with output.MlupsCsv(str(tmp_path), "MI355X", "d2q9_bgk_lid", 128) as csv:
        csv.row(1000, 128 * 128, 0.5)
    lines = (tmp_path / "MI355X_d2q9_bgk_lid_mlups_128.csv").read_text().splitlines()
    assert lines[0] == "Iterations,MLUPS,MLUPS_x10"
    it, m, m10 = lines[1].split(",")
    assert int(it) == 1000
    assert float(m) == pytest.approx(128 * 128 * 1000 / 0.5 / 1e6)
    assert float(m10) == pytest.approx(10 * float(m))


def test_scalar_image_meaning():
    nx, ny = 6, 4
    a = np.zeros(nx * ny)
    a[2 + nx * 3] = 2.0   # the maximum, top row
    a[4 + nx * 0] = -1.0  # the minimum, bottom row
    img = output.scalar_image(nx, ny, a)
    assert tuple(img[0, 2]) == (255, 0, 0)            # 1.0: red
    assert tuple(img[ny - 1, 4]) == tuple(output.spectrum(0.0))
    assert tuple(img[1, 1]) == tuple(output.spectrum(1.0 / 3.0))
    flat = output.scalar_image(nx, ny, np.ones(nx * ny))
    assert (flat == output.spectrum(0.5)).all()
