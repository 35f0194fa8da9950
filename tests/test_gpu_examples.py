"""The C++ host examples on the device: each program (the reference examples' inputs and outputs:
D2Q9 lid, two-fluid Shan-Chen, D3Q19 MRT channel, D3Q27 MRT lid) writes its headless snapshots
and MLUPS CSV, and the fields it writes equal the same lattice stepped through the library from
Python, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from lbm_amd import Lattice, output
from lbm_amd.capi import BC_PHYSICAL, D2Q9, F32, NODE_MOVING, NODE_SOLID

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lbm-arrayfire_amd", "lib", "d2q9_bgk_lid")


@pytest.mark.skipif(not os.path.exists(EXE), reason="examples not built (run build())")
def test_d2q9_example_snapshots_match_library(tmp_path):
    n, steps, re = 64, 300, 100.0
    env = dict(os.environ, LBM_GPU_NAME="T")
    r = subprocess.run([EXE, "0", str(n), str(steps), str(re), "-", "physical", "100"], cwd=tmp_path,
                       env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for it in (100, 200, 300):
        for suffix in (".ppm", "_vectors.csv", ".vtk"):
            assert (tmp_path / f"lid_{it}{suffix}").exists()
    rows = (tmp_path / f"T_d2q9_bgk_lid_mlups_{n}.csv").read_text().splitlines()
    assert rows[0] == "Iterations,MLUPS,MLUPS_x10" and [int(x.split(",")[0]) for x in rows[1:]] == [100, 200, 300]
    dims, rho, u = output.read_vtk(tmp_path / f"lid_{steps}.vtk")
    assert dims == (n, n, 1)

    u_lid = 0.05
    nu = u_lid * n / re
    types = np.zeros((n, n), np.uint8)
    types[0, :] = types[-1, :] = NODE_SOLID
    types[:, 0] = types[:, -1] = NODE_SOLID
    types[-1, :] |= NODE_MOVING
    cfg = dict(lattice=D2Q9, nx=n, ny=n, nz=1, storage=F32, bc_mode=BC_PHYSICAL,
               omega=1.0 / (3.0 * nu + 0.5), u_wall=(u_lid, 0.0, 0.0), rho_bc=1.0)
    with Lattice(cfg, types.reshape(-1)) as lat:
        lat.init_reference(1.0)
        lat.step(steps)
        want = lat.fields()
    assert np.array_equal(rho, want[0])
    assert np.array_equal(u[:, 0], want[1]) and np.array_equal(u[:, 1], want[2])
    img = (tmp_path / f"lid_{steps}.ppm").read_bytes()
    assert img.startswith(f"P6\n{n} {n}\n255\n".encode())


def splitmix_uniform(seed, n):
    """The example's noise: splitmix64(seed + i) >> 11, scaled to [0, 1)."""
    z = (np.arange(n, dtype=np.uint64) + np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15))
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


SC_EXE = os.path.join(ROOT, "lbm-arrayfire_amd", "lib", "d2q9_shanchen")


@pytest.mark.skipif(not os.path.exists(SC_EXE), reason="examples not built (run build())")
def test_d2q9_shanchen_example_matches_library(tmp_path):
    from lbm_amd.capi import ShanChen

    n, steps, seed = 48, 300, 7
    env = dict(os.environ, LBM_GPU_NAME="T")
    r = subprocess.run([SC_EXE, "0", str(n), str(steps), "f32", "100", "out.bin", str(seed)], cwd=tmp_path,
                       env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for it in (100, 200, 300):
        assert (tmp_path / f"shanchen_{it}.ppm").exists() and (tmp_path / f"shanchen_{it}.vtk").exists()
    rows = (tmp_path / f"T_d2q9_shanchen_mlups_{n}.csv").read_text().splitlines()
    assert [int(x.split(",")[0]) for x in rows[1:]] == [100, 200, 300]
    got = np.fromfile(tmp_path / "out.bin", dtype="<f8").reshape(4, n * n)
    d = -0.001 * (1.0 - 2.0 * splitmix_uniform(seed, n * n))
    with ShanChen(n, n, 1.0, 1.0, -1.2, F32) as sc:
        sc.init(1.0 + d, 1.0 - d)
        sc.step(steps)
        want = sc.fields()
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    dims, rho, u = output.read_vtk(tmp_path / f"shanchen_{steps}.vtk")
    assert dims == (n, n, 1) and np.array_equal(rho, want[0]) and np.array_equal(u[:, 0], want[2])
    img = (tmp_path / f"shanchen_{steps}.ppm").read_bytes()
    assert img == f"P6\n{n} {n}\n255\n".encode() + output.scalar_image(n, n, want[0]).tobytes()


CH_EXE = os.path.join(ROOT, "lbm-arrayfire_amd", "lib", "d3q19_mrt_channel")


@pytest.mark.skipif(not os.path.exists(CH_EXE), reason="examples not built (run build())")
def test_d3q19_channel_example_matches_library(tmp_path):
    from lbm_amd import cases
    from lbm_amd.capi import MRT_GS

    nx, ny, nz, r, steps, re = 64, 24, 24, 4, 200, 10.0
    env = dict(os.environ, LBM_GPU_NAME="T")
    res = subprocess.run([CH_EXE, "0", str(nx), str(ny), str(nz), str(r), str(steps), str(re), "f32", "100",
                          "out.bin", "viz"], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "C_d" in res.stdout and (tmp_path / "channel_200.vtk").exists()
    rows = (tmp_path / f"T_d3q19_mrt_channel_mlups_{nx}.csv").read_text().splitlines()
    assert [int(x.split(",")[0]) for x in rows[1:]] == [100, 200]
    got = np.fromfile(tmp_path / "out.bin", dtype="<f8").reshape(4, nx * ny * nz)
    cfg, types = cases.d3q19_sphere_channel(nx, ny, nz, (nx // 4, ny // 2, nz // 2), r, re=re, storage=F32,
                                            collision=MRT_GS)
    with Lattice(cfg, types.reshape(-1)) as lat:
        lat.init_equilibrium(1.0, (0.05, 0.0, 0.0), 1, 0.0)
        lat.step(steps)
        want = lat.fields()
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


MRT_EXE = os.path.join(ROOT, "lbm-arrayfire_amd", "lib", "d3q27_mrt_lid")


@pytest.mark.skipif(not os.path.exists(MRT_EXE), reason="examples not built (run build())")
def test_d3q27_mrt_lid_example_matches_library(tmp_path):
    """The reference benchmark program's inputs (rust/examples/d3q27_mrt_lid.rs: raw moment
    matrix, its rates and equilibrium table, lid 0.1, rho0/27 plus one collision): its VTK fields
    equal the same lattice stepped through the library from Python, bit for bit."""
    from lbm_amd import cases
    from lbm_amd.capi import MRT_SCALAR

    n, steps = 16, 50
    env = dict(os.environ, LBM_GPU_NAME="T")
    r = subprocess.run([MRT_EXE, "0", str(steps), "scalar", "f32", str(n), "viz"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "NON-FINITE" not in r.stdout
    rows = (tmp_path / f"T_d3q27_mrt_lid_mlups_{n}.csv").read_text().splitlines()
    assert rows[0] == "Iterations,MLUPS,MLUPS_x10" and int(rows[-1].split(",")[0]) == steps
    dims, rho, u = output.read_vtk(tmp_path / f"d3q27_mrt_lid_{n}.vtk")
    assert dims == (n, n, n)
    cfg, types = cases.d3q27_mrt_lid_reference(n, storage=F32, engine=MRT_SCALAR)
    cfg = dict(cfg, omega=0.0)  # the program leaves omega unset (MRT uses the rates)
    with Lattice(cfg, types.reshape(-1)) as lat:
        lat.init_reference(1.0)
        lat.step(steps)
        want = lat.fields()
    assert np.array_equal(rho, want[0])
    for k in range(3):
        assert np.array_equal(u[:, k], want[1 + k])
