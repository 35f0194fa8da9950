"""Multi-step launches (K D2Q9 BGK steps per launch from an LDS tile) reproduce K single steps
bit for bit: every launch shape, both boundary modes, every storage (FP64 on the one-LDS-copy
shapes), lattices whose sides are not multiples of the tile, and a lattice smaller than one
tile's halo region (its positions wrap around the lattice more than once).  FP32 and FP64
multi-step runs also equal the oracle bit for bit."""
import numpy as np
import oracle as orc
import pytest

from lbm_amd import Lattice, cases, geometry as geo
from lbm_amd.capi import BF16S, F16S, F32, F64

pytestmark = pytest.mark.gpu

SHAPES = range(10)                      # launch shapes of lbm_multistep.hip
SHAPE_K = [4, 3, 4, 4, 4, 4, 4, 4, 2, 5]  # steps per launch of each shape
SHAPES_F64 = [3, 4, 5, 6, 7]              # the shapes with one LDS copy of the region


def case_list():
    out = []
    for st in (F32, F16S, BF16S, F64):
        cfg, t = cases.d2q9_karman(200, 70, 50, 35, 8, u_max=0.1, re=60.0, storage=st)
        out.append((f"karman_phys_200x70_st{st}", cfg, t, (0.08, 0.0, 0.0)))
    cfg, t = cases.d2q9_lid_physical(70, u_lid=0.1, re=100.0, storage=F16S)
    out.append(("lid_phys_70_f16s", cfg, t, (0.0, 0.0, 0.0)))
    cfg, t = cases.d2q9_lid_reference(66, storage=F32)
    out.append(("lid_ref_66_f32", cfg, t, (0.0, 0.0, 0.0)))
    cfg, t = cases.d2q9_lid_reference(66, storage=F64)
    out.append(("lid_ref_66_f64", cfg, t, (0.0, 0.0, 0.0)))
    cfg, t = cases.d2q9_channel_reference(100, 40, storage=F32)
    out.append(("channel_ref_100x40_f32", cfg, t, (0.05, 0.0, 0.0)))
    cfg, t = cases.d2q9_lid_reference(8, storage=F32)
    cfg = dict(cfg, ny=6)
    out.append(("lid_ref_8x6_f32", cfg, geo.lid_cavity_2d_reference(8, 6), (0.0, 0.0, 0.0)))
    return out


CASES = case_list()


def run(cfg, types, u0, nsteps):
    with Lattice(cfg, types) as lat:
        lat.init_equilibrium(1.0, u0, 0x5EED0007, 1e-3)
        lat.step(nsteps)
        return lat.state(), lat.steps_done


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_multistep_equals_single_steps(case, monkeypatch):
    name, cfg, types, u0 = case
    n = 3 * 8 + 5  # several launches of every shape plus a remainder of single steps
    monkeypatch.setenv("LBM_MULTISTEP", "off")
    ref, done = run(cfg, types, u0, n)
    assert done == n
    for v in (SHAPES_F64 if cfg["storage"] == F64 else SHAPES):
        monkeypatch.setenv("LBM_MULTISTEP", str(v))
        got, done = run(cfg, types, u0, n)
        assert done == n
        assert np.array_equal(got, ref), f"{name}: shape {v} (K = {SHAPE_K[v]}) differs"


@pytest.mark.parametrize("v", SHAPES)
def test_multistep_fp32_equals_oracle(v, monkeypatch):
    cfg, t = cases.d2q9_karman(130, 40, 30, 20, 5, u_max=0.1, re=40.0, storage=F32)
    types = t.reshape(-1)
    rng = np.random.default_rng(11)
    ux = 0.05 + 1e-3 * rng.standard_normal(types.size)
    uy = 1e-3 * rng.standard_normal(types.size)
    monkeypatch.setenv("LBM_MULTISTEP", str(v))
    with Lattice(cfg, types) as lat:
        lat.init_fields(None, ux, uy, None)
        lat.step(41)
        got = lat.populations()
    want = orc.run(cfg, types, orc.init_fields(cfg, types, None, ux, uy, None), 41)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("v", SHAPES_F64)
def test_multistep_fp64_equals_oracle(v, monkeypatch):
    cfg, t = cases.d2q9_karman(130, 40, 30, 20, 5, u_max=0.1, re=40.0, storage=F64)
    types = t.reshape(-1)
    rng = np.random.default_rng(12)
    ux = 0.05 + 1e-3 * rng.standard_normal(types.size)
    uy = 1e-3 * rng.standard_normal(types.size)
    monkeypatch.setenv("LBM_MULTISTEP", str(v))
    with Lattice(cfg, types) as lat:
        assert lat.steps_per_launch == SHAPE_K[v]
        lat.init_fields(None, ux, uy, None)
        lat.step(41)
        got = lat.populations()
    want = orc.run(cfg, types, orc.init_fields(cfg, types, None, ux, uy, None), 41)
    assert np.array_equal(got, want)


def test_fp64_default_shape_and_double_copy_shapes_fall_back(monkeypatch):
    cfg, t = cases.d2q9_lid_reference(64, storage=F64)
    monkeypatch.delenv("LBM_MULTISTEP", raising=False)
    with Lattice(cfg, t.reshape(-1)) as lat:
        assert lat.steps_per_launch == 4  # the one-copy default
    monkeypatch.setenv("LBM_MULTISTEP", "9")  # two FP64 copies would not fit: single steps
    with Lattice(cfg, t.reshape(-1)) as lat:
        assert lat.steps_per_launch == 1


def test_wide_lattice_falls_back_to_single_steps_and_matches_oracle(monkeypatch):
    """Multi-step tiles pack x | y << 16: a side of 65536 or more takes single steps, on the same
    trajectory as the oracle."""
    monkeypatch.delenv("LBM_MULTISTEP", raising=False)
    cfg, _ = cases.d2q9_lid_reference(8, storage=F32)
    cfg = dict(cfg, nx=70000, ny=4)
    types = geo.lid_cavity_2d_reference(70000, 4).reshape(-1)
    rng = np.random.default_rng(13)
    ux = 1e-3 * rng.standard_normal(types.size)
    uy = 1e-3 * rng.standard_normal(types.size)
    with Lattice(cfg, types) as lat:
        assert lat.steps_per_launch == 1
        lat.init_fields(None, ux, uy, None)
        lat.step(7)
        got = lat.populations()
    want = orc.run(cfg, types, orc.init_fields(cfg, types, None, ux, uy, None), 7)
    assert np.array_equal(got, want)
