"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the same inputs.

Tolerances: FP64 within 1e-12 relative L2 on rho/u and populations; FP32 storage
against the FP32 oracle within 1e-5 (same operation order, usually bit-identical); shifted
FP16/BF16 storage within 1e-3 relative L2 on rho/u against the FP64 oracle.
"""
import numpy as np
import pytest

import oracle as orc
from lbm_amd import Lattice, cases, geometry as geo
from lbm_amd.capi import (BC_PHYSICAL, BF16S, BGK, D3Q27, F16S, F32, F64, MRT_GS, MRT_MFMA_BF16X3, MRT_MFMA_F32,
                          MRT_MFMA_FUSED, MRT_RAW, MRT_SCALAR)

pytestmark = pytest.mark.gpu


def lid_phys_64(st, collision=BGK, engine=MRT_SCALAR):
    """A 64 x 20 x 12 PHYSICAL lid cavity: rows a whole number of 64-node wavefronts (the 64 x 4
    tiles), walls across the rows (x = 0 and x = 63) and along them, the lid at z = 11.  Its
    wavefronts at x = 0..63 hold solid and fluid lanes, its wall rows and planes solid lanes only."""
    om = cases.omega_from_nu(0.05 * 62 / 100.0)
    cfg = dict(lattice=D3Q27, nx=64, ny=20, nz=12, collision=collision, storage=st, bc_mode=BC_PHYSICAL,
               omega=om, u_wall=(0.05, 0.0, 0.0))
    if collision != BGK:
        cfg.update(s=cases.shear_mrt_rates(om), mrt_engine=engine)
    return cfg, geo.lid_cavity_3d_physical(64, 20, 12)


def rel_l2(a, b):
    b = np.asarray(b)
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


def noisy_fields(types, Q, seed, u0=(0.0, 0.0, 0.0), amp=1e-3):
    rng = np.random.default_rng(seed)
    N = types.size
    u = np.zeros((3, N))
    for a in range(2 if Q == 9 else 3):
        u[a] = u0[a] + amp * rng.standard_normal(N)
    return 1.0 + 0.1 * amp * rng.standard_normal(N), u


def run_both(cfg, types, steps, seed=0, u0=(0.0, 0.0, 0.0), amp=1e-3, storage=None):
    """Same initial fields for kernel and oracle; returns (gpu f, oracle f, gpu fields, oracle fields)."""
    types = np.ascontiguousarray(types, dtype=np.uint8).reshape(-1)
    rho, u = noisy_fields(types, cfg["lattice"], seed, u0, amp)
    with Lattice(cfg, types) as lat:
        lat.init_fields(rho, *u)
        lat.step(steps)
        fg = lat.populations()
        flds_g = lat.fields()
    f0 = orc.init_fields(cfg, types, rho, *u)
    fo = orc.run(cfg, types, f0, steps)
    flds_o = orc.fields(cfg, types, fo)
    return fg, fo, flds_g, flds_o


PARITY = [
    ("d2q9_lid_ref_128", lambda st: cases.d2q9_lid_reference(128, storage=st), 1000),
    ("d2q9_channel_ref_300x100", lambda st: cases.d2q9_channel_reference(300, 100, storage=st), 500),
    ("d2q9_karman_phys_256x64", lambda st: cases.d2q9_karman(256, 64, 64, 32, 6, re=30.0, storage=st), 500),
    ("d2q9_lid_phys_64", lambda st: cases.d2q9_lid_physical(64, storage=st), 500),
    ("d3q27_bgk_lid_ref_24", lambda st: cases.d3q27_bgk_lid_reference(24, u_lid=0.05, re=100.0, storage=st), 100),
    ("d3q27_mrt_raw_lid_ref_20", lambda st: cases.d3q27_mrt_lid_reference(20, storage=st, engine=MRT_SCALAR), 60),
    ("d3q27_mrt_gs_lid_ref_20", lambda st: cases.d3q27_mrt_lid_gs(20, re=100.0, storage=st, engine=MRT_SCALAR), 60),
    ("d3q27_sphere_phys_40x20x20", lambda st: cases.d3q27_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0, storage=st), 100),
    ("d3q27_sphere_mrt_phys_40x20x20", lambda st: cases.d3q27_sphere_channel(
        40, 20, 20, (12, 10, 10), 4, re=10.0, storage=st, collision=MRT_GS), 60),
    ("d3q27_bgk_channel_ref_40x20x20", lambda st: cases.d3q27_bgk_channel_reference(40, 20, 20, re=20.0, storage=st), 100),
    ("d3q27_mrt_channel_ref_48x20x20", lambda st: cases.d3q27_mrt_channel_reference(48, 20, 20, storage=st), 60),
    ("d3q27_suga_gs_phys_40x18x18", lambda st: cases.d3q27_mrt_channel_suga(40, 18, 18, storage=st), 60),
    ("d3q27_suga_raw_phys_40x18x18", lambda st: cases.d3q27_mrt_channel_suga(
        40, 18, 18, storage=st, collision=MRT_RAW), 60),
    ("d3q19_bgk_lid_ref_20", lambda st: cases.d3q19_lid(20, re=100.0, storage=st), 100),
    ("d3q19_mrt_lid_ref_20", lambda st: cases.d3q19_lid(20, re=100.0, storage=st, collision=MRT_GS), 60),
    ("d3q19_sphere_phys_40x20x20", lambda st: cases.d3q19_sphere_channel(
        40, 20, 20, (12, 10, 10), 4, re=10.0, storage=st), 100),
    ("d3q19_sphere_mrt_phys_40x20x20", lambda st: cases.d3q19_sphere_channel(
        40, 20, 20, (12, 10, 10), 4, re=10.0, storage=st, collision=MRT_GS), 60),
    ("d3q27_bgk_lid_phys_64x20x12", lambda st: lid_phys_64(st), 60),
    ("d3q27_mrt_gs_lid_phys_64x20x12", lambda st: lid_phys_64(st, MRT_GS), 40),
]


@pytest.mark.parametrize("name,make,steps", PARITY, ids=[p[0] for p in PARITY])
def test_fp64_parity(name, make, steps):
    cfg, types = make(F64)
    cfg = dict(cfg, mrt_engine=MRT_SCALAR) if cfg.get("collision", 0) else cfg
    fg, fo, flg, flo = run_both(cfg, types, steps, seed=11)
    e = rel_l2(fg, fo)
    print(f"{name}: FP64 populations rel-L2 {e:.3e}, bitwise {np.array_equal(fg, fo)}")
    assert e <= 1e-12
    # rho relative to rho; each velocity component relative to the whole velocity field (a
    # component that is nearly zero, as u_y in a cavity driven along x, carries only rounding)
    D = 2 if cfg["lattice"] == 9 else 3
    assert rel_l2(flg[0], flo[0]) <= 1e-12
    u_norm = np.linalg.norm(np.stack(flo[1:1 + D]))
    for k in range(1, 1 + D):
        assert np.linalg.norm(flg[k] - flo[k]) <= 1e-12 * u_norm, k


@pytest.mark.parametrize("name,make,steps", PARITY, ids=[p[0] for p in PARITY])
def test_fp32_parity(name, make, steps):
    cfg, types = make(F32)
    cfg = dict(cfg, mrt_engine=MRT_SCALAR) if cfg.get("collision", 0) else cfg
    fg, fo, flg, flo = run_both(cfg, types, steps, seed=12)
    e = rel_l2(fg, fo)
    print(f"{name}: FP32 populations rel-L2 {e:.3e}, bitwise {np.array_equal(fg, fo)}")
    assert e <= 1e-5


# (name, case, steps, bound for F16S, bound for BF16S): flows started from rest (lid) or from a
# fluid at rest with the inlet switched on (channel), compared once they have developed.
SHIFTED = [
    ("d2q9_lid_ref_128", lambda st: cases.d2q9_lid_reference(128, storage=st), 5000, 1e-3, 5e-3),
    ("d3q27_bgk_lid_ref_32", lambda st: cases.d3q27_bgk_lid_reference(32, u_lid=0.05, re=100.0, storage=st),
     1000, 1e-3, 5e-3),
    ("d3q27_mrt_gs_lid_ref_24", lambda st: cases.d3q27_mrt_lid_gs(24, re=100.0, storage=st, engine=MRT_SCALAR),
     600, 1e-3, 5e-3),
    ("d3q27_mrt_gs_lid_ref_24_mfma2", lambda st: cases.d3q27_mrt_lid_gs(24, re=100.0, storage=st,
                                                                         engine=MRT_MFMA_F32), 600, 1e-3, 5e-3),
    ("d3q27_mrt_gs_lid_ref_24_fused", lambda st: cases.d3q27_mrt_lid_gs(24, re=100.0, storage=st,
                                                                         engine=MRT_MFMA_FUSED), 600, 1e-3, 5e-3),
    ("d3q27_mrt_gs_lid_ref_24_bf16x3", lambda st: cases.d3q27_mrt_lid_gs(24, re=100.0, storage=st,
                                                                          engine=MRT_MFMA_BF16X3), 600, 1e-3, 5e-3),
    ("d2q9_channel_phys_256x64", lambda st: cases.d2q9_karman(256, 64, 64, 32, 6, re=20.0, storage=st), 3000,
     1e-3, 5e-3),
    ("d3q19_mrt_lid_ref_24", lambda st: cases.d3q19_lid(24, re=100.0, storage=st, collision=MRT_GS), 600,
     1e-3, 5e-3),
]


def vector_errors(a, b):
    """(rel-L2 of rho, rel-L2 of the velocity vector field) of fields a against fields b."""
    e_rho = rel_l2(a[0], b[0])
    ua, ub = np.stack(a[1:4]), np.stack(b[1:4])
    return e_rho, float(np.linalg.norm(ua - ub) / np.linalg.norm(ub))


@pytest.mark.parametrize("storage", [F16S, BF16S], ids=["f16s", "bf16s"])
@pytest.mark.parametrize("name,make,steps,b16,bbf", SHIFTED, ids=[p[0] for p in SHIFTED])
def test_shifted_storage_against_fp64(name, make, steps, b16, bbf, storage):
    cfg16, types = make(storage)
    cfg64 = dict(cfg16, storage=F64)
    if cfg64.get("collision", 0) and cfg64.get("mrt_engine", 0) != MRT_SCALAR:
        cfg64["mrt_engine"] = MRT_SCALAR
    types = np.ascontiguousarray(types, dtype=np.uint8).reshape(-1)
    out = {}
    for key, cfg in (("s", cfg16), ("d", cfg64)):
        with Lattice(cfg, types) as lat:
            lat.init_fields()
            lat.step(steps)
            out[key] = lat.fields()
    e_rho, e_u = vector_errors(out["s"], out["d"])
    print(f"{name} storage {storage}: rel-L2 vs FP64 rho {e_rho:.2e}, u {e_u:.2e}")
    bound = b16 if storage == F16S else bbf
    assert e_rho <= bound and e_u <= bound


# BGK in shifted storage: the kernels compute on g = f - w_i and store with stochastic rounding;
# the oracle restates both (the same FP32 operations, the same random bits per node, population and
# step, the conversions restated bit for bit), so the populations agree bit for bit -- D2Q9 through
# the multi-step kernel (and its boundary list), the 3D lattices through the single-step kernel.
SHIFTED_BGK = [
    ("d2q9_karman_phys_128x48", lambda st: cases.d2q9_karman(128, 48, 32, 24, 5, re=30.0, storage=st), 300),
    ("d2q9_lid_ref_64", lambda st: cases.d2q9_lid_reference(64, storage=st), 301),
    ("d3q27_bgk_lid_ref_20", lambda st: cases.d3q27_bgk_lid_reference(20, u_lid=0.05, re=100.0, storage=st), 60),
    ("d3q27_sphere_phys_40x20x20", lambda st: cases.d3q27_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0,
                                                                          storage=st), 60),
    ("d3q19_sphere_phys_40x20x20", lambda st: cases.d3q19_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0,
                                                                          storage=st), 60),
    ("d3q27_bgk_lid_phys_64x20x12", lambda st: lid_phys_64(st), 60),
]


@pytest.mark.parametrize("storage", [F16S, BF16S], ids=["f16s", "bf16s"])
@pytest.mark.parametrize("name,make,steps", SHIFTED_BGK, ids=[p[0] for p in SHIFTED_BGK])
def test_shifted_storage_matches_oracle_emulation(name, make, steps, storage):
    cfg, types = make(storage)
    fg, fo, flg, flo = run_both(cfg, types, steps, seed=14, u0=(0.05, 0, 0))
    e = rel_l2(fg, fo)
    print(f"{name} storage {storage}: kernel vs oracle rel-L2 {e:.3e}, bitwise {np.array_equal(fg, fo)}")
    assert np.array_equal(fg, fo)
    for a, b in zip(flg, flo):  # fields: rho = 1 + drho in FP64, u in FP32, as the oracle
        assert np.array_equal(a, b)


# D3Q27 MRT in shifted storage, every engine, against the oracle's run of the same algorithm
# (arithmetic on g = f - w_i, stochastic rounding with the same random bits).  The engines sum the
# transforms in other orders than the oracle's dense M g (tensor-product passes, MFMA
# accumulation), so where the FP32 results differ in their last bits a stored 2-byte value can
# round to the other neighbour; each such flip moves one population by one storage ulp (2^-11 of
# |f - w_i| for F16S, 2^-8 for BF16S), so the difference grows to the size of the storage rounding
# itself: measured 8.6e-6 to 3.3e-5 (F16S) and 6.8e-5 to 2.6e-4 (BF16S) after 60 steps, against
# 2.2e-5 to 2.9e-4 (F16S) and 1.8e-4 to 2.4e-3 (BF16S) from the FP64 oracle.  FP32 storage of the
# same engines agrees with FP64 to 2e-6 (test_mrt_mfma_against_fp64).
MRT_SHIFTED = [
    ("d3q27_mrt_gs_lid_ref_20", lambda st, en: cases.d3q27_mrt_lid_gs(20, re=100.0, storage=st, engine=en), 60),
    ("d3q27_mrt_raw_lid_ref_20", lambda st, en: cases.d3q27_mrt_lid_reference(20, storage=st, engine=en), 60),
    ("d3q27_sphere_mrt_phys_40x20x20", lambda st, en: (
        dict(cases.d3q27_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0, storage=st, collision=MRT_GS)[0],
             mrt_engine=en),
        cases.d3q27_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0, storage=st, collision=MRT_GS)[1]), 60),
    ("d3q27_mrt_gs_lid_phys_64x20x12", lambda st, en: lid_phys_64(st, MRT_GS, en), 40),
]


@pytest.mark.parametrize("engine", [MRT_SCALAR, MRT_MFMA_F32, MRT_MFMA_FUSED, MRT_MFMA_BF16X3],
                         ids=["scalar", "mfma2", "fused", "bf16x3"])
@pytest.mark.parametrize("storage", [F16S, BF16S], ids=["f16s", "bf16s"])
@pytest.mark.parametrize("name,make,steps", MRT_SHIFTED, ids=[p[0] for p in MRT_SHIFTED])
def test_mrt_shifted_storage_matches_oracle_emulation(name, make, steps, storage, engine):
    cfg, types = make(storage, engine)
    assert cfg["mrt_engine"] == engine
    fg, fo, flg, flo = run_both(cfg, types, steps, seed=16, u0=(0.02, 0, 0))
    e = rel_l2(fg, fo)
    types = np.ascontiguousarray(types, dtype=np.uint8).reshape(-1)
    rho, u = noisy_fields(types, 27, 16, (0.02, 0, 0))  # the initial fields run_both used
    cfg64 = dict(cfg, storage=F64, mrt_engine=MRT_SCALAR)
    f64 = orc.run(cfg64, types, orc.init_fields(cfg64, types, rho, *u), steps)
    e64 = rel_l2(fg, f64)
    print(f"{name} storage {storage} engine {engine}: vs oracle codec emulation {e:.2e}, vs FP64 oracle {e64:.2e}")
    assert e <= (5e-5 if storage == F16S else 5e-4)
    assert e64 <= (5e-4 if storage == F16S else 3e-3)


def test_init_equilibrium_noise_generator():
    """lbm_init_equilibrium draws SplitMix64 noise exactly as its contract states."""
    cfg, types = cases.d2q9_karman(64, 32, 16, 16, 4, storage=F64)
    types = types.reshape(-1)
    N = types.size
    mask = 2 ** 64 - 1
    s = 0x5EED0002
    u = np.zeros((3, N))
    u[0] = 0.1
    for n in range(N):
        for a in range(2):
            s = (s + 0x9E3779B97F4A7C15) & mask
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
            z ^= z >> 31
            r = (z >> 11) * 2.0 ** -53
            if not (types[n] & 1):
                u[a, n] += 1e-4 * (2 * r - 1)
    with Lattice(cfg, types) as a, Lattice(cfg, types) as b:
        a.init_equilibrium(1.0, (0.1, 0.0, 0.0), 0x5EED0002, 1e-4)
        b.init_fields(np.ones(N), u[0], u[1])
        assert np.array_equal(a.populations(), b.populations())


def test_monitors_match_host_sums():
    cfg, types = cases.d3q27_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0, storage=F64)
    types = types.reshape(-1)
    with Lattice(cfg, types) as lat:
        lat.init_equilibrium(1.0, (0.05, 0, 0), 7, 1e-3)
        lat.step(20)
        mass, ke, bad = lat.monitors()
        f = lat.populations().reshape(27, -1)
    fluid = (types & 1) == 0
    c, _, _ = orc.tables(27)
    rho = f[:, fluid].sum(0)
    j = c.astype(float) @ f[:, fluid]
    assert bad == 0
    assert abs(mass - rho.sum()) <= 1e-12 * rho.sum()
    assert abs(ke - ((j * j).sum(0) / (2 * rho)).sum()) <= 1e-10 * ke


def test_monitors_flag_nonfinite():
    cfg, types = cases.d2q9_lid_reference(32, storage=F32)
    with Lattice(cfg, types) as lat:
        lat.init_reference(1.0)
        f = lat.populations()
        f[5 + 32 * 10 + 1024 * 3] = np.nan  # a fluid node
        lat.set_populations(f)
        _, _, bad = lat.monitors()
    assert bad == 1


@pytest.mark.parametrize("storage", [F64, F32, F16S, BF16S])
def test_checkpoint_resume_is_bitwise(storage):
    cfg, types = cases.d2q9_karman(128, 48, 32, 24, 5, re=30.0, storage=storage)
    with Lattice(cfg, types) as a:
        a.init_equilibrium(1.0, (0.1, 0, 0), 3, 1e-4)
        a.step(37)
        snap = a.state()
        a.step(41)
        ref = a.state()
    with Lattice(cfg, types) as b:
        b.set_state(snap)
        b.step(41)
        assert np.array_equal(b.state(), ref)


def test_mrt_reference_init_matches_oracle():
    cfg, types = cases.d3q27_mrt_lid_reference(16, storage=F64, engine=MRT_SCALAR)
    types = types.reshape(-1)
    with Lattice(cfg, types) as lat:
        lat.init_reference(1.0)
        g = lat.populations()
    assert np.array_equal(g, orc.init_reference(cfg, types, 1.0))


@pytest.mark.parametrize("sign", [0, 0x80000000])
def test_fp32_reciprocal_equals_division_exhaustively(sign):
    """The FP32 kernels take 1/rho as the hardware reciprocal plus one Newton step; over every
    normal float with a normal reciprocal (2^-126 <= |x| < 2^125) it equals IEEE 1.0f / x."""
    from lbm_amd.capi import check_reciprocal
    assert check_reciprocal(sign | 0x00800000, sign | 0x7E000000) == 0


@pytest.mark.parametrize("storage", [F64, F32, F16S, BF16S])
@pytest.mark.parametrize("q", [9, 19, 27])
def test_bgk_reference_init_matches_oracle(q, storage):
    """f = f_eq(rho0, 0) = w_i rho0 in every storage (exact: the shifted codecs store 0 for it),
    and a few steps from it stay on the oracle's trajectory."""
    if q == 9:
        cfg, types = cases.d2q9_lid_reference(24, storage=storage)
    elif q == 19:
        cfg, types = cases.d3q19_lid(12, u_lid=0.05, re=50.0, storage=storage)
    else:
        cfg, types = cases.d3q27_bgk_lid_reference(12, u_lid=0.05, re=50.0, storage=storage)
    types = types.reshape(-1)
    want = orc.init_reference(cfg, types, 1.0)
    with Lattice(cfg, types) as lat:
        lat.init_reference(1.0)
        got = lat.populations()
        assert np.array_equal(got, want)
        lat.step(5)
        got = lat.populations()
    ref = orc.run(cfg, types, want, 5)
    assert np.all(np.isfinite(got))
    tol = {F64: 1e-12, F32: 1e-6, F16S: 1e-3, BF16S: 1e-2}[storage]
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)


MFMA_CASES = [
    ("d3q27_mrt_raw_lid_ref_20", lambda st, en: cases.d3q27_mrt_lid_reference(20, storage=st, engine=en), 60),
    ("d3q27_mrt_gs_lid_ref_20", lambda st, en: cases.d3q27_mrt_lid_gs(20, re=100.0, storage=st, engine=en), 60),
    ("d3q27_mrt_gs_lid_ref_70x20x12", lambda st, en: (
        dict(cases.d3q27_mrt_lid_gs(20, re=100.0, storage=st, engine=en)[0], nx=70, ny=20, nz=12),
        geo.lid_cavity_3d_reference(70, 20, 12)), 40),
    ("d3q27_sphere_mrt_phys_40x20x20", lambda st, en: cases.d3q27_sphere_channel(
        40, 20, 20, (12, 10, 10), 4, re=10.0, storage=st, collision=MRT_GS), 60),
    ("d3q27_suga_gs_phys_40x18x18", lambda st, en: cases.d3q27_mrt_channel_suga(40, 18, 18, storage=st), 60),
    ("d3q27_mrt_channel_ref_48x20x20", lambda st, en: cases.d3q27_mrt_channel_reference(
        48, 20, 20, storage=st, engine=en), 60),
    ("d3q27_mrt_gs_lid_phys_64x20x12", lambda st, en: lid_phys_64(st, MRT_GS, en), 40),
]


@pytest.mark.parametrize("name,make,steps", MFMA_CASES, ids=[p[0] for p in MFMA_CASES])
def test_mrt_mfma_against_fp64(name, make, steps):
    """The MFMA engines (two-stage FP32, fused-operator FP32, BF16x3 split) and the FP32
    scalar-FMA variant, each against the FP64 oracle."""
    errs = {}
    for en in (MRT_MFMA_F32, MRT_MFMA_FUSED, MRT_MFMA_BF16X3, MRT_SCALAR):
        cfg, types = make(F32, en)
        cfg = dict(cfg, mrt_engine=en)
        types = np.ascontiguousarray(types, dtype=np.uint8).reshape(-1)
        rho, u = noisy_fields(types, 27, 15, (0.02, 0, 0))
        with Lattice(cfg, types) as lat:
            lat.init_fields(rho, *u)
            lat.step(steps)
            fg = lat.populations()
        cfg64 = dict(cfg, storage=F64, mrt_engine=MRT_SCALAR)
        fo = orc.run(cfg64, types, orc.init_fields(cfg64, types, rho, *u), steps)
        errs[en] = rel_l2(fg, fo)
    print(f"{name}: rel-L2 vs FP64 oracle: MFMA {errs[MRT_MFMA_F32]:.2e}, fused MFMA "
          f"{errs[MRT_MFMA_FUSED]:.2e}, BF16x3 {errs[MRT_MFMA_BF16X3]:.2e}, scalar {errs[MRT_SCALAR]:.2e}")
    assert max(errs.values()) <= 1e-5


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_force_closed_form_on_device(Q):
    """An isolated solid node in a uniform equilibrium flow: F = 2 rho u exactly (FP64)."""
    shape = (1, 9, 11) if Q == 9 else (7, 8, 9)
    nz, ny, nx = shape
    types = np.zeros(shape, np.uint8)
    types[nz // 2, ny // 2, nx // 2] = 1
    cfg = dict(lattice=Q, nx=nx, ny=ny, nz=nz, collision=0, storage=F64, bc_mode=1, omega=1.0)
    u = (0.03, -0.02, 0.01 if Q != 9 else 0.0)
    rho = 1.1
    n = nx * ny * nz
    with Lattice(cfg, types) as lat:
        lat.init_fields(np.full(n, rho), *[np.full(n, v) for v in u])
        got = lat.force()
    assert np.allclose(got, 2 * rho * np.array(u), rtol=0, atol=1e-14)


@pytest.mark.parametrize("storage", [F64, F32, F16S])
def test_force_matches_restatement(storage):
    """Kernel force on a developed sphere flow equals the numpy restatement applied to the same
    populations (only the summation order differs)."""
    from np_reference import force as np_force
    cfg, types = cases.d3q27_sphere_channel(48, 24, 24, (14, 12, 12), 4, re=10.0, storage=storage)
    box = (9, 20, 7, 18, 7, 18)
    with Lattice(cfg, types) as lat:
        lat.init_equilibrium(1.0, (0.05, 0.0, 0.0), 7, 1e-4)
        lat.step(300)
        got = lat.force(box)
        f = lat.populations().reshape(27, 24, 24, 48)
    want = np_force(f, types.reshape(24, 24, 48), cfg, box)
    print(f"force {got} restatement {want}")
    assert abs(got[0]) > 1e-4
    assert np.allclose(got, want, rtol=1e-12, atol=1e-15)


# The step kernels with non-temporal population stores run on 3D lattices over 256 MiB per buffer
# (the benchmark's); LBM_NT_STORE=1 forces them on the small parity lattices.  Only the store kind
# differs, so the populations must be bit-identical to the ordinary kernels' and, in FP64, to the
# oracle's.
NT_CASES = [c for c in PARITY if c[0].startswith("d3q")]


@pytest.mark.parametrize("storage", [F64, F32, F16S, BF16S])
@pytest.mark.parametrize("name,make,steps", NT_CASES, ids=[p[0] for p in NT_CASES])
def test_nontemporal_store_kernels_are_bitwise(monkeypatch, name, make, steps, storage):
    cfg, types = make(storage)
    cfg = dict(cfg, mrt_engine=MRT_SCALAR) if cfg.get("collision", 0) else cfg
    types = np.ascontiguousarray(types, dtype=np.uint8).reshape(-1)
    rho, u = noisy_fields(types, cfg["lattice"], 21)
    out = {}
    for nt in ("0", "1"):
        monkeypatch.setenv("LBM_NT_STORE", nt)
        with Lattice(cfg, types) as lat:
            lat.init_fields(rho, *u)
            lat.step(min(steps, 40))
            out[nt] = lat.populations()
    assert np.array_equal(out["0"], out["1"]), name
    if storage == F64:
        fo = orc.run(cfg, types, orc.init_fields(cfg, types, rho, *u), min(steps, 40))
        assert rel_l2(out["1"], fo) <= 1e-12


@pytest.mark.parametrize("storage", [F64, F32])
def test_reference_3d_channel_from_program_start(storage):
    """The reference's 3D MRT channel started from its program's initial fields (rho = 1,
    u = (0.1, 0, 0), both zero at solid nodes): kernel against oracle after 100 steps."""
    cfg, types = cases.d3q27_mrt_channel_reference(48, 20, 20, storage=storage, engine=MRT_SCALAR)
    types = types.reshape(-1)
    start = cases.reference_channel_start(types, cfg["u_in"])
    with Lattice(cfg, types) as lat:
        lat.init_fields(*start)
        lat.step(100)
        fg = lat.populations()
    cfg64 = dict(cfg, storage=F64)
    fo = orc.run(cfg64, types, orc.init_fields(cfg64, types, *start), 100)
    assert rel_l2(fg, fo) <= (1e-12 if storage == F64 else 1e-5)
