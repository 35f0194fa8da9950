"""Pins the CPU oracle (oracle/lbm_oracle.cpp) before anything is checked against it.

Each check is chosen so that a plausible mistake (wrong sign of c, wrong opposite map, wrong
weight, dropped override, transposed moment matrix, wrong rounding) fails one of them:
closed forms of the lattice, brute-force streaming permutations, an independent numpy
restatement on tiny grids, exact MRT algebra, and codecs against numpy/torch conversions.
"""
import os
import re

import numpy as np
import pytest

import np_reference as npr
import oracle as orc
from lbm_amd import cases, geometry as geo
from lbm_amd.capi import BC_PHYSICAL, BC_REFERENCE, BGK, F16S, F32, F64, BF16S, MRT_GS, MRT_RAW

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_lattice_closed_forms(Q):
    c, w, opp = orc.tables(Q)
    cT = c.T.astype(float)
    assert abs(w.sum() - 1) < 1e-15
    assert np.abs(w @ cT).max() < 1e-16
    D = 2 if Q == 9 else 3
    second = np.einsum("i,ia,ib->ab", w, cT, cT)
    assert np.allclose(second[:D, :D], np.eye(D) / 3, atol=1e-16)
    # fourth-order isotropy: sum w c_a^4 = 1/3, sum w c_a^2 c_b^2 = 1/9
    assert abs(np.sum(w * cT[:, 0] ** 4) - 1 / 3) < 1e-15
    assert abs(np.sum(w * cT[:, 0] ** 2 * cT[:, 1] ** 2) - 1 / 9) < 1e-15
    for i in range(Q):
        assert (c[:, opp[i]] == -c[:, i]).all()
    # the oracle's tables are the reference's listing
    cr, wr, oppr = npr.lattice(Q)
    assert (cr.T == c).all() and np.allclose(wr, w, rtol=0, atol=0) and (oppr == opp).all()


@pytest.mark.parametrize("Q,shape", [(9, (1, 4, 5)), (19, (3, 4, 3)), (27, (3, 4, 3))])
def test_pure_streaming_is_pull_permutation(Q, shape):
    """omega = 0 in a fully periodic fluid box: the step is exactly the pull permutation
    f_i(x) <- f_i(x - c_i mod dims), checked slot by slot with distinct values."""
    nz, ny, nx = shape
    cfg = dict(lattice=Q, nx=nx, ny=ny, nz=nz, collision=BGK, storage=F64, bc_mode=BC_REFERENCE, omega=0.0)
    types = geo.periodic_box(nx, ny, nz).reshape(-1)
    N = nx * ny * nz
    f = np.arange(N * Q, dtype=np.float64) + 1.0
    out = orc.run(cfg, types, f, 1).reshape(Q, nz, ny, nx)
    F = f.reshape(Q, nz, ny, nx)
    c, _, _ = orc.tables(Q)
    for i in range(Q):
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    sx, sy, sz = (x - c[0, i]) % nx, (y - c[1, i]) % ny, (z - c[2, i]) % nz
                    assert out[i, z, y, x] == F[i, sz, sy, sx]
    assert out.sum() == F.sum()


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_equilibrium_moments(Q):
    nx, ny, nz = (6, 5, 1) if Q == 9 else (4, 3, 3)
    N = nx * ny * nz
    rng = np.random.default_rng(1)
    rho = 1 + 0.1 * rng.standard_normal(N)
    u = 0.05 * rng.standard_normal((3, N))
    if Q == 9:
        u[2] = 0
    cfg = dict(lattice=Q, nx=nx, ny=ny, nz=nz, storage=F64)
    f = orc.init_fields(cfg, np.zeros(N, np.uint8), rho, u[0], u[1], u[2]).reshape(Q, N)
    c, w, _ = orc.tables(Q)
    cf = c.astype(float)
    assert np.allclose(f.sum(0), rho, rtol=1e-14)
    assert np.allclose(cf @ f, rho * u, atol=1e-15)
    D = 2 if Q == 9 else 3
    for a in range(D):
        for b in range(D):
            Pab = (cf[a] * cf[b]) @ f
            assert np.allclose(Pab, rho * (u[a] * u[b] + (a == b) / 3), atol=1e-14)


@pytest.mark.parametrize("Q", [9, 19, 27])
def test_bgk_conserves_mass_and_momentum(Q):
    nx, ny, nz = (8, 6, 1) if Q == 9 else (5, 4, 4)
    N = nx * ny * nz
    rng = np.random.default_rng(2)
    u = 0.03 * rng.standard_normal((3, N))
    if Q == 9:
        u[2] = 0
    cfg = dict(lattice=Q, nx=nx, ny=ny, nz=nz, collision=BGK, storage=F64, bc_mode=BC_REFERENCE, omega=1.7)
    types = np.zeros(N, np.uint8)
    f0 = orc.init_fields(cfg, types, 1 + 0.01 * rng.standard_normal(N), u[0], u[1], u[2])
    f1 = orc.run(cfg, types, f0, 20)
    c, _, _ = orc.tables(Q)
    F0, F1 = f0.reshape(Q, N), f1.reshape(Q, N)
    assert abs(F1.sum() - F0.sum()) < 1e-12
    assert np.allclose(c.astype(float) @ F1.sum(1), c.astype(float) @ F0.sum(1), atol=1e-13)


REF_CASES = {
    "d2q9_lid": lambda: (dict(cases.d2q9_lid_reference(8, re=5.0, storage=F64)[0]),
                         geo.lid_cavity_2d_reference(8, 6)),
    "d2q9_channel": lambda: (dict(cases.d2q9_channel_reference(14, 10, storage=F64)[0]),
                             geo.channel_2d_reference(14, 10)),
    "d3q27_bgk_lid": lambda: (dict(cases.d3q27_bgk_lid_reference(5, nu=0.05, storage=F64)[0]),
                              geo.lid_cavity_3d_reference(5, 5, 5)),
    "d3q27_mrt_raw": lambda: (dict(cases.d3q27_mrt_lid_reference(5, storage=F64)[0]),
                              geo.lid_cavity_3d_reference(5, 5, 5)),
    "d3q27_mrt_gs": lambda: (dict(cases.d3q27_mrt_lid_gs(5, re=10.0, storage=F64)[0]),
                             geo.lid_cavity_3d_reference(5, 5, 5)),
    "d3q19_bgk_lid": lambda: (dict(cases.d3q19_lid(5, re=10.0, storage=F64)[0]),
                              geo.lid_cavity_3d_reference(5, 5, 5)),
    "d3q19_mrt_lid": lambda: (dict(cases.d3q19_lid(5, re=10.0, storage=F64, collision=MRT_GS)[0]),
                              geo.lid_cavity_3d_reference(5, 5, 5)),
    "d3q27_bgk_channel_ref": lambda: cases.d3q27_bgk_channel_reference(10, 7, 7, re=10.0, storage=F64),
    "d3q27_mrt_channel_ref": lambda: cases.d3q27_mrt_channel_reference(10, 8, 8, storage=F64),
    "d3q27_mrt_gs_channel_ref": lambda: cases.d3q27_mrt_channel_reference(10, 8, 8, storage=F64, collision=MRT_GS),
}


def _fix_dims(cfg, types):
    nz, ny, nx = types.shape
    cfg.update(nx=nx, ny=ny, nz=nz)
    return cfg


@pytest.mark.parametrize("name", list(REF_CASES))
def test_oracle_matches_numpy_restatement_reference_mode(name):
    cfg, types = REF_CASES[name]()
    cfg = _fix_dims(cfg, types)
    Q = cfg["lattice"]
    rng = np.random.default_rng(3)
    shape = types.shape
    u = np.zeros((3,) + shape)
    u[0] = 0.02 * rng.standard_normal(shape)
    u[1] = 0.02 * rng.standard_normal(shape)
    if Q != 9:
        u[2] = 0.02 * rng.standard_normal(shape)
    rho = 1 + 0.01 * rng.standard_normal(shape)
    F0 = npr.init_feq(shape, cfg, rho, u)
    f_or = orc.run(cfg, types.reshape(-1), F0.reshape(-1), 4)
    f_np = npr.run(F0, types, cfg, 4).reshape(-1)
    assert rel_l2(f_or, f_np) < 1e-13, rel_l2(f_or, f_np)


PHYS_CASES = {
    "d2q9_lid_phys": lambda: (dict(cases.d2q9_lid_physical(9, re=5.0, storage=F64)[0]),
                              geo.lid_cavity_2d_physical(9, 7)),
    "d2q9_karman": lambda: (dict(cases.d2q9_karman(24, 12, 7, 6, 2, storage=F64)[0]),
                            geo.channel_2d_physical(24, 12, 7, 6, 2)),
    "d3q27_sphere": lambda: (dict(cases.d3q27_sphere_channel(12, 7, 7, (4, 3, 3), 1, storage=F64)[0]),
                             geo.channel_3d_physical(12, 7, 7, 4, 3, 3, 1)),
    "d3q27_sphere_mrt": lambda: (dict(cases.d3q27_sphere_channel(12, 7, 7, (4, 3, 3), 1, storage=F64,
                                                                  collision=MRT_GS)[0]),
                                 geo.channel_3d_physical(12, 7, 7, 4, 3, 3, 1)),
    "d3q19_sphere": lambda: (dict(cases.d3q19_sphere_channel(12, 7, 7, (4, 3, 3), 1, storage=F64)[0]),
                             geo.channel_3d_physical(12, 7, 7, 4, 3, 3, 1)),
    "d3q19_sphere_mrt": lambda: (dict(cases.d3q19_sphere_channel(12, 7, 7, (4, 3, 3), 1, storage=F64,
                                                                  collision=MRT_GS)[0]),
                                 geo.channel_3d_physical(12, 7, 7, 4, 3, 3, 1)),
    # both inlet closures on every lattice
    "d2q9_karman_extrapolated_inlet": lambda: (dict(cases.d2q9_karman(24, 12, 7, 6, 2, storage=F64)[0],
                                                    inlet_closure=1), geo.channel_2d_physical(24, 12, 7, 6, 2)),
    "d3q27_sphere_zou_he_inlet": lambda: (dict(cases.d3q27_sphere_channel(12, 7, 7, (4, 3, 3), 1, storage=F64)[0],
                                               inlet_closure=0), geo.channel_3d_physical(12, 7, 7, 4, 3, 3, 1)),
    "d3q19_sphere_zou_he_inlet": lambda: (dict(cases.d3q19_sphere_channel(12, 7, 7, (4, 3, 3), 1, storage=F64)[0],
                                               inlet_closure=0), geo.channel_3d_physical(12, 7, 7, 4, 3, 3, 1)),
}


@pytest.mark.parametrize("name", list(PHYS_CASES))
def test_oracle_matches_numpy_restatement_physical_mode(name):
    cfg, types = PHYS_CASES[name]()
    cfg = _fix_dims(cfg, types)
    shape = types.shape
    rng = np.random.default_rng(4)
    u = np.zeros((3,) + shape)
    u[0] = 0.05 + 0.01 * rng.standard_normal(shape)
    u[1] = 0.01 * rng.standard_normal(shape)
    if cfg["lattice"] != 9:
        u[2] = 0.01 * rng.standard_normal(shape)
    F0 = npr.init_feq(shape, cfg, None, u)
    f_or = orc.run(cfg, types.reshape(-1), F0.reshape(-1), 5)
    f_np = npr.run(F0, types, cfg, 5).reshape(-1)
    assert rel_l2(f_or, f_np) < 1e-13, rel_l2(f_or, f_np)


@pytest.mark.parametrize("name", ["d3q27_sphere", "d3q27_sphere_mrt", "d3q19_sphere", "d3q27_sphere_zou_he_inlet",
                                  "d2q9_karman", "d2q9_lid_phys"])
def test_solid_nodes_feed_nothing_back(name):
    """PHYSICAL mode: the fluid nodes' values do not depend on what the solid nodes held before
    the run (garbage in every solid slot gives the same fluid state bit for bit), since every link
    into a solid node is bounced back from the node's own populations.  The solid slots then hold
    the rest equilibrium w_i in 3D lattices (the value the kernels store there with the fluid
    lanes) when their aligned block of 64 node indices holds a non-solid node, and keep their
    values otherwise and in 2D lattices."""
    cfg, types = PHYS_CASES[name]()
    cfg = _fix_dims(cfg, types)
    shape = types.shape
    Q = cfg["lattice"]
    _, w, _ = npr.lattice(Q)
    rng = np.random.default_rng(7)
    u = np.zeros((3,) + shape)
    u[0] = 0.05 + 0.01 * rng.standard_normal(shape)
    u[1] = 0.01 * rng.standard_normal(shape)
    if Q != 9:
        u[2] = 0.01 * rng.standard_normal(shape)
    F0 = npr.init_feq(shape, cfg, None, u).reshape(-1)
    t = types.reshape(-1)
    N = t.size
    solid = (t & 1) != 0
    assert solid.any() and (~solid).any()
    G0 = F0.reshape(Q, N).copy()
    G0[:, solid] = rng.uniform(-5, 5, (Q, int(solid.sum())))
    f = orc.run(cfg, t, F0, 3).reshape(Q, N)
    g = orc.run(cfg, t, G0.reshape(-1), 3).reshape(Q, N)
    assert np.array_equal(f[:, ~solid], g[:, ~solid])
    if Q >= 19:
        blk = np.array([not solid[b:b + 64].all() for b in range(0, N, 64)])[np.arange(N) // 64]
        rest = solid & blk
        assert rest.any() and (solid & ~blk).any()
        want = np.broadcast_to(np.asarray(w, dtype=float)[:, None], (Q, int(rest.sum())))
        assert np.array_equal(g[:, rest], want)
        assert np.array_equal(g[:, solid & ~blk], G0[:, solid & ~blk])
    else:
        assert np.array_equal(g[:, solid], G0[:, solid])


def _read_int_table(name, n=27):
    s = open(os.path.join(ROOT, "lbm-arrayfire_amd", "csrc", "mrt_tables.h")).read()
    body = re.search(name + r"\[%d\]\[%d\] = \{(.*?)\};" % (n, n), s, re.S).group(1)
    return np.array([[int(v) for v in r.split(",") if v.strip()] for r in re.findall(r"\{([^{}]*)\}", body)])


def test_mrt_matrices_exact():
    M, Mi = orc.mrt_matrices(1)
    assert np.array_equal(M, npr.raw_moment_matrix())
    assert np.allclose(M @ Mi, np.eye(27), atol=1e-13)
    # 24 M^-1 is an integer matrix; the kernels' compile-time table equals it
    assert np.abs(24 * Mi - np.round(24 * Mi)).max() < 1e-12
    assert np.array_equal(np.round(24 * Mi).astype(int), _read_int_table("MRAW_INV24"))
    assert np.array_equal(M.astype(int), _read_int_table("MRAW"))
    assert (M != 0).sum() == 339 and (np.round(24 * Mi) != 0).sum() == 226
    # Gram-Schmidt basis: orthogonal rows spanning the same row space, integer table is a row
    # scaling of the oracle's rows
    G, Gi = orc.mrt_matrices(2)
    gram = G @ G.T
    assert np.abs(gram - np.diag(np.diag(gram))).max() < 1e-10
    Gint = _read_int_table("MGS")
    for k in range(27):
        ratio = Gint[k] @ G[k] / (G[k] @ G[k])
        assert np.allclose(Gint[k], ratio * G[k], atol=1e-9)


def _mrt_cfg(s, meq_mode=0, coll=MRT_RAW):
    return dict(lattice=27, nx=4, ny=3, nz=3, collision=coll, storage=F64, bc_mode=BC_REFERENCE,
                meq_mode=meq_mode, s=list(s))


def test_mrt_zero_rates_is_pure_streaming():
    cfg = _mrt_cfg([0.0] * 27)
    types = np.zeros(36, np.uint8)
    rng = np.random.default_rng(5)
    f = 1 / 27 + 0.01 * rng.random(36 * 27)
    cfg_bgk = dict(cfg, collision=BGK, omega=0.0)
    assert rel_l2(orc.run(cfg, types, f, 2), orc.run(cfg_bgk, types, f, 2)) < 1e-14


def test_mrt_unit_rates_returns_equilibrium():
    """s = 1 with meq = M f_eq: f_out = M^-1 M f_eq(rho, u) = f_eq, i.e. BGK with omega = 1."""
    cfg = _mrt_cfg([1.0] * 27, meq_mode=1)
    types = np.zeros(36, np.uint8)
    rng = np.random.default_rng(6)
    u = 0.03 * rng.standard_normal((3, 36))
    f = orc.init_fields(cfg, types, 1 + 0.01 * rng.standard_normal(36), *u)
    f = f + 1e-3 * rng.standard_normal(f.size)
    out = orc.run(cfg, types, f, 1)
    assert rel_l2(out, orc.run(dict(cfg, collision=BGK, omega=1.0), types, f, 1)) < 1e-13


@pytest.mark.parametrize("basis", [MRT_RAW, MRT_GS])
def test_mrt_with_projected_meq_and_uniform_rate_is_bgk(basis):
    """With meq = M f_eq and every rate = omega, MRT is BGK: M^-1 (I - omega)(M f) + omega f_eq."""
    om = 1.37
    cfg = _mrt_cfg([om] * 27, meq_mode=1, coll=basis)
    types = np.zeros(36, np.uint8)
    rng = np.random.default_rng(7)
    f = orc.init_fields(cfg, types, 1 + 0.01 * rng.standard_normal(36), *(0.03 * rng.standard_normal((3, 36))))
    f = f + 1e-4 * rng.standard_normal(f.size)
    a = orc.run(cfg, types, f, 3)
    b = orc.run(dict(cfg, collision=BGK, omega=om), types, f, 3)
    assert rel_l2(a, b) < 1e-13


def test_reference_meq_differs_from_projection_only_in_row_26():
    M = npr.raw_moment_matrix()
    rng = np.random.default_rng(8)
    rho = 1 + 0.1 * rng.standard_normal(5)
    u = 0.05 * rng.standard_normal((3, 5))
    c, w, _ = npr.lattice(27)
    proj = M @ npr.feq(rho[None, None], u[:, None, None], c, w).reshape(27, -1)
    ref = npr.meq_reference(rho, u)
    assert np.allclose(proj[:26], ref[:26], atol=1e-14)
    # row 26: projection gives rho (1 + 3|u|^2) / 27; the reference table |u|^2 + rho/27
    usq = (u * u).sum(0)
    assert np.allclose(proj[26], rho * (1 + 3 * usq) / 27, atol=1e-15)
    assert np.allclose(ref[26], usq + rho / 27, atol=1e-15)


def test_mrt_reference_initial_state():
    """rho0/27 followed by one collision at rho0, u = 0 (rust/examples/d3q27_mrt_lid.rs:228-287)."""
    cfg = dict(cases.d3q27_mrt_lid_reference(4, storage=F64)[0])
    types = geo.lid_cavity_3d_reference(4, 4, 4).reshape(-1)
    f = orc.init_reference(cfg, types, 1.0).reshape(27, -1)
    M = npr.raw_moment_matrix()
    m = M @ np.full(27, 1 / 27)
    me = npr.meq_reference(np.array([1.0]), np.zeros((3, 1)))[:, 0]
    s = np.array(cfg["s"])
    expect = np.linalg.inv(M) @ (m - s * (m - me))
    assert np.allclose(f, expect[:, None], atol=1e-15)


def test_half_codec_matches_numpy():
    rng = np.random.default_rng(9)
    vals = np.concatenate([rng.standard_normal(3000) * 10.0 ** rng.integers(-9, 5, 3000),
                           [0.0, -0.0, 6.1e-5, 6.0e-8, 2.98e-8, 65504.0, 65519.0, 65520.0, 1e6, -1e-10],
                           # exact ties between half values
                           np.float16(np.arange(1, 200)).astype(np.float32) + np.float32(2.0 ** -11)])
    for v in vals.astype(np.float32):
        h = orc.f2h(v)
        assert h == np.float32(v).astype(np.float16).view(np.uint16), v
        assert orc.h2f(h) == np.float32(np.uint16(h).view(np.float16))


def test_bf16_codec_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(10)
    vals = np.concatenate([rng.standard_normal(3000) * 10.0 ** rng.integers(-30, 30, 3000),
                           [0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -3.0e-39]]).astype(np.float32)
    t = torch.from_numpy(vals).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    for v, ref in zip(vals, t):
        assert orc.f2bf(v) == ref, v
        assert orc.bf2f(ref) == np.float32((np.uint32(ref) << 16).view(np.float32))


@pytest.mark.parametrize("storage", [F16S, BF16S])
def test_shifted_storage_roundtrip_bound(storage):
    """decode(encode(f)) differs from f by at most half an ulp of (f - w_i) in the codec plus the
    float rounding of the shift."""
    cfg = dict(lattice=9, nx=6, ny=5, nz=1, storage=storage, collision=BGK, omega=0.0, bc_mode=BC_REFERENCE)
    types = np.zeros(30, np.uint8)
    rng = np.random.default_rng(11)
    u = 0.05 * rng.standard_normal((3, 30))
    f64 = orc.init_fields(dict(cfg, storage=F64), types, 1 + 0.02 * rng.standard_normal(30), *u)
    _, w, _ = orc.tables(9)
    g = f64 - np.repeat(w.astype(np.float32).astype(np.float64), 30)
    bits = 11 if storage == F16S else 8
    rt = orc.run(cfg, types, f64, 0)  # load encodes, store decodes
    err = np.abs(rt - f64)
    assert (err <= np.abs(g) * 2.0 ** -bits + 2.0 ** -24 * np.abs(f64) + 2.0 ** -24).all()
    assert err.max() > 0


def test_outlet_keeps_bgk_duct_stable_at_low_viscosity():
    """D3Q27 BGK duct from rest at tau = 0.536 with a uniform inlet: the outlet closure must not
    seed the instability that an NEBB closure with the neighbour's tangential velocity did."""
    from lbm_amd.capi import F32
    nx, ny, nz = 64, 24, 24
    cfg, _ = cases.d3q27_sphere_channel(nx, ny, nz, (16, 12, 12), 3, u_in=0.05, re=25.0, storage=F32)
    assert abs(cfg["omega"] - 1.8656716) < 1e-6
    types = geo.channel_3d_physical(nx, ny, nz, -100, -100, -100, 1).reshape(-1)
    f = orc.run(cfg, types, orc.init_fields(cfg, types), 600)
    assert np.isfinite(f).all()


def test_zou_he_inlet_recovers_prescribed_velocity():
    cfg, types = cases.d2q9_karman(32, 14, 10, 7, 2, storage=F64)
    cfg = dict(cfg)
    types = types.reshape(-1)
    rng = np.random.default_rng(12)
    u = np.zeros((3, types.size))
    u[0] = 0.05 + 0.01 * rng.standard_normal(types.size)
    u[1] = 0.01 * rng.standard_normal(types.size)
    f = orc.run(cfg, types, orc.init_fields(cfg, types, None, *u), 3)
    rho, ux, uy, _ = orc.fields(cfg, types, f)
    inlet = (types & 4) != 0
    y = np.arange(types.size) // 32
    H = 14 - 2
    expect = 4 * 0.1 * (y - 0.5) * (H - (y - 0.5)) / H ** 2
    assert np.abs(ux[inlet] - expect[inlet]).max() < 1e-14
    assert np.abs(uy[inlet]).max() < 1e-14


@pytest.mark.parametrize("Q", [9, 27])
def test_extrapolated_inlet_holds_prescribed_velocity_and_downstream_density(Q):
    """The extrapolated inlet node's moments are exactly u_in and the density of its downstream
    neighbour's previous populations (the non-equilibrium part carries neither mass nor
    momentum); a wall-adjacent inlet node included."""
    if Q == 9:
        cfg, types = cases.d2q9_karman(32, 14, 10, 7, 2, storage=F64)
    else:
        cfg, types = cases.d3q27_sphere_channel(16, 9, 9, (6, 4, 4), 1, storage=F64)
    cfg = dict(cfg, inlet_closure=1)
    types = types.reshape(-1)
    N = types.size
    rng = np.random.default_rng(13)
    u = np.zeros((3, N))
    u[0] = 0.05 + 0.01 * rng.standard_normal(N)
    u[1:3] = 0.01 * rng.standard_normal((2, N))
    f = orc.run(cfg, types, orc.init_fields(cfg, types, 1 + 0.01 * rng.standard_normal(N), *u), 3)
    rho, ux, uy, uz = orc.fields(cfg, types, f)
    inlet = np.flatnonzero((types & 4) != 0)
    assert inlet.size > 0
    down = f.reshape(Q, N)[:, inlet + 1].sum(0)
    nx, ny = cfg["nx"], cfg["ny"]
    y = (inlet // nx) % ny
    if Q == 9:
        H = ny - 2
        expect = 4 * 0.1 * (y - 0.5) * (H - (y - 0.5)) / H ** 2
    else:
        expect = np.full(inlet.size, cfg["u_in"])
    assert np.abs(rho[inlet] - down).max() < 1e-14
    assert np.abs(ux[inlet] - expect).max() < 1e-14
    assert np.abs(uy[inlet]).max() < 1e-14 and np.abs(uz[inlet]).max() < 1e-14


def test_extrapolated_inlet_keeps_sphere_duct_stable():
    """The 3D duct that the Zou-He inlet destabilises: 80 x 40 x 40, sphere r = 4, BGK at
    tau = 0.536 (u = 0.05, Re 33).  With the extrapolated inlet it stays bounded for 2,000 steps;
    the velocity never exceeds twice the inlet speed."""
    cfg, types = cases.d3q27_sphere_channel(80, 40, 40, (20, 20, 20), 4, re=33.333333333333336, storage=F32)
    assert abs(1 / cfg["omega"] - 0.536) < 1e-9 and cfg["inlet_closure"] == 1
    types = types.reshape(-1)
    N = types.size
    f = orc.init_fields(cfg, types, None, np.full(N, 0.05) * ((types & 1) == 0))
    f = orc.run(cfg, types, f, 2000)
    rho, ux, uy, uz = orc.fields(cfg, types, f)
    sp = np.sqrt(ux ** 2 + uy ** 2 + uz ** 2)
    assert np.isfinite(sp).all() and sp.max() < 0.1


def test_poiseuille_profile_develops():
    """Empty channel, parabolic velocity inlet, zero-gradient outlet at rho_bc, half-way walls:
    the flow becomes steady with a parabolic cross-section (walls at y = 0.5 and ny - 1.5), the
    mass flux sum_y rho u_x is the same at every section, and the outlet node holds rho_bc and
    the velocity of its upstream neighbour."""
    nx, ny = 60, 12
    cfg = dict(lattice=9, nx=nx, ny=ny, nz=1, collision=BGK, storage=F64, bc_mode=BC_PHYSICAL,
               omega=1.0, u_in=0.04, inlet_profile=1)
    types = geo.channel_2d_physical(nx, ny, -100, -100, 1).reshape(-1)
    f = orc.run(cfg, types, orc.init_fields(cfg, types), 4000)
    rho, ux, uy, _ = orc.fields(cfg, types, f)
    rho, ux = rho.reshape(ny, nx), ux.reshape(ny, nx)
    y = np.arange(1, ny - 1)
    H = ny - 2
    shape = (y - 0.5) * (H - (y - 0.5))
    for x in (30, nx - 10):
        prof = ux[1:-1, x]
        assert np.abs(prof / prof.max() - shape / shape.max()).max() < 2e-2
    flux = (rho[1:-1] * ux[1:-1]).sum(0)
    inflow = (4 * 0.04 * shape / H ** 2).sum() * rho[1:-1, 0].mean()
    assert np.abs(flux[1:nx // 2] / flux[1] - 1).max() < 1e-4
    assert np.abs(flux[1:] / flux[1] - 1).max() < 5e-3  # odd-even ripple next to the outlet
    assert abs(flux[1] / inflow - 1) < 2e-2
    # the outlet node holds rho_bc and its upstream neighbour's velocity
    assert np.abs(rho[1:-1, nx - 1] - 1.0).max() < 1e-14
    assert np.abs(ux[1:-1, nx - 1] - ux[1:-1, nx - 2]).max() < 1e-2 * ux.max()


def test_f32_index_compat_only_changes_large_lattices():
    cfg = dict(lattice=9, nx=8, ny=6, nz=1, collision=BGK, storage=F32, bc_mode=BC_REFERENCE, omega=1.2)
    types = geo.lid_cavity_2d_reference(8, 6).reshape(-1)
    f = orc.init_fields(cfg, types)
    a = orc.run(cfg, types, f, 3, f32_index_compat=False)
    b = orc.run(cfg, types, f, 3, f32_index_compat=True)
    assert np.array_equal(a, b)


def test_force_restatement_closed_form():
    """An isolated solid node in a uniform equilibrium flow: the links into it carry
    2 sum_{i>0} c_i f_eq,i = 2 rho u."""
    from np_reference import force, init_feq
    for Q, shape in ((9, (1, 7, 9)), (27, (5, 6, 7))):
        types = np.zeros(shape, np.uint8)
        types[shape[0] // 2, shape[1] // 2, shape[2] // 2] = 1
        u = np.array([0.03, -0.02, 0.01 if Q == 27 else 0.0])
        rho = 1.1
        F = init_feq(shape, dict(lattice=Q), np.full(shape, rho), np.broadcast_to(u[:, None, None, None], (3,) + shape))
        got = force(F, types, dict(lattice=Q))
        assert np.allclose(got, 2 * rho * u, rtol=0, atol=1e-15)
        # nothing inside an empty box
        assert np.all(force(F, types, dict(lattice=Q), box=(0, 1, 0, 1, 0, 1)) == 0)


def _d19_cfg(s, coll=MRT_GS):
    return dict(lattice=19, nx=4, ny=3, nz=3, collision=coll, storage=F64, bc_mode=BC_REFERENCE, s=list(s))


def test_d3q19_basis_known_answers():
    """d'Humieres et al. (2002): rows mutually orthogonal with squared norms 19, 2394, 252, 10, 40
    (j, q), 36, 72, 12, 24, 4, 8; e = -30/-11/8 and eps = 12/-4/1 on rest/face/edge directions."""
    M, Mi = orc.mrt_matrices(3)
    G = M @ M.T
    assert np.allclose(G, np.diag(np.diag(G)), atol=1e-12)
    assert np.allclose(np.diag(G), [19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8])
    assert np.allclose(Mi, M.T / np.diag(G), atol=1e-15)
    assert list(M[1, [0, 1, 7]]) == [-30, -11, 8] and list(M[2, [0, 1, 7]]) == [12, -4, 1]
    assert np.array_equal(M, npr.d3q19_moment_matrix())
    gen = _read_int_table("M19", 19)
    assert np.array_equal(gen, M)


def test_d3q19_mrt_zero_rates_is_pure_streaming():
    cfg = _d19_cfg([0.0] * 27)
    types = np.zeros(36, np.uint8)
    rng = np.random.default_rng(15)
    f = 1 / 19 + 0.01 * rng.random(36 * 19)
    assert rel_l2(orc.run(cfg, types, f, 2), orc.run(dict(cfg, collision=BGK, omega=0.0), types, f, 2)) < 1e-14


def test_d3q19_mrt_uniform_rate_is_bgk():
    om = 1.37
    cfg = _d19_cfg([om] * 27)
    types = np.zeros(36, np.uint8)
    rng = np.random.default_rng(16)
    f = orc.init_fields(cfg, types, 1 + 0.01 * rng.standard_normal(36), *(0.03 * rng.standard_normal((3, 36))))
    f = f + 1e-4 * rng.standard_normal(f.size)
    assert rel_l2(orc.run(cfg, types, f, 3), orc.run(dict(cfg, collision=BGK, omega=om), types, f, 3)) < 1e-13


def test_d3q19_mrt_conserves_mass_and_momentum():
    cfg = _d19_cfg(cases.dhumieres_rates(1.6))
    types = np.zeros(36, np.uint8)
    rng = np.random.default_rng(17)
    f = orc.init_fields(cfg, types, 1 + 0.01 * rng.standard_normal(36), *(0.03 * rng.standard_normal((3, 36))))
    f = f + 1e-4 * rng.standard_normal(f.size)
    g = orc.run(cfg, types, f, 5)
    c, _, _ = orc.tables(19)
    F0, F1 = f.reshape(19, -1), g.reshape(19, -1)
    assert abs(F1.sum() - F0.sum()) < 1e-13
    assert np.allclose(c.astype(float) @ F1.sum(1), c.astype(float) @ F0.sum(1), atol=1e-14)


def test_d3q19_mrt_duct_rates_are_stable():
    """The D3Q19 MRT duct case (uniform velocity inlet, zero-gradient outlet, sphere) stays bounded
    with its rates (s_m = 1.2 on the ghost rows) under both inlet closures; d'Humieres et al.'s
    s_m = 1.98 diverges within 100 steps with the Zou-He inlet and stays bounded for 1,000 steps
    with the extrapolated one."""
    from lbm_amd.capi import MRT_GS
    cfg, t = cases.d3q19_sphere_channel(40, 20, 20, (10, 10, 10), 4, re=10.0, storage=F64, collision=MRT_GS)
    t = t.reshape(-1)
    N = t.size
    f0 = orc.init_fields(cfg, t, np.ones(N), np.full(N, 0.05), np.zeros(N), np.zeros(N))
    fluid = (t & 1) == 0  # the fields are 0 at solid nodes

    def bounded(c, steps):
        rho = orc.fields(c, t, orc.run(c, t, f0, steps))[0][fluid]
        return np.isfinite(rho).all() and np.abs(rho - 1).max() < 0.3

    for closure in (0, 1):
        assert bounded(dict(cfg, inlet_closure=closure), 300)
    dh = cases.dhumieres_rates(cfg["omega"])
    assert not bounded(dict(cfg, s=dh, inlet_closure=0), 100)
    assert bounded(dict(cfg, s=dh, inlet_closure=1), 1000)


def test_suga_channel_case_has_the_programs_setup():
    """cases.d3q27_mrt_channel_suga carries the constants of cpp/d3q19_mrt_channel.cpp: a 320 x 80
    x 80 duct (:24-26), a sphere of radius ny / 9 = 8 at (nx / 4, ny / 2 + ny / 30, nz / 2 + nz / 30)
    = (80, 42, 42) (:31-34), walls on the four sides (:164-167), and its rate groups (:71-110)
    laid on the 27 rows in the order of the raw matrix (rust/examples/d3q27_mrt_lid.rs:167-193)."""
    from lbm_amd.capi import D3Q27, MRT_GS
    cfg, t = cases.d3q27_mrt_channel_suga()
    assert (cfg["lattice"], cfg["nx"], cfg["ny"], cfg["nz"], cfg["collision"]) == (D3Q27, 320, 80, 80, MRT_GS)
    t = t.reshape(80, 80, 320)
    solid = (t & 1) != 0
    assert solid[42, 42, 80] and solid[42 + 8, 42, 80] and not solid[42 + 9, 42, 80]
    assert solid[42, 42 - 8, 80] and not solid[42, 42 - 9, 80] and not solid[42, 42, 80 + 9]
    assert solid[0].all() and solid[-1].all() and solid[:, 0].all() and solid[:, -1].all()
    assert (t[1:-1, 1:-1, 0] & 4).all() and (t[1:-1, 1:-1, -1] & 8).all()
    s = np.array(cfg["s"])
    expect = {1.0: [0, 1, 2, 3], 1.2: [4, 5, 6, 8, 9] + list(range(10, 16)) + list(range(17, 27)),
              0.6: [7, 16]}
    for v, rows in expect.items():
        assert np.all(s[rows] == v), v
    assert sorted(sum(expect.values(), [])) == list(range(27))
    assert abs((1 / cfg["omega"] - 0.5) / 3 - 1 / 9) < 1e-15  # nu from s_nu = 1.2 (:114-115)
    s2 = cases.suga_rates(omega=1.7)
    assert [s2[k] for k in (4, 5, 6, 8, 9)] == [1.7] * 5 and s2[7] == 0.6


def test_suga_channel_reduced_gs_is_steady_raw_diverges():
    """The Suga-rate duct at 40 x 18 x 18 from a uniform start at u = 0.1: relaxed in the
    Gram-Schmidt basis it settles to the steady flow of BGK at the same s_nu (velocity within a few
    percent; the rates differ only on modes that steady low-Re flow barely excites); relaxed in the
    raw basis, as the reference program does, the same groups couple non-orthogonal moments and
    the run diverges within 300 steps."""
    from lbm_amd.capi import BGK, MRT_RAW
    out = {}
    for name, kw in (("gs", {}), ("bgk", dict(collision=BGK)), ("raw", dict(collision=MRT_RAW))):
        cfg, t = cases.d3q27_mrt_channel_suga(40, 18, 18, storage=F64, **kw)
        t = t.reshape(-1)
        N = t.size
        f0 = orc.init_fields(cfg, t, np.ones(N), np.full(N, 0.1), np.zeros(N), np.zeros(N))
        out[name] = orc.fields(cfg, t, orc.run(cfg, t, f0, 400 if name != "raw" else 300))
    fluid = (t & 1) == 0
    rho, ux = out["gs"][0][fluid], out["gs"][1][fluid]
    assert np.isfinite(rho).all() and np.abs(rho - 1).max() < 1.2
    ux_bgk = out["bgk"][1][fluid]
    assert np.linalg.norm(ux - ux_bgk) / np.linalg.norm(ux_bgk) < 0.05
    assert not np.isfinite(out["raw"][0][fluid]).all()


@pytest.mark.parametrize("basis", ["gs", "raw"])
def test_suga_rates_conserve_mass_in_a_closed_box(basis):
    """Steps from a perturbed state inside a closed box (the duct with inlet and outlet walled)
    conserve mass to round-off: the conserved rows carry rate 1, and mass moves only by streaming
    and bounce-back."""
    from lbm_amd.capi import MRT_RAW
    kw = {} if basis == "gs" else dict(collision=MRT_RAW)
    cfg, t = cases.d3q27_mrt_channel_suga(40, 18, 18, storage=F64, **kw)
    t = t.reshape(-1)
    N = t.size
    box = np.where((t & 1) == 0, 0, 1).astype(np.uint8)
    box.reshape(18, 18, 40)[:, :, [0, -1]] = 1
    rng = np.random.default_rng(5)
    fb0 = orc.init_fields(cfg, box, 1 + 0.01 * rng.standard_normal(N), 0.01 * rng.standard_normal(N),
                          np.zeros(N), np.zeros(N))
    fb = orc.run(cfg, box, fb0, 20)
    inside = box == 0
    m0 = orc.fields(cfg, box, fb0)[0][inside].sum()
    m1 = orc.fields(cfg, box, fb)[0][inside].sum()
    assert abs(m1 - m0) < 1e-9 * m0


def test_reference_3d_channels_have_the_programs_setup():
    """The 3D channel cases carry the constants of cpp/d3q27_bgk_channel.cpp (200 x 60 x 60, sphere
    r = ny/10+1 = 7 at (nx/4+1, ny/2+ny/30, nz/2+nz/30) = (51, 32, 32), Re 220, :57-79) and of
    cpp/d3q27_mrt_channel.cpp (320 x 80 x 80, sphere r = 9 at (80, 45, 40), pipe
    (y-40)^2+(z-40)^2 >= 1599, :60-72,206-208), the channel rates of
    rust/examples/d3q27_mrt_channel.rs:106-144, and the boundary planes of both (:181-187)."""
    from lbm_amd.capi import BC_REFERENCE, BGK, MRT_RAW
    cfg, t = cases.d3q27_bgk_channel_reference()
    assert (cfg["nx"], cfg["ny"], cfg["nz"], cfg["collision"], cfg["bc_mode"]) == (200, 60, 60, BGK, BC_REFERENCE)
    assert abs(cfg["omega"] - 1 / (3 * 0.1 * 14 / 220 + 0.5)) < 1e-12
    t = t.reshape(60, 60, 200)
    solid = (t & 1) != 0
    assert solid.sum() == sum((x - 51) ** 2 + (y - 32) ** 2 + (z - 32) ** 2 <= 49
                              for x in range(44, 59) for y in range(25, 40) for z in range(25, 40))
    assert solid[32, 32, 51] and solid[32, 32, 58] and not solid[32, 32, 59] and not solid[0].any()
    assert ((t[:, :, 0] & 0x14) == 0x14).all() and ((t[:, :, -1] & 0x10) == 0x10).all()
    assert not (t[:, :, 1:-1] & 0x14).any()
    cfg, t = cases.d3q27_mrt_channel_reference()
    assert (cfg["nx"], cfg["ny"], cfg["nz"], cfg["collision"]) == (320, 80, 80, MRT_RAW)
    t = t.reshape(80, 80, 320)
    solid = (t & 1) != 0
    assert solid[40, 45, 80] and solid[40, 45 + 9, 80] and not solid[40, 45 + 10, 80]
    y, z = np.meshgrid(np.arange(80), np.arange(80))
    assert np.array_equal(solid[:, :, 200], (y - 40) ** 2 + (z - 40) ** 2 >= 1599)
    assert not solid[40, 1, 200] and solid[40, 0, 200] and not solid[1, 40, 200] and solid[0, 40, 200]
    s = np.array(cfg["s"])
    expect = {1.3: [0, 1, 2, 3], 1.0: [4, 5, 6, 8, 9], 0.55: [7, 16],
              1.2: list(range(10, 16)) + list(range(17, 27))}
    for v, rows in expect.items():
        assert np.all(s[rows] == v), v
    assert sorted(sum(expect.values(), [])) == list(range(27))


def test_reference_3d_channels_reach_steady_flow():
    """Reduced 3D channels in the reference's boundary semantics, started from the programs' initial
    fields, settle: the density excursion and the peak speed at 400 steps equal those at 300 steps
    to 5e-3 (the BGK box, periodic in y and z, still accelerates by about 1e-3 per 100 steps)."""
    from lbm_amd.capi import MRT_GS
    for cfg, t in (cases.d3q27_bgk_channel_reference(40, 20, 20, re=20.0, storage=F64),
                   cases.d3q27_mrt_channel_reference(48, 20, 20, storage=F64),
                   cases.d3q27_mrt_channel_reference(48, 20, 20, storage=F64, collision=MRT_GS)):
        t = t.reshape(-1)
        fluid = (t & 1) == 0
        f = orc.run(cfg, t, orc.init_fields(cfg, t, *cases.reference_channel_start(t, cfg["u_in"])), 300)
        a = orc.fields(cfg, t, f)
        b = orc.fields(cfg, t, orc.run(cfg, t, f, 100))
        for k in (0, 1):
            assert np.isfinite(b[k]).all()
            assert abs(np.abs(b[k][fluid] - (k == 0)).max() - np.abs(a[k][fluid] - (k == 0)).max()) < 5e-3
        assert np.abs(b[0][fluid] - 1).max() < 0.3 and np.abs(b[1][fluid]).max() < 0.2


def _fp32_cases():
    """Every lattice x collision x boundary mode the oracle runs, on small grids."""
    out = []
    for bc in (BC_REFERENCE, BC_PHYSICAL):
        ref = bc == BC_REFERENCE
        out.append((f"d2q9_bgk_{bc}", cases.d2q9_lid_reference(16, storage=F32) if ref else
                    cases.d2q9_karman(40, 20, 10, 10, 3, re=20.0, storage=F32)))
        for q, cols in ((19, (BGK, MRT_GS)), (27, (BGK, MRT_RAW, MRT_GS))):
            for co in cols:
                if q == 19:
                    c = cases.d3q19_lid(8, re=20.0, storage=F32, collision=co) if ref else \
                        cases.d3q19_sphere_channel(20, 10, 10, (6, 5, 5), 2, re=5.0, storage=F32, collision=co)
                elif ref:
                    c = (cases.d3q27_mrt_lid_reference(8, storage=F32) if co == MRT_RAW else
                         cases.d3q27_mrt_lid_gs(8, re=20.0, storage=F32) if co == MRT_GS else
                         cases.d3q27_bgk_lid_reference(8, u_lid=0.05, re=20.0, storage=F32))
                else:
                    c = cases.d3q27_sphere_channel(20, 10, 10, (6, 5, 5), 2, re=5.0, storage=F32, collision=co)
                out.append((f"d3q{q}_co{co}_bc{bc}", c))
    return out


@pytest.mark.parametrize("name,case", _fp32_cases(), ids=[c[0] for c in _fp32_cases()])
def test_fp32_oracle_tracks_fp64_oracle(name, case):
    """The oracle's FP32 path (FMA-contracted equilibrium and relaxation, u = j * (1/rho), FP32
    moment transforms) against its FP64 path, from the same fields: 10 steps stay within 1e-6
    rel-L2, the FP32 rounding level.  A term dropped or changed only in the FP32 path (a sign, an
    FMA operand, a weight) moves it by far more."""
    cfg, types = case
    types = types.reshape(-1)
    rng = np.random.default_rng(3)
    N = types.size
    u = [0.02 * rng.standard_normal(N) * 0.05 for _ in range(3)]
    rho = 1 + 1e-3 * rng.standard_normal(N)
    f32 = orc.run(cfg, types, orc.init_fields(cfg, types, rho, *u), 10)
    cfg64 = dict(cfg, storage=F64)
    f64 = orc.run(cfg64, types, orc.init_fields(cfg64, types, rho, *u), 10)
    e = rel_l2(f32, f64)
    assert 1e-9 < e <= 1e-6, (name, e)  # FP32 rounding present, and no larger


def test_fp32_oracle_pin_sees_small_slips():
    """The 1e-6 band of the pin above is tight enough to see slips far smaller than a wrong term:
    the same FP32 run with one shear rate 0.1 % off, or the lid speed 0.01 % off, lands outside
    it (4e-6 each)."""
    cfg, types = cases.d3q27_mrt_lid_gs(8, re=20.0, storage=F32)
    types = types.reshape(-1)
    rng = np.random.default_rng(3)
    N = types.size
    u = [0.02 * rng.standard_normal(N) * 0.05 for _ in range(3)]
    rho = 1 + 1e-3 * rng.standard_normal(N)
    cfg64 = dict(cfg, storage=F64)
    ref = orc.run(cfg64, types, orc.init_fields(cfg64, types, rho, *u), 10)
    for bad in (dict(cfg, s=[v * 1.001 if k == 4 else v for k, v in enumerate(cfg["s"])]),
                dict(cfg, u_wall=(0.05 * 1.0001, 0.0, 0.0))):
        got = orc.run(bad, types, orc.init_fields(bad, types, rho, *u), 10)
        assert rel_l2(got, ref) > 1e-6


def test_f32_index_compat_rounds_pull_indices_as_float():
    """Above 2^24 populations the reference's FP32 builds index through float arrays
    (cpp/d2q9_bgk_lid.cpp:86): brute force over a 1365 x 1367 periodic D2Q9 box (N*Q = 16,793,595)
    that the gathered index of every population equals fl32(its true index).  With omega = 0 and
    F[p] = p (FP64 values, exact), one step leaves F'[p] = the index it was gathered from."""
    nx, ny = 1365, 1367
    N = nx * ny
    cfg = dict(lattice=9, nx=nx, ny=ny, nz=1, collision=BGK, storage=F64, bc_mode=BC_REFERENCE, omega=0.0)
    types = np.zeros(N, np.uint8)
    f = np.arange(9 * N, dtype=np.float64)
    got = orc.run(cfg, types, f, 1, f32_index_compat=True).astype(np.int64)
    c, _, _ = orc.tables(9)
    x = np.tile(np.arange(nx), ny)
    y = np.repeat(np.arange(ny), nx)
    want = np.empty(9 * N, np.int64)
    for i in range(9):
        src = (x - c[0, i]) % nx + nx * ((y - c[1, i]) % ny) + N * i
        want[N * i:N * (i + 1)] = src.astype(np.float32).astype(np.int64)
    assert np.array_equal(got, want)
    exact = orc.run(cfg, types, f, 1, f32_index_compat=False).astype(np.int64)
    # floats are 2 apart in [2^24, 2^25): exactly the odd indices there move (8,189 of them)
    moved = (exact >= 2 ** 24) & (exact % 2 == 1)
    assert np.array_equal(exact != want, moved) and moved.sum() > 8000


def test_f32_index_compat_refuses_an_index_past_the_array():
    """1366 x 1366: N*Q - 1 = 16,793,603 rounds up to N*Q as a float, an index past the lattice; the
    oracle refuses instead of reading past its array."""
    nx = ny = 1366
    cfg = dict(lattice=9, nx=nx, ny=ny, nz=1, collision=BGK, storage=F64, bc_mode=BC_REFERENCE, omega=0.0)
    assert int(np.float32(9 * nx * ny - 1)) == 9 * nx * ny
    with pytest.raises(ValueError, match="outside"):
        orc.run(cfg, np.zeros(nx * ny, np.uint8), np.zeros(9 * nx * ny), 1, f32_index_compat=True)


# ---- shifted storages: stochastic rounding and arithmetic on deviations ------------------------

def test_stochastic_rounding_matches_mi355x():
    """The oracle's restatement of v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32 equals the conversions'
    outputs on MI355X for 16,384 (x, r) pairs (tests/golden/conversions/sr_conversions.npz: normal and
    subnormal half magnitudes, zeros, the largest halves, overflow).  Taking the random bits from
    the bottom of r instead of the top (a plausible misreading) fails it."""
    d = np.load(os.path.join(GOLDEN, "conversions", "sr_conversions.npz"))
    x, r = d["x"], d["r"]
    got16 = np.array([orc.f2h_sr(a, b) for a, b in zip(x, r)], np.uint16)
    gotbf = np.array([orc.f2bf_sr(a, b) for a, b in zip(x, r)], np.uint16)
    assert np.array_equal(got16, d["f16"]) and np.array_equal(gotbf, d["bf16"])
    low = np.array([orc.f2h_sr(a, (int(b) & 0x1FFF) << 19) for a, b in zip(x[:2000], r[:2000])], np.uint16)
    assert (low != d["f16"][:2000]).mean() > 0.2


def test_stochastic_rounding_is_unbiased_and_exact_on_representable_values():
    """Over all 2^13 (half) or 2^16 (bfloat16) values of the bits the conversion uses, the mean of
    the rounded value is x (the defining property of stochastic rounding): exactly for normal
    halves and bfloat16, within 2^-13 of the ulp for half subnormals (whose bits below that are
    cut first); and a value the format holds comes back unchanged for every r."""
    rng = np.random.default_rng(5)
    for x in np.float32(rng.uniform(-3e3, 3e3, 6)).tolist() + [np.float32(3.3e-6), np.float32(-7.7e-7)]:
        vals = [orc.h2f(orc.f2h_sr(x, k << 19)) for k in range(1 << 13)]
        ulp = 2.0 ** -24 if abs(x) < 2.0 ** -14 else 2.0 ** (np.floor(np.log2(abs(x))) - 10)
        tol = 0.0 if abs(x) >= 2.0 ** -14 else ulp * 2.0 ** -13
        assert abs(np.mean(np.array(vals, np.float64)) - np.float64(np.float32(x))) <= tol
    for x in np.float32(rng.uniform(-1, 1, 4)).tolist():
        vals = [orc.bf2f(orc.f2bf_sr(x, k << 16)) for k in range(1 << 16)]
        assert abs(np.mean(np.array(vals, np.float64)) - np.float64(np.float32(x))) <= 1e-12
    for h in (0x3C00, 0x0001, 0x83FF, 0x7BFF, 0x1234):
        x = orc.h2f(h)
        assert all(orc.f2h_sr(x, int(r)) == h for r in rng.integers(0, 2 ** 32, 50, dtype=np.uint64))


def _lid_2d(side, re, storage, **extra):
    nu = 0.05 * (side - 2) / re
    cfg = dict(lattice=9, nx=side, ny=side, nz=1, collision=BGK, storage=storage, bc_mode=BC_PHYSICAL,
               omega=cases.omega_from_nu(nu), u_wall=(0.05, 0.0, 0.0), **extra)
    return cfg, geo.lid_cavity_2d_physical(side, side).reshape(-1)


def test_shifted_storage_holds_a_long_lid_run_to_fp64():
    """D2Q9 lid cavity 128^2 at tau = 1 (Re 40), 4,000 steps from rest: the F16S state computed on
    deviations and stored with stochastic rounding stays within 1e-3 rel-L2 of the FP64 run in u
    and 1e-5 in rho, and its mean density within 1e-6 (measured 3.1e-4, 3.6e-6, 4.1e-7).  Rounded to
    nearest instead (the oracle's test switch), the frozen rounding errors of the quasi-steady flow
    act as a source: the mean density drifts 40 times as far (1.7e-5) and u more than twice as
    far."""
    cfg64, types = _lid_2d(128, 40.0, F64)
    fluid = (types & 1) == 0
    ref = orc.fields(cfg64, types, orc.run(cfg64, types, orc.init_fields(cfg64, types), 4000))
    out = {}
    for rn in (False, True):
        cfg, _ = _lid_2d(128, 40.0, F16S, oracle_round_nearest=rn)
        got = orc.fields(cfg, types, orc.run(cfg, types, orc.init_fields(cfg, types), 4000))
        e_u = np.linalg.norm(np.stack(got[1:3]) - np.stack(ref[1:3])) / np.linalg.norm(np.stack(ref[1:3]))
        out[rn] = (rel_l2(got[0], ref[0]), e_u, abs((got[0] - ref[0])[fluid].mean()))
    (er, eu, dm), (_, eu_rn, _) = out[False], out[True]
    assert eu <= 1e-3 and er <= 1e-5 and dm <= 1e-6, out
    (_, eu_rn, dm_rn) = out[True]
    assert dm_rn > 5e-6 and eu_rn > 2 * eu, out


@pytest.mark.parametrize("collision", [BGK, MRT_GS], ids=["bgk", "mrt_gs"])
def test_shifted_step_is_the_fp64_step_to_the_storage_rounding(collision):
    """One step on deviations, from a representable F16S state of a D3Q27 sphere duct, rounded to
    nearest: every population is within half a storage ulp of the FP64 step from the same state,
    plus 2e-9 for the FP32 arithmetic on deviations up to 0.03 (measured 1.2e-9).  The unshifted
    FP32 step, whose rounding is relative to f ~ w_i, misses the same bound by up to 1e-7 -- the
    error that the arithmetic on deviations removes; a sign or weight wrong in the shifted
    equilibrium would move whole populations.  Fluid nodes only: the solid ones hold the rest
    equilibrium, w_i as each storage represents it."""
    cfg, types = cases.d3q27_sphere_channel(20, 10, 10, (6, 5, 5), 2, re=5.0, storage=F16S, collision=collision)
    types = types.reshape(-1)
    rng = np.random.default_rng(8)
    N = types.size
    u = [0.03 * rng.standard_normal(N) for _ in range(3)]
    f16 = orc.init_fields(cfg, types, 1 + 1e-3 * rng.standard_normal(N), *u)
    a = orc.run(dict(cfg, oracle_round_nearest=True), types, f16, 1).reshape(27, -1)
    ref = orc.run(dict(cfg, storage=F64), types, f16, 1).reshape(27, -1)
    b32 = orc.run(dict(cfg, storage=F32), types, f16, 1).reshape(27, -1)
    _, w, _ = orc.tables(27)
    half_ulp = np.abs(ref - w[:, None]).clip(2.0 ** -29) * 2.0 ** -11
    fluid = (types & 1) == 0  # solid nodes hold the rest equilibrium, w_i as each storage holds it
    assert np.all((np.abs(a - ref) <= half_ulp + 2e-9)[:, fluid])
    assert np.abs(b32 - ref)[:, fluid].max() > 2e-8
