"""Physics checks of the physical-boundary path on the GPU, with each reduced-precision result
stated against the FP64 run.

* Lid-driven cavity, Re = 100: u/U on the vertical centre line against Ghia, Ghia & Shin,
  J. Comput. Phys. 48 (1982) 387, Table I.
* Channel with a cylinder (1024 x 256, r = 20, Re = 150, parabolic inlet u_max = 0.1, started
  from the developed channel profile): the vortex-shedding Strouhal number St = f D / u_max from
  a transverse-velocity sensor in the wake; FP32 and FP16-shifted storage agree with FP64 within
  1e-2, BF16-shifted within 5e-2, and St lies in the range reported for confined cylinders at
  this Reynolds number (loose bound).  The shedding period must be regular (jitter < 5 %).
"""
import numpy as np
import pytest

from lbm_amd import Lattice, cases
from lbm_amd.capi import BF16S, F16S, F32, F64

pytestmark = pytest.mark.gpu

# Ghia et al. (1982), Re = 100: (y, u/U) on the vertical line through the cavity centre
GHIA_RE100 = [(0.9766, 0.84123), (0.9688, 0.78871), (0.9609, 0.73722), (0.9531, 0.68717),
              (0.8516, 0.23151), (0.7344, 0.00332), (0.6172, -0.13641), (0.5000, -0.20581),
              (0.4531, -0.21090), (0.2813, -0.15662), (0.1719, -0.10150), (0.1016, -0.06434),
              (0.0703, -0.04775), (0.0625, -0.04192), (0.0547, -0.03717)]


def test_lid_cavity_matches_ghia_re100():
    n = 130                      # walls at y = 0.5 and n - 1.5: cavity side L = n - 2 = 128
    L, U = n - 2, 0.1
    cfg, types = cases.d2q9_lid_physical(n, u_lid=U, re=100.0, storage=F32)
    cfg["omega"] = cases.omega_from_nu(U * L / 100.0)
    with Lattice(cfg, types) as lat:
        lat.init_fields()
        lat.step(80000)
        rho, ux, uy, _ = lat.fields()
        _, _, bad = lat.monitors()
    assert bad == 0
    ux = ux.reshape(n, n)
    col = 0.5 * (ux[:, n // 2 - 1] + ux[:, n // 2])  # x = 64.5: the centre line
    y = (np.arange(n) - 0.5) / L
    prof = col[1:-1] / U
    worst = 0.0
    for yg, ug in GHIA_RE100:
        got = np.interp(yg, y[1:-1], prof)
        worst = max(worst, abs(got - ug))
    print(f"lid cavity Re 100, L = {L}: max |u/U - Ghia| = {worst:.4f}")
    assert worst < 0.02


def strouhal(storage, steps=60000, sample_every=20, start=30000):
    cfg, types = cases.d2q9_karman(1024, 256, 256, 128, 20, u_max=0.1, re=150.0, storage=storage)
    probe = np.array([600 + 1024 * 128, 700 + 1024 * 140], dtype=np.int64)
    series = []
    with Lattice(cfg, types) as lat:
        # developed channel flow + 1e-4 seeded noise: an impulsive start from rest leaves an
        # acoustic mode ringing between inlet and outlet that modulates the shedding period
        lat.init_fields(*cases.poiseuille_start(1024, 256, 0.1, 0x5EED0002, 1e-4))
        lat.step(start)
        for _ in range((steps - start) // sample_every):
            lat.step(sample_every)
            series.append(lat.fields_at(probe)[2][0])
        _, _, bad = lat.monitors()
    assert bad == 0
    s = np.asarray(series) - np.mean(series)
    # period from upward zero crossings (linear interpolation)
    idx = np.where((s[:-1] < 0) & (s[1:] >= 0))[0]
    t = idx + s[idx] / (s[idx] - s[idx + 1])
    assert len(t) >= 5, "no periodic shedding"
    period = np.mean(np.diff(t)) * sample_every
    return (1.0 / period) * 40 / 0.1, float(np.std(np.diff(t)) / np.mean(np.diff(t)))


def test_cylinder_strouhal_across_precisions():
    st64, jit = strouhal(F64)
    print(f"Strouhal FP64 {st64:.4f} (period jitter {jit:.3f})")
    assert 0.17 <= st64 <= 0.26
    assert jit < 0.05
    for storage, tol in ((F32, 0.01), (F16S, 0.01), (BF16S, 0.05)):
        st, _ = strouhal(storage)
        print(f"Strouhal storage {storage}: {st:.4f}, relative error vs FP64 {abs(st - st64) / st64:.2e}")
        assert abs(st - st64) <= tol * st64


def sphere_drag(storage):
    """Drag coefficient C_d = 2 F_x / (rho u^2 pi r^2) of a sphere (r = 6) in a 128 x 48 x 48 duct,
    uniform inlet u = 0.05, Re = 30, after 4000 steps from a uniform start."""
    r, u = 6, 0.05
    cfg, types = cases.d3q27_sphere_channel(128, 48, 48, (32, 24, 24), r, u_in=u, re=30.0, storage=storage)
    with Lattice(cfg, types) as lat:
        lat.init_equilibrium(1.0, (u, 0.0, 0.0), 0x5EED0004, 1e-4)
        lat.step(4000)
        F = lat.force((32 - r - 1, 32 + r + 2, 24 - r - 1, 24 + r + 2, 24 - r - 1, 24 + r + 2))
        _, _, bad = lat.monitors()
    assert bad == 0
    return 2.0 * F[0] / (u * u * np.pi * r * r)


def test_sphere_drag_across_precisions():
    """The reduced-size sphere case: drag coefficient of each storage format against FP64."""
    cd64 = sphere_drag(F64)
    print(f"sphere C_d FP64 {cd64:.5f}")
    assert 1.0 < cd64 < 20.0  # confined sphere (blockage 25 %) at Re 30: loose physical range
    for storage, tol in ((F32, 1e-4), (F16S, 1e-2), (BF16S, 5e-2)):
        cd = sphere_drag(storage)
        print(f"sphere C_d storage {storage}: {cd:.5f}, relative error vs FP64 {abs(cd - cd64) / cd64:.2e}")
        assert abs(cd - cd64) <= tol * cd64
