"""Two-component Shan-Chen D2Q9 (cpp/d2q9_shanchen.cpp:39-181, with its work-in-progress defects
corrected as listed in oracle_sc_run's header): the oracle's restatement pinned on the CPU by

- an independent numpy restatement (whole-array np.roll streaming and reductions);
- G = 0 with equal rates: the sum of the two fluids is exactly a single-fluid periodic BGK run of
  the summed populations (the single-fluid oracle), since f_eq is linear in rho and u' is then the
  barycentric velocity;
- conservation: each fluid's mass, and the total momentum, whose change per step is
  u'(w1 rho1 + w2 rho2) - (w1 j1 + w2 j2) (zero only for the momentum-weighted u') plus the pair
  interaction, which cancels between x and x + c_i;
- a point perturbation: from local rest (populations w_i rho after streaming), one extra unit of
  fluid 2 at x0 kicks fluid 1's post-collision momentum at x0 + c_j by exactly -g w_j delta c_j
  rho1 and changes nothing elsewhere;
- the sign of G: g = -1.2 separates a 50/50 mixture from small noise; below the critical coupling
  the noise decays, the faster the larger g (cross-attraction for g > 0).
"""
import numpy as np
import pytest

import np_reference as R
import oracle as O
from lbm_amd.capi import F16S, LbmError, ShanChen


def rest_state(nx, ny, rho1, rho2):
    """Populations that stream into local rest: f_i(x) = w_i rho(x + c_i), so that after the pull
    f_i(x) = w_i rho(x)."""
    c, w, _ = R.lattice(9)
    out = []
    for r in (rho1, rho2):
        r = r.reshape(ny, nx)
        out.append(np.stack([w[i] * np.roll(r, shift=(-c[i][1], -c[i][0]), axis=(0, 1)) for i in range(9)]))
    return np.concatenate([o.reshape(9, -1) for o in out]).reshape(-1)


def noisy(nx, ny, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    _, w, _ = R.lattice(9)
    F = w[:, None, None] * (1 + amp * rng.standard_normal((9, ny, nx)))
    G = w[:, None, None] * (1 + amp * rng.standard_normal((9, ny, nx)))
    return F, G


@pytest.mark.parametrize("w1,w2,g", [(1.0, 1.0, -1.2), (0.9, 1.3, -1.2), (1.7, 0.6, 0.8)])
def test_oracle_matches_numpy(w1, w2, g):
    nx, ny = 13, 7
    F, G = noisy(nx, ny, 1)
    cfg = dict(nx=nx, ny=ny, omega1=w1, omega2=w2, g=g)
    a, b = F, G
    for _ in range(6):
        a, b = R.shanchen_step(a, b, cfg)
    f0 = np.concatenate([F.reshape(9, -1), G.reshape(9, -1)]).reshape(-1)
    o = O.sc_run(cfg, f0, 6).reshape(18, -1)
    np.testing.assert_allclose(o[:9], a.reshape(9, -1), rtol=0, atol=2e-15)
    np.testing.assert_allclose(o[9:], b.reshape(9, -1), rtol=0, atol=2e-15)
    o32 = O.sc_run(dict(cfg, storage=1), f0, 6).reshape(18, -1)
    np.testing.assert_allclose(o32[:9], a.reshape(9, -1), rtol=0, atol=2e-6)


def test_zero_coupling_equal_rates_is_single_fluid_bgk():
    nx, ny, om = 17, 11, 0.8
    F, G = noisy(nx, ny, 2)
    f0 = np.concatenate([F.reshape(9, -1), G.reshape(9, -1)]).reshape(-1)
    o = O.sc_run(dict(nx=nx, ny=ny, omega1=om, omega2=om, g=0.0), f0, 20).reshape(18, -1)
    cfg = dict(lattice=9, nx=nx, ny=ny, collision=0, storage=0, omega=om)
    types = np.zeros(nx * ny, np.uint8)
    single = O.run(cfg, types, (F + G).reshape(-1), 20).reshape(9, -1)
    np.testing.assert_allclose(o[:9] + o[9:], single, rtol=0, atol=1e-14)


def test_mass_and_total_momentum_conserved():
    nx, ny = 24, 16
    F, G = noisy(nx, ny, 3, amp=0.1)
    c, _, _ = R.lattice(9)
    f0 = np.concatenate([F.reshape(9, -1), G.reshape(9, -1)]).reshape(-1)
    cfg = dict(nx=nx, ny=ny, omega1=0.7, omega2=1.6, g=-1.1)
    m0 = (F.sum(), G.sum())
    j0 = np.einsum("ia,iyx->a", c[:, :2].astype(float), F + G)
    o = O.sc_run(cfg, f0, 50).reshape(18, -1)
    assert o[:9].sum() == pytest.approx(m0[0], rel=1e-13)
    assert o[9:].sum() == pytest.approx(m0[1], rel=1e-13)
    j = np.einsum("ia,in->a", c[:, :2].astype(float), o[:9] + o[9:])
    np.testing.assert_allclose(j, j0, rtol=0, atol=1e-12)
    # the momentum weighting of u' is what makes the total conserved: weighting velocities
    # (the reference's j_s / rho_s) breaks it
    a, b = F, G
    for _ in range(50):
        a, b = R.shanchen_step(a, b, cfg)
    jn = np.einsum("ia,iyx->a", c[:, :2].astype(float), a + b)
    np.testing.assert_allclose(jn, j0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("j", range(1, 9))
def test_point_perturbation_kick(j):
    nx, ny, delta, g = 9, 8, 0.25, -1.2
    c, w, _ = R.lattice(9)
    rho1 = np.full(nx * ny, 1.1)
    rho2 = np.full(nx * ny, 0.9)
    x0, y0 = 4, 3
    rho2[x0 + nx * y0] += delta
    cfg = dict(nx=nx, ny=ny, omega1=0.8, omega2=1.2, g=g)
    o = O.sc_run(cfg, rest_state(nx, ny, rho1, rho2), 1).reshape(18, -1)
    j1 = np.einsum("ia,in->an", c[:, :2].astype(float), o[:9])
    expect = np.zeros((2, nx * ny))
    for i in range(1, 9):  # every neighbour of x0 gets its own kick
        n = (x0 + c[i][0]) % nx + nx * ((y0 + c[i][1]) % ny)
        expect[:, n] = -g * w[i] * delta * c[i][:2] * 1.1
    n = (x0 + c[j][0]) % nx + nx * ((y0 + c[j][1]) % ny)
    np.testing.assert_allclose(j1[:, n], expect[:, n], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(j1, expect, rtol=0, atol=1e-14)
    assert np.abs(o[:9].sum(0) - rho1).max() < 1e-14  # fluid 1's densities are untouched


def separation(g, steps=1500):
    nx = ny = 48
    rng = np.random.default_rng(5)
    r1 = 1 + 0.01 * rng.standard_normal(nx * ny)
    r2 = 2 - r1
    cfg = dict(nx=nx, ny=ny, omega1=1.0, omega2=1.0, g=g)
    _, w, _ = R.lattice(9)
    f0 = np.concatenate([np.outer(w, r1), np.outer(w, r2)]).reshape(-1)
    o = O.sc_run(cfg, f0, steps).reshape(18, -1)
    return np.std(r1 - r2), np.std(o[:9].sum(0) - o[9:].sum(0))


def test_coupling_sign_separates_or_mixes():
    s0, s = separation(-1.2)
    assert s > 20 * s0  # demixed domains
    decay = [separation(g, 300)[1] / s0 for g in (-0.4, 0.0, 0.4)]
    assert decay[0] > decay[1] > decay[2] and decay[0] < 0.1


@pytest.mark.parametrize("kw", [dict(nx=0, ny=4), dict(nx=8, ny=8, omega1=2.5), dict(nx=8, ny=8, omega2=0.0),
                                dict(nx=8, ny=8, storage=F16S), dict(nx=8, ny=8, g=float("nan"))])
def test_invalid_descriptors_raise_before_any_device_call(lbm_lib, kw):
    with pytest.raises(LbmError):
        ShanChen(**kw)
