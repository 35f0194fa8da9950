"""GPU parity of the two-component Shan-Chen D2Q9 kernels (lbm_sc_*, through the C ABI) against
the oracle's restatement on the same populations.  Tolerances: FP64 within 1e-13 absolute on the
populations (the kernels evaluate the oracle's operation sequence: bit-identical in practice);
FP32 against the FP32 oracle within 1e-6 absolute (same sequence, FP32 reciprocal of the common
velocity's denominator correctly rounded).  At sizes the oracle does not reach, the run's
size-independent properties: each fluid's mass and the total momentum."""
import numpy as np
import pytest

import np_reference as R
import oracle as orc
from lbm_amd.capi import F32, F64, ShanChen

pytestmark = pytest.mark.gpu

_, W, _ = R.lattice(9)
C = R.lattice(9)[0][:, :2].astype(float)


def noisy_populations(nx, ny, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    f = np.concatenate([np.outer(W, np.ones(nx * ny)) * (1 + amp * rng.standard_normal((9, nx * ny))),
                        np.outer(W, np.ones(nx * ny)) * (1 + amp * rng.standard_normal((9, nx * ny)))])
    return f.reshape(-1)


def as_storage(f, storage):
    return f.astype(np.float32).astype(np.float64) if storage == F32 else f


@pytest.mark.parametrize("storage", [F64, F32])
@pytest.mark.parametrize("nx,ny,w1,w2,g,steps", [
    (1, 1, 1.0, 1.0, -1.2, 5),
    (3, 5, 0.9, 1.3, -1.2, 20),
    (64, 48, 1.0, 1.0, -1.2, 200),
    (257, 130, 1.7, 0.6, 0.8, 50),
    (100, 100, 1.0, 1.0, -1.2, 300),
])
def test_parity_against_oracle(storage, nx, ny, w1, w2, g, steps):
    f0 = as_storage(noisy_populations(nx, ny, nx + ny), storage)
    with ShanChen(nx, ny, w1, w2, g, storage) as sc:
        sc.set_populations(f0)
        assert np.array_equal(sc.populations(), f0)
        sc.step(steps)
        fg = sc.populations()
    fo = orc.sc_run(dict(nx=nx, ny=ny, omega1=w1, omega2=w2, g=g, storage=0 if storage == F64 else 1), f0, steps)
    err = np.abs(fg - fo).max()
    print(f"shanchen {nx}x{ny} storage={storage} steps={steps}: max |gpu - oracle| = {err:.3e}, "
          f"bitwise={np.array_equal(fg, fo)}")
    assert err <= (1e-13 if storage == F64 else 1e-6)


@pytest.mark.parametrize("storage", [F64, F32])
def test_init_and_fields(storage):
    nx, ny = 37, 21
    rng = np.random.default_rng(4)
    r1 = 1 + 0.001 * (1 - 2 * rng.random(nx * ny))
    r2 = 2 - r1
    with ShanChen(nx, ny, 1.0, 1.0, -1.2, storage) as sc:
        sc.init(r1, r2)
        f = sc.populations().reshape(18, -1)
        if storage == F64:
            expect = np.concatenate([np.outer(W, r1), np.outer(W, r2)])
        else:
            w32 = W.astype(np.float32)
            expect = np.concatenate([w32[:, None] * r1.astype(np.float32)[None],
                                     w32[:, None] * r2.astype(np.float32)[None]]).astype(np.float64)
        assert np.array_equal(f, expect)
        sc.step(7)
        f = sc.populations().reshape(18, -1)
        rho1, rho2, ux, uy = sc.fields()
    np.testing.assert_allclose(rho1, f[:9].sum(0), rtol=1e-6 if storage == F32 else 1e-14)
    np.testing.assert_allclose(rho2, f[9:].sum(0), rtol=1e-6 if storage == F32 else 1e-14)
    j = C.T @ (f[:9] + f[9:])
    np.testing.assert_allclose(ux, j[0] / (rho1 + rho2), rtol=0, atol=1e-7 if storage == F32 else 1e-15)
    np.testing.assert_allclose(uy, j[1] / (rho1 + rho2), rtol=0, atol=1e-7 if storage == F32 else 1e-15)


@pytest.mark.parametrize("storage", [F32, F64])
def test_reference_setup_separates_and_conserves(storage):
    """The reference's example (cpp/d2q9_shanchen.cpp:34-46): 100 x 100, G = -1.2, omega = 1,
    rho = 1 +- drho uniform noise with drho = 0.001; 15000 iterations (its maxiter).  FP32 rounds
    each collision's mass sum: about 1e-8 relative drift per step, 2e-4 over the run."""
    nx = ny = 100
    rng = np.random.default_rng(0)
    d = -0.001 * (1 - 2 * rng.random(nx * ny))
    with ShanChen(nx, ny, 1.0, 1.0, -1.2, storage) as sc:
        sc.init(1 + d, 1 - d)
        m1, m2 = sc.fields()[:2]
        sc.step(15000)
        r1, r2, ux, uy = sc.fields()
        f = sc.populations().reshape(18, -1)
    assert np.isfinite(f).all()
    tol = 5e-4 if storage == F32 else 1e-12
    assert r1.sum() == pytest.approx(m1.sum(), rel=tol) and r2.sum() == pytest.approx(m2.sum(), rel=tol)
    assert np.std(r1 - r2) > 100 * np.std(2 * d)  # demixed domains
    j = (C.T @ (f[:9] + f[9:])).sum(1)
    assert np.abs(j).max() / (nx * ny) < (1e-6 if storage == F32 else 1e-14)  # total momentum stays 0


@pytest.mark.parametrize("storage", [F64, F32])
def test_conservation_at_bench_size(storage):
    nx, ny, steps = 1024, 1024, 200
    f0 = as_storage(noisy_populations(nx, ny, 9, amp=0.01), storage)
    with ShanChen(nx, ny, 0.8, 1.1, -1.2, storage) as sc:
        sc.set_populations(f0)
        sc.step(steps)
        f = sc.populations().reshape(18, -1)
    f0 = f0.reshape(18, -1)
    tol = 1e-12 if storage == F64 else 2e-5
    assert f[:9].sum() == pytest.approx(f0[:9].sum(), rel=tol)
    assert f[9:].sum() == pytest.approx(f0[9:].sum(), rel=tol)
    j0, j = (C.T @ (f0[:9] + f0[9:])).sum(1), (C.T @ (f[:9] + f[9:])).sum(1)
    assert np.abs(j - j0).max() <= (1e-8 if storage == F64 else 0.05)



@pytest.mark.parametrize("shape", range(1, 9))
@pytest.mark.parametrize("storage", [F64, F32])
def test_every_tile_shape_matches_oracle(monkeypatch, shape, storage):
    """LBM_SC_SHAPE selects the step kernel's tile shape at create: every shape, on a lattice that
    leaves partial tiles in both directions, equals the oracle."""
    nx, ny, steps = 131, 45, 40
    monkeypatch.setenv("LBM_SC_SHAPE", str(shape))
    f0 = as_storage(noisy_populations(nx, ny, shape), storage)
    with ShanChen(nx, ny, 0.9, 1.2, -1.2, storage) as sc:
        sc.set_populations(f0)
        sc.step(steps)
        fg = sc.populations()
    fo = orc.sc_run(dict(nx=nx, ny=ny, omega1=0.9, omega2=1.2, g=-1.2, storage=0 if storage == F64 else 1), f0, steps)
    assert np.array_equal(fg, fo)
