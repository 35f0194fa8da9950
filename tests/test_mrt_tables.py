"""The generated D3Q27 moment tables (lbm-arrayfire_amd/csrc/mrt_tables.h) that the FP32 MRT
collision uses: the header is what the generator writes, the closed-form equilibrium moments equal
M f_eq computed numerically, T_GS maps the raw rows onto the Gram-Schmidt rows, and the FP32 form
(monomial moments by three 1D passes, raw rows from monomials, T_GS^-1 diag(s) T_GS, P M^-1,
three inverse passes weighted 2^-|c|) equals M^-1 (m - S (m - meq))."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lbm-arrayfire_amd", "csrc")
sys.path.insert(0, CSRC)

import gen_mrt_tables as gen  # noqa: E402
from np_reference import lattice  # noqa: E402


def tables():
    from fractions import Fraction
    M = gen.raw_rows()
    Mi = gen.inverse(M)
    G = [[Fraction(v) for v in gen.primitive(r)] for r in gen.gram_schmidt(M)]
    T = [[sum(G[k][i] * Mi[i][j] for i in range(27)) for j in range(27)] for k in range(27)]
    f = lambda A: np.array([[float(v) for v in r] for r in A])  # noqa: E731
    return f(M), f(Mi), f(G), f(T), gen.meq_coefficients(M), gen.meq_coefficients(G)


def feq(rho, u, c, w):
    cu = c @ u
    return w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u))


def phi(u):
    ux, uy, uz = u
    return np.array([1, ux, uy, uz, ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz])


def test_header_is_the_generator_output():
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_mrt_tables.py")], check=True,
                         capture_output=True, text=True).stdout
    assert out == open(os.path.join(CSRC, "mrt_tables.h")).read()


def test_closed_form_equilibrium_moments():
    M, _, G, _, eraw, egs = tables()
    c, w, _ = lattice(27)
    rng = np.random.default_rng(3)
    for _ in range(20):
        rho, u = 1 + 0.1 * rng.standard_normal(), 0.1 * rng.standard_normal(3)
        fe = feq(rho, u, c.astype(float), w)
        for A, E in ((M, eraw), (G, egs)):
            closed = rho * np.array([[float(v) for v in row] for row in E]) @ phi(u)
            assert np.allclose(closed, A @ fe, rtol=0, atol=1e-14)


def test_gs_rows_from_raw_rows():
    M, _, G, T, _, _ = tables()
    assert np.array_equal(T @ M, G)
    assert np.all(np.triu(T, 1) == 0)
    Ti = np.linalg.inv(T)
    assert np.allclose(Ti @ T, np.eye(27), atol=1e-13)


def test_fp32_form_equals_reference_relaxation():
    """Both bases, random populations and rates, projected equilibrium, in float64."""
    M, Mi, G, T, eraw, _ = tables()
    c, w, _ = lattice(27)
    rng = np.random.default_rng(4)
    Ti = np.linalg.inv(T)
    E = np.array([[float(v) for v in row] for row in eraw])
    for basis in ("raw", "gs"):
        A = M if basis == "raw" else G
        for _ in range(10):
            f = w * (1 + 0.05 * rng.standard_normal(27))
            s = rng.uniform(0.5, 1.8, 27)
            rho = f.sum()
            u = (c.T.astype(float) @ f) / rho
            meq = A @ feq(rho, u, c.astype(float), w)
            ref = np.linalg.solve(A, A @ f - s * (A @ f - meq))
            d = M @ f - rho * (E @ phi(u))
            if basis == "gs":
                d = Ti @ ((s / 24) * (T @ d))
            else:
                d = (s / 24) * d
            got = f - (24 * Mi) @ d
            assert np.allclose(got, ref, rtol=0, atol=1e-14)


def passes(v, inverse=False):
    """The kernel's three 1D passes over the 3 x 3 x 3 grid (slot (cx+1) + 3(cy+1) + 9(cz+1))."""
    v = np.array(v, dtype=float).reshape(3, 3, 3)  # [z][y][x]
    for axis in (2, 1, 0):
        a = np.moveaxis(v, axis, 0).copy()
        if not inverse:
            fm, f0, fp = a
            out = np.stack([fp + fm + f0, fp - fm, fp + fm])
        else:
            r0, r1, r2 = a
            out = np.stack([r2 - r1, r0 - r2, r2 + r1])
        v = np.moveaxis(out, 0, axis)
    return v.reshape(27)


def test_tensor_product_fp32_form_equals_reference_relaxation():
    """The factorized form the FP32 kernels evaluate, restated step by step in float64."""
    M, Mi, G, T, eraw, _ = tables()
    terms, Ainv = gen.raw_in_monomials(gen.raw_rows())
    Ainv = np.array([[float(v) for v in row] for row in Ainv])
    c, w, _ = lattice(27)
    grid = [(c[i, 0] + 1) + 3 * (c[i, 1] + 1) + 9 * (c[i, 2] + 1) for i in range(27)]
    nz = np.abs(c).sum(1)
    rng = np.random.default_rng(5)
    Ti = np.linalg.inv(T)
    E = np.array([[float(v) for v in row] for row in eraw])
    for basis in ("raw", "gs"):
        A = M if basis == "raw" else G
        for _ in range(10):
            f = w * (1 + 0.05 * rng.standard_normal(27))
            s = rng.uniform(0.5, 1.8, 27)
            rho = f.sum()
            u = (c.T.astype(float) @ f) / rho
            meq = A @ feq(rho, u, c.astype(float), w)
            ref = np.linalg.solve(A, A @ f - s * (A @ f - meq))
            v = np.zeros(27)
            v[grid] = f
            mu = passes(v)
            raw = np.array([sum(k * mu[j] for j, k in t) for t in terms])
            assert np.allclose(raw, M @ f, rtol=0, atol=1e-14)
            d = raw - rho * (E @ phi(u))
            r = Ti @ (s * (T @ d)) if basis == "gs" else s * d
            back = passes(Ainv @ r, inverse=True)
            got = f - back[grid] * 0.5 ** nz
            assert np.allclose(got, ref, rtol=0, atol=1e-14)


def test_d3q19_tables():
    """M19 rows are the d'Humieres polynomials (integral, orthogonal, norms as generated) and MEQ19
    is M19 f_eq / rho in closed form."""
    D = np.array([[float(v) for v in row] for row in gen.d3q19_rows()])
    assert np.allclose(D @ D.T, np.diag(np.diag(D @ D.T)))
    c, w, _ = lattice(19)
    E = np.array([[float(v) for v in row] for row in gen.meq_coefficients_q(
        gen.d3q19_rows(), 19, lambda i: gen.W19[abs(gen.CX[i]) + abs(gen.CY[i]) + abs(gen.CZ[i])])])
    rng = np.random.default_rng(9)
    for _ in range(20):
        rho, u = 1 + 0.1 * rng.standard_normal(), 0.1 * rng.standard_normal(3)
        fe = feq(rho, u, c.astype(float), w)
        assert np.allclose(rho * E @ phi(u), D @ fe, rtol=0, atol=1e-14)


def test_d3q19_pair_form_equals_dense_transforms():
    """The FP32 kernels evaluate m = M19 f and M19^T d on the opposite pairs: even rows over
    f_0 and the 9 sums f_i + f_opp(i), odd rows over the 9 differences, and per pair
    out_i = E + O, out_opp = E - O.  Restated here from the rows' parities (computed from the
    generated matrix, with the opposite map found by search) and checked against the dense
    products on random vectors; a row put in the wrong class fails it."""
    D = np.array([[float(v) for v in row] for row in gen.d3q19_rows()])
    c, _, opp = lattice(19)
    even = [bool(np.all(D[k] == D[k][opp])) for k in range(19)]
    odd = [bool(np.all(D[k] == -D[k][opp])) for k in range(19)]
    assert all(e != o for e, o in zip(even, odd))  # every row has exactly one parity
    reps = [i for i in range(1, 19) if i < opp[i]]
    assert len(reps) == 9
    rng = np.random.default_rng(19)
    for _ in range(10):
        f, d = rng.standard_normal(19), rng.standard_normal(19)
        ps = np.array([f[i] + f[opp[i]] for i in reps])
        pd = np.array([f[i] - f[opp[i]] for i in reps])
        m = np.array([D[k][0] * f[0] + D[k][reps] @ ps if even[k] else D[k][reps] @ pd for k in range(19)])
        assert np.allclose(m, D @ f, rtol=0, atol=1e-12)
        out = np.zeros(19)
        out[0] = sum(D[k][0] * d[k] for k in range(19) if even[k])
        for i in reps:
            E = sum(D[k][i] * d[k] for k in range(19) if even[k])
            O = sum(D[k][i] * d[k] for k in range(19) if odd[k])
            out[i], out[opp[i]] = E + O, E - O
        assert np.allclose(out, D.T @ d, rtol=0, atol=1e-12)
