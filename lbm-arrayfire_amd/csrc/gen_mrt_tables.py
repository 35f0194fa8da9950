#!/usr/bin/env python3
"""Generates mrt_tables.h: exact D3Q27 moment matrices as integer tables.

* MRAW  — the raw monomial moment matrix, rows in the order of
  rust/examples/d3q27_mrt_lid.rs:167-193 (row 0 = 1, 1-3 = c, 4-6 = products, 7 = |c|^2, ...).
* MRAW_INV24 — 24 * MRAW^-1, which is an integer matrix (MRAW^-1 = MRAW_INV24 / 24).
* MGS  — the same rows Gram-Schmidt orthogonalised in row order (unweighted inner product over
  the 27 velocities), each row scaled to primitive integers.  MGS_NORM2[k] = |row k|^2, so
  MGS^-1 = MGS^T diag(1 / MGS_NORM2).
* T_GS, TINV_GS — the Gram-Schmidt rows in terms of the raw ones: MGS = T_GS MRAW (T_GS is
  integral and lower triangular), TINV_GS = T_GS^-1, so MGS^-1 = MRAW^-1 TINV_GS.
* MEQ_RAW, MEQ_GS — the moments of the second-order equilibrium in closed form:
  (M f_eq)_k = rho * sum_j MEQ[k][j] phi_j with phi = (1, ux, uy, uz, ux^2, uy^2, uz^2, ux uy,
  ux uz, uy uz), from f_eq,i = w_i rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u); exact rationals
  printed as doubles.

Exact rational arithmetic (fractions.Fraction); run from this directory:
    python3 gen_mrt_tables.py > mrt_tables.h
"""
from fractions import Fraction
from math import gcd
import sys

# D3Q27 velocities in the reference's order (rust/examples/d3q27_mrt_lid.rs:98-100).
CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1]
CY = [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1]


def raw_rows():
    rows = []
    fns = [
        lambda x, y, z: 1, lambda x, y, z: x, lambda x, y, z: y, lambda x, y, z: z,
        lambda x, y, z: x * y, lambda x, y, z: x * z, lambda x, y, z: y * z,
        lambda x, y, z: x * x + y * y + z * z, lambda x, y, z: x * x - y * y,
        lambda x, y, z: x * x - z * z, lambda x, y, z: x * y * y, lambda x, y, z: x * z * z,
        lambda x, y, z: y * x * x, lambda x, y, z: z * x * x, lambda x, y, z: y * z * z,
        lambda x, y, z: z * y * y, lambda x, y, z: x * y * z, lambda x, y, z: x * x * y * y,
        lambda x, y, z: x * x * z * z, lambda x, y, z: y * y * z * z,
        lambda x, y, z: x * x * y * z, lambda x, y, z: x * y * y * z,
        lambda x, y, z: x * y * z * z, lambda x, y, z: x * y * y * z * z,
        lambda x, y, z: x * x * y * z * z, lambda x, y, z: x * x * y * y * z,
        lambda x, y, z: x * x * y * y * z * z,
    ]
    for f in fns:
        rows.append([Fraction(f(CX[i], CY[i], CZ[i])) for i in range(27)])
    return rows


def inverse(A):
    n = len(A)
    a = [list(A[r]) + [Fraction(int(r == c)) for c in range(n)] for r in range(n)]
    for c in range(n):
        p = next(r for r in range(c, n) if a[r][c] != 0)
        a[c], a[p] = a[p], a[c]
        piv = a[c][c]
        a[c] = [v / piv for v in a[c]]
        for r in range(n):
            if r != c and a[r][c] != 0:
                f = a[r][c]
                a[r] = [vr - f * vc for vr, vc in zip(a[r], a[c])]
    return [row[n:] for row in a]


def gram_schmidt(rows):
    out = []
    for r in rows:
        v = list(r)
        for q in out:
            num = sum(a * b for a, b in zip(r, q))
            den = sum(b * b for b in q)
            c = num / den
            v = [a - c * b for a, b in zip(v, q)]
        out.append(v)
    return out


def primitive(row):
    den = 1
    for v in row:
        den = den * v.denominator // gcd(den, v.denominator)
    ints = [int(v * den) for v in row]
    g = 0
    for v in ints:
        g = gcd(g, abs(v))
    g = g or 1
    ints = [v // g for v in ints]
    # sign convention: first non-zero entry in the same direction as the raw row
    return ints


W = {0: Fraction(8, 27), 1: Fraction(2, 27), 2: Fraction(1, 54), 3: Fraction(1, 216)}


def meq_coefficients(M):
    """MEQ[k][j]: coefficient of phi_j in (M f_eq)_k / rho."""
    out = []
    for row in M:
        co = [Fraction(0)] * 10
        for i in range(27):
            c = (CX[i], CY[i], CZ[i])
            w = W[abs(c[0]) + abs(c[1]) + abs(c[2])] * row[i]
            co[0] += w
            for a in range(3):
                co[1 + a] += 3 * w * c[a]
                co[4 + a] += w * (Fraction(9, 2) * c[a] * c[a] - Fraction(3, 2))
            co[7] += 9 * w * c[0] * c[1]
            co[8] += 9 * w * c[0] * c[2]
            co[9] += 9 * w * c[1] * c[2]
        out.append(co)
    return out


def fmt_rational(name, mat):
    lines = [f"constexpr double {name}[27][{len(mat[0])}] = {{"]
    for row in mat:
        lines.append("    {" + ", ".join(f"{float(v)!r}" if v.denominator == 1 else
                                         f"{v.numerator}.0 / {v.denominator}.0" for v in row) + "},")
    lines.append("};")
    return "\n".join(lines)


W19 = {0: Fraction(1, 3), 1: Fraction(1, 18), 2: Fraction(1, 36)}


def d3q19_rows():
    """The orthogonal D3Q19 moment basis of d'Humieres, Ginzburg, Krafczyk, Lallemand & Luo,
    Phil. Trans. R. Soc. A 360 (2002) 437, over the first 19 directions of the D3Q27 table:
    rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pi_xx, pww, pi_ww, pxy, pyz, pxz, mx, my, mz."""
    rows = [[] for _ in range(19)]
    for i in range(19):
        x, y, z = CX[i], CY[i], CZ[i]
        c2 = x * x + y * y + z * z
        vals = [1, 19 * c2 - 30, Fraction(21 * c2 * c2 - 53 * c2 + 24, 2), x, (5 * c2 - 9) * x, y,
                (5 * c2 - 9) * y, z, (5 * c2 - 9) * z, 3 * x * x - c2, (3 * c2 - 5) * (3 * x * x - c2),
                y * y - z * z, (3 * c2 - 5) * (y * y - z * z), x * y, y * z, x * z, (y * y - z * z) * x,
                (z * z - x * x) * y, (x * x - y * y) * z]
        for k, v in enumerate(vals):
            rows[k].append(Fraction(v))
    return rows


def meq_coefficients_q(M, n, weight):
    """MEQ[k][j]: coefficient of phi_j in (M f_eq)_k / rho over the first n directions."""
    out = []
    for row in M:
        co = [Fraction(0)] * 10
        for i in range(n):
            c = (CX[i], CY[i], CZ[i])
            w = weight(i) * row[i]
            co[0] += w
            for a in range(3):
                co[1 + a] += 3 * w * c[a]
                co[4 + a] += w * (Fraction(9, 2) * c[a] * c[a] - Fraction(3, 2))
            co[7] += 9 * w * c[0] * c[1]
            co[8] += 9 * w * c[0] * c[2]
            co[9] += 9 * w * c[1] * c[2]
        out.append(co)
    return out


def monomial_rows():
    """P[a + 3b + 9c][i] = cx_i^a cy_i^b cz_i^c: the tensor-product moment basis."""
    rows = []
    for c in range(3):
        for b in range(3):
            for a in range(3):
                rows.append([Fraction(CX[i] ** a * CY[i] ** b * CZ[i] ** c) for i in range(27)])
    return rows


def raw_in_monomials(M):
    """A = M P^-1 (raw rows as integer combinations of monomial rows): per row the (monomial,
    coefficient) terms, and P M^-1 (monomial moments from raw moments)."""
    P = monomial_rows()
    A = [[sum(M[k][i] * v for i, v in enumerate(col)) for col in zip(*inverse(P))] for k in range(27)]
    assert all(v.denominator == 1 for row in A for v in row), "raw rows are not integral in monomials"
    terms = [[(j, int(v)) for j, v in enumerate(row) if v != 0] for row in A]
    Ainv = [[sum(P[u][i] * Mi_ij for i, Mi_ij in enumerate(col)) for col in zip(*inverse(M))] for u in range(27)]
    return terms, Ainv


def fmt(name, ctype, mat):
    lines = [f"constexpr {ctype} {name}[27][27] = {{"]
    for row in mat:
        lines.append("    {" + ", ".join(str(v) for v in row) + "},")
    lines.append("};")
    return "\n".join(lines)


def main():
    M = raw_rows()
    Mi = inverse(M)
    Mi24 = [[v * 24 for v in row] for row in Mi]
    assert all(v.denominator == 1 for row in Mi24 for v in row), "24*M^-1 is not integral"
    G = gram_schmidt(M)
    Gi = [primitive(row) for row in G]
    # rows must be mutually orthogonal
    for a in range(27):
        for b in range(a):
            assert sum(x * y for x, y in zip(Gi[a], Gi[b])) == 0
    norm2 = [sum(v * v for v in row) for row in Gi]
    out = sys.stdout
    out.write("// Generated by gen_mrt_tables.py — exact D3Q27 moment matrices.  Do not edit.\n")
    out.write("// MRAW rows follow rust/examples/d3q27_mrt_lid.rs:167-193 of the reference.\n")
    out.write("#pragma once\n\nnamespace lbm {\n\n")
    out.write(fmt("MRAW", "int", [[int(v) for v in row] for row in M]) + "\n\n")
    out.write(fmt("MRAW_INV24", "int", [[int(v) for v in row] for row in Mi24]) + "\n\n")
    out.write(fmt("MGS", "int", Gi) + "\n\n")
    out.write("constexpr long long MGS_NORM2[27] = {" + ", ".join(str(v) for v in norm2) + "};\n\n")
    G = [[Fraction(v) for v in row] for row in Gi]
    T = [[sum(G[k][i] * Mi[i][j] for i in range(27)) for j in range(27)] for k in range(27)]
    assert all(v.denominator == 1 for row in T for v in row), "T_GS is not integral"
    assert all(T[k][j] == 0 for k in range(27) for j in range(k + 1, 27)), "T_GS is not lower triangular"
    Ti = inverse(T)
    out.write(fmt("T_GS", "int", [[int(v) for v in row] for row in T]) + "\n\n")
    out.write(fmt_rational("TINV_GS", Ti) + "\n\n")
    out.write(fmt_rational("MEQ_RAW", meq_coefficients(M)) + "\n\n")
    out.write(fmt_rational("MEQ_GS", meq_coefficients([[Fraction(v) for v in row] for row in Gi])) + "\n\n")
    terms, Ainv = raw_in_monomials(M)
    nt = max(len(t) for t in terms)
    out.write("// Raw row k = sum_t MRT_RAW_COEF[k][t] * monomial MRT_RAW_MONO[k][t] (t < MRT_RAW_NT[k]),\n"
              "// monomial a + 3b + 9c = sum_i cx^a cy^b cz^c f_i.\n")
    out.write("constexpr int MRT_RAW_NT[27] = {" + ", ".join(str(len(t)) for t in terms) + "};\n")
    out.write(f"constexpr int MRT_RAW_MONO[27][{nt}] = {{" + ", ".join(
        "{" + ", ".join(str(j) for j, _ in t + [(0, 0)] * (nt - len(t))) + "}" for t in terms) + "};\n")
    out.write(f"constexpr int MRT_RAW_COEF[27][{nt}] = {{" + ", ".join(
        "{" + ", ".join(str(v) for _, v in t + [(0, 0)] * (nt - len(t))) + "}" for t in terms) + "};\n\n")
    out.write("// Monomial moments from raw moments (P M^-1).\n")
    out.write(fmt_rational("MRT_AINV", Ainv) + "\n\n")
    D = d3q19_rows()
    assert all(v.denominator == 1 for row in D for v in row), "D3Q19 rows are not integral"
    for a in range(19):
        for b in range(a):
            assert sum(x * y for x, y in zip(D[a], D[b])) == 0, "D3Q19 rows are not orthogonal"
    w19 = lambda i: W19[abs(CX[i]) + abs(CY[i]) + abs(CZ[i])]  # noqa: E731
    out.write("// D3Q19 MRT: the orthogonal basis of d'Humieres et al. (2002) over the first 19 directions;\n"
              "// its inverse is M19^T diag(1 / M19_NORM2).  MEQ19: (M19 f_eq)_k / rho in phi.\n")
    out.write("constexpr int M19[19][19] = {\n" + "\n".join(
        "    {" + ", ".join(str(int(v)) for v in row) + "}," for row in D) + "\n};\n\n")
    out.write("constexpr long long M19_NORM2[19] = {" + ", ".join(
        str(int(sum(v * v for v in row))) for row in D) + "};\n\n")
    out.write(fmt_rational("MEQ19", meq_coefficients_q(D, 19, w19)).replace("[27][10]", "[19][10]") + "\n\n")
    out.write("}  // namespace lbm\n")


if __name__ == "__main__":
    main()
