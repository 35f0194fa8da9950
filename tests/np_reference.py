"""Independent numpy restatement of the reference time step, used to pin the C++ oracle.

Written from the reference programs directly (not from the oracle): whole-array np.roll streaming
(af::shift semantics, out[x] = in[x - c], cpp/d2q9_bgk_lid.cpp:18-28), numpy reductions, boolean
mask overrides, and for MRT a dense moment matrix inverted with numpy.linalg.inv
(rust/examples/d3q27_mrt_lid.rs:167-199,336-387).  Arrays are (Q, nz, ny, nx), FP64.
"""
import numpy as np

# velocity sets as listed in the reference (cpp/d2q9_bgk_lid.cpp:72-73,
# rust/examples/d3q27_mrt_lid.rs:98-100)
D2Q9_C = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0),
                   (-1, -1, 0), (1, -1, 0)])
D2Q9_W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
_ex = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1]
_ey = [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1]
_ez = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1]
D3Q27_C = np.array(list(zip(_ex, _ey, _ez)))
D3Q27_W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)

# D3Q19: rest, the six faces, the twelve edges (xy, xz, yz pairs), listed in D3Q27's order
D3Q19_C = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                    (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1),
                    (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)])
D3Q19_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

SOLID, MOVING, INLET, OUTLET, RHO_FIX = 1, 2, 4, 8, 16


def lattice(Q):
    c, w = {9: (D2Q9_C, D2Q9_W), 19: (D3Q19_C, D3Q19_W), 27: (D3Q27_C, D3Q27_W)}[Q]
    # opposite direction found by search, not taken from a table
    opp = np.array([int(np.where((c == -ci).all(axis=1))[0][0]) for ci in c])
    return c, w, opp


def stream(F, c):
    return np.stack([np.roll(F[i], shift=(c[i][2], c[i][1], c[i][0]), axis=(0, 1, 2))
                     for i in range(len(c))])


def feq(rho, u, c, w):
    cu = np.einsum("ia,a...->i...", c.astype(float), u)
    usq = (u * u).sum(0)
    return w[:, None, None, None] * rho[None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq[None])


def inlet_profile(cfg, shape):
    nz, ny, nx = shape
    if cfg.get("inlet_profile", 0) == 0:
        return np.full(shape, cfg.get("u_in", 0.0))
    _, y, _ = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    H = ny - 2
    eta = y - 0.5
    return 4 * cfg["u_in"] * eta * (H - eta) / H ** 2


def raw_moment_matrix():
    x, y, z = (D3Q27_C[:, k].astype(float) for k in range(3))
    rows = [np.ones(27), x, y, z, x * y, x * z, y * z, x * x + y * y + z * z, x * x - y * y,
            x * x - z * z, x * y * y, x * z * z, y * x * x, z * x * x, y * z * z, z * y * y, x * y * z,
            x * x * y * y, x * x * z * z, y * y * z * z, x * x * y * z, x * y * y * z, x * y * z * z,
            x * y * y * z * z, x * x * y * z * z, x * x * y * y * z, x * x * y * y * z * z]
    return np.array(rows)


def gs_moment_matrix():
    R = raw_moment_matrix()
    G = []
    for r in R:
        v = r.copy()
        for q in G:
            v = v - (r @ q) / (q @ q) * q
        G.append(v)
    return np.array(G)


def d3q19_moment_matrix():
    """d'Humieres et al. (2002) D3Q19 rows: rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pi_xx,
    pww, pi_ww, pxy, pyz, pxz, mx, my, mz."""
    x, y, z = (D3Q19_C[:, k].astype(float) for k in range(3))
    c2 = x * x + y * y + z * z
    rows = [np.ones(19), 19 * c2 - 30, (21 * c2 ** 2 - 53 * c2 + 24) / 2, x, (5 * c2 - 9) * x, y,
            (5 * c2 - 9) * y, z, (5 * c2 - 9) * z, 3 * x * x - c2, (3 * c2 - 5) * (3 * x * x - c2),
            y * y - z * z, (3 * c2 - 5) * (y * y - z * z), x * y, y * z, x * z, (y * y - z * z) * x,
            (z * z - x * x) * y, (x * x - y * y) * z]
    return np.array(rows)


def meq_reference(rho, u):
    ux, uy, uz = u
    cs2 = 1 / 3
    r = rho
    m = np.zeros((27,) + rho.shape)
    m[0] = r
    m[1], m[2], m[3] = r * ux, r * uy, r * uz
    m[4], m[5], m[6] = r * ux * uy, r * ux * uz, r * uy * uz
    m[7] = r + r * (ux * ux + uy * uy + uz * uz)
    m[8], m[9] = r * (ux * ux - uy * uy), r * (ux * ux - uz * uz)
    m[10], m[11], m[12], m[13], m[14], m[15] = (cs2 * r * v for v in (ux, ux, uy, uz, uy, uz))
    m[17] = cs2 * r * (cs2 + ux * ux + uy * uy)
    m[18] = cs2 * r * (cs2 + ux * ux + uz * uz)
    m[19] = cs2 * r * (cs2 + uy * uy + uz * uz)
    m[20], m[21], m[22] = cs2 * r * uy * uz, cs2 * r * ux * uz, cs2 * r * ux * uy
    m[23], m[24], m[25] = cs2 * cs2 * r * ux, cs2 * cs2 * r * uy, cs2 * cs2 * r * uz
    m[26] = ux * ux + uy * uy + uz * uz + cs2 ** 3 * r
    return m


def collide(S, rho, u, cfg, c, w):
    coll = cfg.get("collision", 0)
    if coll == 0:
        om = cfg["omega"]
        return om * feq(rho, u, c, w) + (1 - om) * S
    Q = len(c)
    M = d3q19_moment_matrix() if Q == 19 else (raw_moment_matrix() if coll == 1 else gs_moment_matrix())
    Minv = np.linalg.inv(M)
    m = np.einsum("ki,i...->k...", M, S)
    if Q == 27 and coll == 1 and cfg.get("meq_mode", 0) == 0:
        me = meq_reference(rho, u)
    else:
        me = np.einsum("ki,i...->k...", M, feq(rho, u, c, w))
    s = np.asarray(cfg["s"], dtype=float)[:Q, None, None, None]
    return np.einsum("ik,k...->i...", Minv, m - s * (m - me))


def step_reference(F, types, cfg):
    """One REFERENCE-mode step (cpp/d2q9_bgk_lid.cpp:124-155)."""
    Q = cfg["lattice"]
    c, w, opp = lattice(Q)
    S = stream(F, c)
    rho = S.sum(0)
    u = np.einsum("ia,i...->a...", c.astype(float), S) / rho
    solid = (types & SOLID) != 0
    mov = (types & MOVING) != 0
    for a in range(3):
        u[a][mov] = cfg.get("u_wall", (0, 0, 0))[a]
    inl = (types & INLET) != 0
    u[0][inl] = inlet_profile(cfg, types.shape)[inl]
    for a in range(3):
        u[a][solid] = 0
    rho[solid] = 0
    rho[(types & RHO_FIX) != 0] = cfg.get("rho_bc", 1.0)
    if Q == 9:
        u[2] = 0
    out = collide(S, rho, u, cfg, c, w)
    for i in range(1, Q):
        out[opp[i]][solid] = S[i][solid]
    return out


def step_physical(F, types, cfg):
    """One PHYSICAL-mode step: half-way / moving-wall bounce-back, velocity inlet closed by
    non-equilibrium bounce-back with transverse correction, zero-gradient outlet at rho_bc
    (every population = f_eq(rho_bc, u_nb) + non-equilibrium part of the upstream neighbour)."""
    Q = cfg["lattice"]
    c, w, opp = lattice(Q)
    solid = (types & SOLID) != 0
    outlet = (types & OUTLET) != 0
    inl = (types & INLET) != 0
    moving = (types & MOVING) != 0
    uw = np.asarray(cfg.get("u_wall", (0, 0, 0)), dtype=float)
    rw = cfg.get("rho_bc", 1.0)
    f = np.empty_like(F)
    for i in range(Q):
        shift = (c[i][2], c[i][1], c[i][0])
        pulled = np.roll(F[i], shift=shift, axis=(0, 1, 2))
        src_solid = np.roll(solid, shift=shift, axis=(0, 1, 2))
        src_moving = np.roll(moving, shift=shift, axis=(0, 1, 2))
        bb = F[opp[i]] + np.where(src_moving, 6 * w[i] * rw * float(c[i] @ uw), 0.0)
        f[i] = np.where(src_solid, bb, pulled)
    zero = [i for i in range(Q) if c[i][0] == 0]

    def close(f, sel, sgn, rho, u):
        Ny = sum(c[i][1] * f[i] for i in zero) - 2 / 3 * rho * u[1]
        Nz = sum(c[i][2] * f[i] for i in zero) - 2 / 3 * rho * u[2]
        for i in range(Q):
            if c[i][0] == sgn:
                cu = c[i][0] * u[0] + c[i][1] * u[1] + c[i][2] * u[2]
                new = f[opp[i]] + 6 * w[i] * rho * cu - 18 * w[i] * (c[i][1] * Ny + c[i][2] * Nz)
                f[i] = np.where(sel, new, f[i])

    if inl.any() and cfg.get("inlet_closure", 0) == 1:
        # f_eq(rho_e, u_in) + (g - f_eq(rho_e, u_e)): g the downstream neighbour's previous
        # post-collision populations, rho_e / u_e their moments (f_eq(rho_bc, u_in) without one)
        uin = inlet_profile(cfg, types.shape)
        Fp = np.roll(F, shift=-1, axis=3)
        rp = Fp.sum(0)
        up = np.einsum("ia,i...->a...", c.astype(float), Fp) / np.where(rp == 0, 1.0, rp)
        nb_ok = ~np.roll(solid, shift=-1, axis=2)
        nb_ok[..., -1] = False
        rb = np.where(nb_ok, rp, rw)
        ub = np.stack([uin, 0 * uin, 0 * uin])
        fb = feq(rb, ub, c, w)
        fg = feq(rp, up, c, w)
        for i in range(Q):
            new = np.where(nb_ok, Fp[i] + (fb[i] - fg[i]), fb[i])
            f[i] = np.where(inl, new, f[i])
    elif inl.any():
        uin = inlet_profile(cfg, types.shape)
        neg = [i for i in range(Q) if c[i][0] == -1]
        rho_in = (sum(f[i] for i in zero) + 2 * sum(f[i] for i in neg)) / (1 - uin)
        close(f, inl, 1, rho_in, np.stack([uin, 0 * uin, 0 * uin]))
    if outlet.any():
        # upstream neighbour (x - 1): previous post-collision populations and their moments
        Fm = np.roll(F, shift=1, axis=3)
        rm = Fm.sum(0)
        um = np.einsum("ia,i...->a...", c.astype(float), Fm) / rm
        nb_ok = ~np.roll(solid, shift=1, axis=2)
        nb_ok[..., 0] = False
        um = np.where(nb_ok[None], um, 0.0)
        fe = feq(np.full(types.shape, rw), um, c, w)
        fn = feq(rm, um, c, w)
        for i in range(Q):  # f_eq(rho_bc, u) + (g - f_eq(rho_nb, u)), the definition
            new = np.where(nb_ok, fe[i] + (Fm[i] - fn[i]), fe[i])
            f[i] = np.where(outlet, new, f[i])
    rho = f.sum(0)
    u = np.einsum("ia,i...->a...", c.astype(float), f) / np.where(solid, 1.0, rho)
    out = collide(f, rho, u, cfg, c, w)
    # solid nodes: in 3D lattices those whose aligned block of 64 node indices holds a non-solid
    # node take the rest equilibrium w_i; the others, and those of 2D lattices, keep their values
    new_solid = F
    if Q >= 19:
        flat = solid.reshape(-1)
        pad = (-flat.size) % 64
        blocks = np.concatenate([flat, np.ones(pad, bool)]).reshape(-1, 64)
        fluid_block = np.repeat(~blocks.all(axis=1), 64)[:flat.size].reshape(solid.shape)
        rest = np.broadcast_to(np.asarray(w)[:, None, None, None], F.shape)
        new_solid = np.where(fluid_block[None], rest, F)
    return np.where(solid[None], new_solid, out)


def run(F, types, cfg, nsteps):
    step = step_reference if cfg.get("bc_mode", 0) == 0 else step_physical
    for _ in range(nsteps):
        F = step(F, types, cfg)
    return F


def init_feq(types_shape, cfg, rho=None, u=None):
    Q = cfg["lattice"]
    c, w, _ = lattice(Q)
    rho = np.ones(types_shape) if rho is None else rho
    u = np.zeros((3,) + types_shape) if u is None else u
    return feq(rho, u, c, w)


def force(F, types, cfg, box=None):
    """Momentum-exchange force on the solid nodes inside box = (x0, x1, y0, y1, z0, z1)
    (half-open): sum over links from non-solid nodes x into such solid nodes x + c_i of
    c_i (2 f_i(x) - 6 w_i rho_bc c_i.u_wall), the last term for moving walls only."""
    Q = cfg["lattice"]
    c, w, _ = lattice(Q)
    nz, ny, nx = types.shape
    solid = (types & SOLID) != 0
    moving = (types & MOVING) != 0
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    if box is None:
        box = (0, nx, 0, ny, 0, nz)
    inbox = (x >= box[0]) & (x < box[1]) & (y >= box[2]) & (y < box[3]) & (z >= box[4]) & (z < box[5])
    target = solid & inbox
    uw = np.asarray(cfg.get("u_wall", (0, 0, 0)), dtype=float)
    rw = cfg.get("rho_bc", 1.0)
    out = np.zeros(3)
    for i in range(1, Q):
        shift = (-c[i][2], -c[i][1], -c[i][0])  # value at x + c_i, seen from x
        hit = np.roll(target, shift=shift, axis=(0, 1, 2)) & ~solid
        mov = np.roll(moving, shift=shift, axis=(0, 1, 2))
        m = 2.0 * F[i] - np.where(mov, 6.0 * w[i] * rw * float(c[i] @ uw), 0.0)
        out += c[i].astype(float) * m[hit].sum()
    return out


def shanchen_step(F, G, cfg):
    """One step of two-component Shan-Chen D2Q9 (cpp/d2q9_shanchen.cpp:108-176) on periodic
    (9, ny, nx) arrays F, G of the two fluids, with the reference's work-in-progress defects
    corrected: both fluids are streamed before the collision, fluid s feels the other fluid's
    density, the interaction sums restart every step and the common velocity weights momenta:
        u' = (w1 j1 + w2 j2) / (w1 rho1 + w2 rho2)
        u_1 = u' - (g / w1) sum_{i>=1} w_i rho2(x - c_i) c_i,  u_2 likewise with rho1
        out = f - w (f - f_eq(rho, u_s))."""
    c, w, _ = lattice(9)
    w1, w2, g = cfg["omega1"], cfg["omega2"], cfg["g"]

    def pull(A):
        return np.stack([np.roll(A[i], shift=(c[i][1], c[i][0]), axis=(0, 1)) for i in range(9)])

    F, G = pull(F), pull(G)
    r1, r2 = F.sum(0), G.sum(0)
    j1 = np.einsum("ia,iyx->ayx", c[:, :2].astype(float), F)
    j2 = np.einsum("ia,iyx->ayx", c[:, :2].astype(float), G)
    up = (w1 * j1 + w2 * j2) / (w1 * r1 + w2 * r2)

    def contrib(r):
        out = np.zeros((2,) + r.shape)
        for i in range(1, 9):
            sh = w[i] * np.roll(r, shift=(c[i][1], c[i][0]), axis=(0, 1))  # w_i rho(x - c_i)
            out[0] += sh * c[i][0]
            out[1] += sh * c[i][1]
        return out

    u1 = up - (g / w1) * contrib(r2)
    u2 = up - (g / w2) * contrib(r1)

    def eq(r, u):
        u3 = np.concatenate([u, np.zeros((1,) + r.shape)])[:, None]
        return feq(r[None], u3, c, w)[:, 0]

    return F - w1 * (F - eq(r1, u1)), G - w2 * (G - eq(r2, u2))
