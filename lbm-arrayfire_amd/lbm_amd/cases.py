"""Descriptor builders for the solver cases: the reference examples and the benchmark workloads.

omega follows the reference's unit convention nu = (tau - 1/2)/3, omega = 1/tau
(cpp/d2q9_bgk_lid.cpp:45-49)."""
import numpy as np

from . import geometry as g
from .capi import (BC_PHYSICAL, BC_REFERENCE, BGK, D2Q9, D3Q19, D3Q27, F32, INLET_EXTRAPOLATED, INLET_PARABOLIC,
                   INLET_UNIFORM, MEQ_REFERENCE, MRT_AUTO, MRT_GS, MRT_RAW)


def omega_from_nu(nu):
    return 1.0 / (3.0 * nu + 0.5)


# Rates of the reference MRT lid case, one per moment row (rust/examples/d3q27_mrt_lid.rs:111-150).
def reference_mrt_rates():
    s0, s1, sv, sb, s3, s3b, s4, s4b, s5, s6 = 1.3, 1.3, 1.1, 0.9, 1.2, 0.9, 1.2, 1.2, 1.1, 1.1
    return [s0, s1, s1, s1, sv, sv, sv, sb, sv, sv, s3, s3, s3, s3, s3, s3, s3b,
            s4, s4, s4, s4b, s4b, s4b, s5, s5, s5, s6]


def shear_mrt_rates(omega, conserved=1.0, bulk=1.0, third=1.2, fourth=1.2, fifth=1.1, sixth=1.1):
    """Rates for a given shear viscosity: s_nu = omega on the shear rows (4-6, 8, 9), the other
    groups fixed (recorded with each result)."""
    s = [conserved] * 4 + [omega] * 3 + [bulk] + [omega] * 2 + [third] * 7 + [fourth] * 6 + \
        [fifth] * 3 + [sixth]
    assert len(s) == 27
    return s


def d2q9_lid_reference(n=128, u_lid=0.05, re=100.0, storage=F32):
    """D2Q9 BGK lid cavity, reference semantics; nu = u*nx/Re as cpp/d2q9_bgk_lid.cpp:45."""
    nu = u_lid * n / re
    cfg = dict(lattice=D2Q9, nx=n, ny=n, nz=1, collision=BGK, storage=storage, bc_mode=BC_REFERENCE,
               omega=omega_from_nu(nu), u_wall=(u_lid, 0.0, 0.0))
    return cfg, g.lid_cavity_2d_reference(n, n)


def d2q9_lid_physical(n=128, u_lid=0.05, re=100.0, storage=F32):
    nu = u_lid * n / re
    cfg = dict(lattice=D2Q9, nx=n, ny=n, nz=1, collision=BGK, storage=storage, bc_mode=BC_PHYSICAL,
               omega=omega_from_nu(nu), u_wall=(u_lid, 0.0, 0.0))
    return cfg, g.lid_cavity_2d_physical(n, n)


def d2q9_channel_reference(nx=300, ny=100, u_max=0.1, re=220.0, storage=F32):
    """rust/examples/d2q9_bgk_channel.rs:36-60: nu = u_max * 2r / Re with r = ny/10+1."""
    r = ny // 10 + 1
    nu = u_max * 2.0 * r / re
    cfg = dict(lattice=D2Q9, nx=nx, ny=ny, nz=1, collision=BGK, storage=storage, bc_mode=BC_REFERENCE,
               omega=omega_from_nu(nu), u_in=u_max, inlet_profile=INLET_UNIFORM)
    return cfg, g.channel_2d_reference(nx, ny)


def d2q9_karman(nx=1024, ny=256, cx=256, cy=128, r=20, u_max=0.1, re=150.0, storage=F32):
    """Channel with a cylinder: Zou-He parabolic inlet, zero-gradient outlet at rho_bc = 1,
    no-slip walls.
    Re = u_max * D / nu with D = 2r."""
    nu = u_max * 2 * r / re
    cfg = dict(lattice=D2Q9, nx=nx, ny=ny, nz=1, collision=BGK, storage=storage, bc_mode=BC_PHYSICAL,
               omega=omega_from_nu(nu), u_in=u_max, inlet_profile=INLET_PARABOLIC)
    return cfg, g.channel_2d_physical(nx, ny, cx, cy, r)


def poiseuille_start(nx, ny, u_max=0.1, seed=0x5EED0002, amp=1e-4):
    """rho, ux, uy of the developed channel flow between the half-way walls at y = 0.5 and
    ny - 1.5 (parabola of peak u_max), plus seeded uniform noise of amplitude amp on ux and uy.
    Starting a channel from this state instead of rest avoids the strong pressure pulse of an
    impulsive start, whose longitudinal acoustic mode rings between the velocity inlet and the
    pressure outlet for tens of thousands of steps."""
    H = ny - 2
    y = np.arange(ny) - 0.5
    prof = 4.0 * u_max * y * (H - y) / H ** 2
    prof[0] = prof[-1] = 0.0
    rng = np.random.default_rng(seed)
    ux = np.repeat(prof[:, None], nx, 1) + amp * (2 * rng.random((ny, nx)) - 1)
    uy = amp * (2 * rng.random((ny, nx)) - 1)
    return np.ones(nx * ny), ux.reshape(-1), uy.reshape(-1)


def d3q27_mrt_lid_reference(n=64, u_lid=0.1, storage=F32, engine=MRT_AUTO):
    """rust/examples/d3q27_mrt_lid.rs: raw monomial M, reference rates and meq table."""
    cfg = dict(lattice=D3Q27, nx=n, ny=n, nz=n, collision=MRT_RAW, storage=storage,
               bc_mode=BC_REFERENCE, mrt_engine=engine, meq_mode=MEQ_REFERENCE,
               s=reference_mrt_rates(), u_wall=(u_lid, 0.0, 0.0))
    return cfg, g.lid_cavity_3d_reference(n, n, n)


def d3q27_mrt_lid_gs(n=128, u_lid=0.05, re=400.0, storage=F32, engine=MRT_AUTO, bc_mode=BC_REFERENCE):
    """D3Q27 MRT lid cavity with the Gram-Schmidt basis, s_nu = omega, nu = u*n/Re."""
    nu = u_lid * n / re
    om = omega_from_nu(nu)
    geo = g.lid_cavity_3d_reference(n, n, n) if bc_mode == BC_REFERENCE else g.lid_cavity_3d_physical(n, n, n)
    cfg = dict(lattice=D3Q27, nx=n, ny=n, nz=n, collision=MRT_GS, storage=storage, bc_mode=bc_mode,
               mrt_engine=engine, omega=om, s=shear_mrt_rates(om), u_wall=(u_lid, 0.0, 0.0))
    return cfg, geo


def d3q27_bgk_lid_reference(n=64, u_lid=0.1, re=None, nu=None, storage=F32):
    if nu is None:
        nu = u_lid * n / (re if re else 100.0)
    cfg = dict(lattice=D3Q27, nx=n, ny=n, nz=n, collision=BGK, storage=storage, bc_mode=BC_REFERENCE,
               omega=omega_from_nu(nu), u_wall=(u_lid, 0.0, 0.0))
    return cfg, g.lid_cavity_3d_reference(n, n, n)


def channel_mrt_rates(s0=1.3, s1=1.3, sv=1.0, sb=0.55, s3=1.2, s3b=0.55, s4=1.2, s4b=1.2, s5=1.2, s6=1.2):
    """The 27 rates of the reference's D3Q27 MRT channels (rust/examples/d3q27_mrt_channel.rs:106-144,
    cpp/d3q27_mrt_channel.cpp:116-150) in the raw row order; nu = (1/sv - 1/2)/3 = 1/6."""
    return [s0, s1, s1, s1, sv, sv, sv, sb, sv, sv] + [s3] * 6 + [s3b] + [s4] * 3 + [s4b] * 3 + \
        [s5] * 3 + [s6]


def d3q27_bgk_channel_reference(nx=200, ny=60, nz=60, u_in=0.1, re=220.0, storage=F32):
    """cpp/d3q27_bgk_channel.cpp:57-79,118-124,181-187: a sphere of radius ny/10+1 at
    (nx/4+1, ny/2+ny/30, nz/2+nz/30) in a box periodic in y and z, inlet u = 0.1 on x = 0, rho = 1 on
    x = 0 and x = nx-1, nu = u * 2r / Re (Re 220), in the reference's boundary semantics.  The
    program's D3Q27 weights are wrong (:88-91); the lattice here has the correct ones.  Start with
    lbm_init_fields of reference_channel_start (the program's f_eq of its initial fields, :134-160)."""
    r = ny // 10 + 1
    cfg = dict(lattice=D3Q27, nx=nx, ny=ny, nz=nz, collision=BGK, storage=storage, bc_mode=BC_REFERENCE,
               omega=omega_from_nu(u_in * 2 * r / re), u_in=u_in, inlet_profile=INLET_UNIFORM)
    return cfg, g.channel_3d_reference(nx, ny, nz, nx // 4 + 1, ny // 2 + ny // 30, nz // 2 + nz // 30, r)


def d3q27_mrt_channel_reference(nx=320, ny=80, nz=80, u_in=0.1, storage=F32, collision=MRT_RAW,
                                engine=MRT_AUTO):
    """cpp/d3q27_mrt_channel.cpp:60-72,116-150,204-210,327-333: pipe (y-ny/2)^2+(z-nz/2)^2 >=
    (ny/2)^2-1 and a sphere of radius ny/10+1 at (nx/4, ny/2+ny/15, nz/2), inlet u = 0.1 on x = 0,
    rho = 1 on x = 0 and x = nx-1, the channel rates (s_nu = 1.0, nu = 1/6) on the raw basis with
    the reference equilibrium table, reference boundary semantics.  The program's moment matrix
    never sets row 4 (:176-177, singular); the raw matrix here has row 4 = xy as the Rust
    programs do.  Start with lbm_init_fields of reference_channel_start."""
    r = ny // 10 + 1
    s = channel_mrt_rates()
    cfg = dict(lattice=D3Q27, nx=nx, ny=ny, nz=nz, collision=collision, storage=storage,
               bc_mode=BC_REFERENCE, mrt_engine=engine, meq_mode=MEQ_REFERENCE,
               omega=s[4], s=s, u_in=u_in, inlet_profile=INLET_UNIFORM)
    return cfg, g.channel_3d_reference(nx, ny, nz, nx // 4, ny // 2 + ny // 15, nz // 2, r, pipe=True)


def reference_channel_start(types, u_in=0.1):
    """The initial fields of the reference's channel programs (cpp/d3q27_bgk_channel.cpp:134-142,
    cpp/d2q9_bgk_channel.cpp:107-118): rho = 1 and u = (u_in, 0, 0) everywhere, then rho = u = 0
    at solid nodes.  Returns (rho, ux, uy, uz) for lbm_init_fields."""
    t = np.asarray(types).reshape(-1)
    solid = (t & 1) != 0
    rho = np.where(solid, 0.0, 1.0)
    ux = np.where(solid, 0.0, u_in)
    z = np.zeros(t.size)
    return rho, ux, z, z.copy()


def d3q27_sphere_channel(nx=512, ny=256, nz=256, center=(128, 128, 128), r=24, u_in=0.05, re=200.0,
                         storage=F32, collision=BGK):
    """Duct with a sphere: uniform velocity inlet closed by non-equilibrium extrapolation (the
    Zou-He closure destabilises this duct at tau = 0.536), zero-gradient outlet; Re = u * D / nu,
    D = 2r."""
    nu = u_in * 2 * r / re
    om = omega_from_nu(nu)
    cfg = dict(lattice=D3Q27, nx=nx, ny=ny, nz=nz, collision=collision, storage=storage,
               bc_mode=BC_PHYSICAL, omega=om, s=shear_mrt_rates(om), u_in=u_in,
               inlet_profile=INLET_UNIFORM, inlet_closure=INLET_EXTRAPOLATED)
    return cfg, g.channel_3d_physical(nx, ny, nz, center[0], center[1], center[2], r)


def dhumieres_rates(omega, s_e=1.19, s_eps=1.4, s_q=1.2, s_pi=1.4, s_m=1.98, conserved=1.0):
    """D3Q19 rates in the row order of the orthogonal basis (rho, e, eps, jx, qx, jy, qy, jz, qz,
    3pxx, 3pi_xx, pww, pi_ww, pxy, pyz, pxz, mx, my, mz): s_nu = omega on the five stress rows and
    the values of d'Humieres et al. (2002) on the others (padded to 27 entries)."""
    c = conserved
    s = [c, s_e, s_eps, c, s_q, c, s_q, c, s_q, omega, s_pi, omega, s_pi, omega, omega, omega,
         s_m, s_m, s_m]
    return s + [1.0] * (27 - len(s))


def suga_rates(omega=None, s0=1.0, s1=1.0, sv=1.2, sb=0.6, s3=1.2, s3b=0.6, s4=1.2, s4b=1.2, s5=1.2,
               s6=1.2):
    """The rates of the reference's work-in-progress 27-velocity MRT channel, after Suga et al.
    (2015) (cpp/d3q19_mrt_channel.cpp:71-110), in the D3Q27 row order shared by the raw and the
    Gram-Schmidt bases: rho; j (3); the off-diagonal stresses xy, xz, yz; the trace; the normal
    stress differences x^2-y^2, x^2-z^2; the six third-order rows; xyz; x^2y^2, x^2z^2, y^2z^2;
    x^2yz, xy^2z, xyz^2; the three fifth-order rows; x^2y^2z^2.  omega, when given, replaces s_nu
    on the five stress rows (the program's own s_nu = 1.2 gives nu = 1/9)."""
    if omega is not None:
        sv = omega
    return [s0, s1, s1, s1, sv, sv, sv, sb, sv, sv] + [s3] * 6 + [s3b] + [s4] * 3 + [s4b] * 3 + \
        [s5] * 3 + [s6]


def d3q27_mrt_channel_suga(nx=320, ny=80, nz=80, u_in=0.1, storage=F32, collision=MRT_GS):
    """The setup of the reference's work-in-progress MRT channel (cpp/d3q19_mrt_channel.cpp:22-34,
    71-121, 160-167; named D3Q19 but run on the 27-velocity set): a duct nx x ny x nz with walls on
    y = 0, ny - 1, z = 0, nz - 1, a sphere of radius ny / 9 at (nx / 4, ny / 2 + ny / 30,
    nz / 2 + nz / 30), inlet speed 0.1 and Suga's rates (s_nu = 1.2, nu = 1/9, Re = u * 2r / nu).
    The program's inlet and outlet are unfinished; here they are the uniform velocity inlet
    (closed by non-equilibrium extrapolation) and the zero-gradient outlet of BC_PHYSICAL, and with MRT_GS the equilibrium moments are M f_eq (the
    program's singular moment matrix, which never sets row 4, is the raw matrix with row 4 = xy).
    collision MRT_GS relaxes Suga's groups in the orthogonal basis, MRT_RAW in the raw one."""
    cx, cy, cz, r = nx // 4, ny // 2 + ny // 30, nz // 2 + nz // 30, ny // 9
    s = suga_rates()
    cfg = dict(lattice=D3Q27, nx=nx, ny=ny, nz=nz, collision=collision, storage=storage,
               bc_mode=BC_PHYSICAL, omega=s[4], s=s, u_in=u_in, inlet_profile=INLET_UNIFORM,
               inlet_closure=INLET_EXTRAPOLATED)
    return cfg, g.channel_3d_physical(nx, ny, nz, cx, cy, cz, r)


def d3q19_lid(n=64, u_lid=0.05, re=100.0, storage=F32, collision=BGK, bc_mode=BC_REFERENCE):
    """D3Q19 lid cavity (BGK or MRT in the orthogonal basis), nu = u*n/Re."""
    nu = u_lid * n / re
    om = omega_from_nu(nu)
    geo = g.lid_cavity_3d_reference(n, n, n) if bc_mode == BC_REFERENCE else g.lid_cavity_3d_physical(n, n, n)
    cfg = dict(lattice=D3Q19, nx=n, ny=n, nz=n, collision=collision, storage=storage, bc_mode=bc_mode,
               omega=om, s=dhumieres_rates(om), u_wall=(u_lid, 0.0, 0.0))
    return cfg, geo


def d3q19_sphere_channel(nx=512, ny=256, nz=256, center=(128, 128, 128), r=24, u_in=0.05, re=200.0,
                         storage=F32, collision=BGK):
    """D3Q19 duct with a sphere: uniform velocity inlet (non-equilibrium extrapolation),
    zero-gradient outlet; Re = u * D / nu.
    MRT rates: d'Humieres et al.'s except s_m = 1.2 on the ghost rows mx, my, mz.  With the Zou-He
    inlet their 1.98 drives those modes unstable in a duct (the oracle, FP64, 40 x 20 x 20, r 4,
    Re 10: divergence within 100 steps); with the extrapolated inlet used here both stay bounded."""
    nu = u_in * 2 * r / re
    om = omega_from_nu(nu)
    cfg = dict(lattice=D3Q19, nx=nx, ny=ny, nz=nz, collision=collision, storage=storage,
               bc_mode=BC_PHYSICAL, omega=om, s=dhumieres_rates(om, s_m=1.2), u_in=u_in,
               inlet_profile=INLET_UNIFORM, inlet_closure=INLET_EXTRAPOLATED)
    return cfg, g.channel_3d_physical(nx, ny, nz, center[0], center[1], center[2], r)
