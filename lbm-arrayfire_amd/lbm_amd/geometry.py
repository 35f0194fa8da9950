"""Node-type masks (one byte per node, x fastest) for the reference cases and the physical-BC
variants.  Arrays are returned with shape (nz, ny, nx) so that ``.reshape(-1)`` is the library's
node order n = x + nx*(y + ny*z)."""
import numpy as np

from .capi import NODE_INLET, NODE_MOVING, NODE_OUTLET, NODE_RHO_FIX, NODE_SOLID


def _grid(nx, ny, nz=1):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x, y, z


def lid_cavity_2d_reference(nx, ny):
    """cpp/d2q9_bgk_lid.cpp:64,90-91: BOUND = 1 except x in [1, nx-2], y in [1, ny-1]; lid nodes
    y = ny-1, x in [1, nx-2] get the lid velocity (:141-142)."""
    x, y, _ = _grid(nx, ny)
    t = np.zeros((1, ny, nx), np.uint8)
    fluid = (x >= 1) & (x <= nx - 2) & (y >= 1)
    t[~fluid] = NODE_SOLID
    t[fluid & (y == ny - 1)] |= NODE_MOVING
    return t


def lid_cavity_3d_reference(nx, ny, nz):
    """rust/examples/d3q27_mrt_lid.rs:202-212 (and d3q27_bgk_lid.rs:128-136): solid except
    x in [1, nx-2], y in [1, ny-2], z in [1, nz-1]; lid = z = nz-1 plane of that range (:323-325)."""
    x, y, z = _grid(nx, ny, nz)
    t = np.zeros((nz, ny, nx), np.uint8)
    fluid = (x >= 1) & (x <= nx - 2) & (y >= 1) & (y <= ny - 2) & (z >= 1)
    t[~fluid] = NODE_SOLID
    t[fluid & (z == nz - 1)] |= NODE_MOVING
    return t


def channel_2d_reference(nx, ny, ox=None, oy=None, r=None):
    """rust/examples/d2q9_bgk_channel.rs:47-49,100-125,190-201: disc (x-ox)^2+(y-oy)^2 <= r^2 plus
    walls y = 0 and y = ny-1 are solid; the x = 0 column gets the inlet velocity; the x = 0 and
    x = nx-1 columns get rho = 1 after the solid override.  Defaults follow the reference:
    ox = nx/5+1, oy = ny/2+ny/30, r = ny/10+1 (integer division)."""
    ox = nx // 5 + 1 if ox is None else ox
    oy = ny // 2 + ny // 30 if oy is None else oy
    r = ny // 10 + 1 if r is None else r
    x, y, _ = _grid(nx, ny)
    t = np.zeros((1, ny, nx), np.uint8)
    solid = ((x - ox) ** 2 + (y - oy) ** 2 <= r * r) | (y == 0) | (y == ny - 1)
    t[solid] = NODE_SOLID
    t[x == 0] |= NODE_INLET | NODE_RHO_FIX
    t[x == nx - 1] |= NODE_RHO_FIX
    return t


def lid_cavity_2d_physical(nx, ny):
    """Walls x = 0, x = nx-1, y = 0 (half-way bounce-back); the row y = ny-1 is a wall moving with
    the lid velocity (moving-wall bounce-back)."""
    x, y, _ = _grid(nx, ny)
    t = np.zeros((1, ny, nx), np.uint8)
    wall = (x == 0) | (x == nx - 1) | (y == 0) | (y == ny - 1)
    t[wall] = NODE_SOLID
    t[y == ny - 1] |= NODE_MOVING
    return t


def lid_cavity_3d_physical(nx, ny, nz):
    x, y, z = _grid(nx, ny, nz)
    t = np.zeros((nz, ny, nx), np.uint8)
    wall = (x == 0) | (x == nx - 1) | (y == 0) | (y == ny - 1) | (z == 0) | (z == nz - 1)
    t[wall] = NODE_SOLID
    t[z == nz - 1] |= NODE_MOVING
    return t


def channel_2d_physical(nx, ny, cx, cy, r):
    """Channel with a circular obstacle (x-cx)^2+(y-cy)^2 <= r^2: walls y = 0 and y = ny-1, Zou-He
    velocity inlet on x = 0, zero-gradient outlet on x = nx-1."""
    x, y, _ = _grid(nx, ny)
    t = np.zeros((1, ny, nx), np.uint8)
    solid = ((x - cx) ** 2 + (y - cy) ** 2 <= r * r) | (y == 0) | (y == ny - 1)
    t[solid] = NODE_SOLID
    t[(x == 0) & ~solid] |= NODE_INLET
    t[(x == nx - 1) & ~solid] |= NODE_OUTLET
    return t


def channel_3d_reference(nx, ny, nz, ox, oy, oz, r, pipe=False):
    """The 3D channels of cpp/d3q27_bgk_channel.cpp:118-124 and cpp/d3q27_mrt_channel.cpp:204-210
    in the reference's boundary semantics: solid = sphere (x-ox)^2+(y-oy)^2+(z-oz)^2 <= r^2, with
    pipe=True also every node with (y-ny/2)^2+(z-nz/2)^2 >= (ny/2)^2-1 (integer ny/2, nz/2); the
    x = 0 plane gets the inlet velocity and the x = 0 and x = nx-1 planes rho = 1 after the solid
    override (:181-187 / :327-333).  y and z are periodic (the BGK program's pipe mask is
    overwritten by its sphere, :121-124)."""
    x, y, z = _grid(nx, ny, nz)
    t = np.zeros((nz, ny, nx), np.uint8)
    solid = (x - ox) ** 2 + (y - oy) ** 2 + (z - oz) ** 2 <= r * r
    if pipe:
        solid |= (y - ny // 2) ** 2 + (z - nz // 2) ** 2 >= (ny // 2) ** 2 - 1
    t[solid] = NODE_SOLID
    t[x == 0] |= NODE_INLET | NODE_RHO_FIX
    t[x == nx - 1] |= NODE_RHO_FIX
    return t


def channel_3d_physical(nx, ny, nz, cx, cy, cz, r):
    """Duct with a spherical obstacle: walls y = 0, ny-1 and z = 0, nz-1, Zou-He inlet on x = 0,
    zero-gradient outlet on x = nx-1."""
    x, y, z = _grid(nx, ny, nz)
    t = np.zeros((nz, ny, nx), np.uint8)
    solid = ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= r * r) | (y == 0) | (y == ny - 1) | \
        (z == 0) | (z == nz - 1)
    t[solid] = NODE_SOLID
    t[(x == 0) & ~solid] |= NODE_INLET
    t[(x == nx - 1) & ~solid] |= NODE_OUTLET
    return t


def periodic_box(nx, ny, nz=1):
    """All-fluid fully periodic lattice (no boundaries)."""
    return np.zeros((nz, ny, nx), np.uint8)
