"""ctypes loader for the CPU oracle (liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg, never by the product library.  The oracle restates the reference's time step
(see the header of lbm_oracle.cpp for the file:line anchors it follows).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liblbm_oracle.so")


class OracleDesc(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32), ("collision", ctypes.c_int32), ("storage", ctypes.c_int32),
        ("bc_mode", ctypes.c_int32), ("mrt_engine", ctypes.c_int32), ("meq_mode", ctypes.c_int32),
        ("inlet_profile", ctypes.c_int32), ("inlet_closure", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("omega", ctypes.c_double),
        ("s", ctypes.c_double * 27), ("u_wall", ctypes.c_double * 3), ("u_in", ctypes.c_double),
        ("rho_bc", ctypes.c_double),
    ]


class ScDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("omega1", ctypes.c_double), ("omega2", ctypes.c_double),
                ("g", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        dsc = ctypes.POINTER(OracleDesc)
        L.oracle_init_fields.argtypes = [dsc, u8p, dp, dp, dp, dp, dp]
        L.oracle_init_reference.argtypes = [dsc, u8p, ctypes.c_double, dp]
        L.oracle_run.argtypes = [dsc, u8p, dp, ctypes.c_int64, ctypes.c_int]
        L.oracle_run_from.argtypes = [dsc, u8p, dp, ctypes.c_int64, ctypes.c_int, ctypes.c_uint32]
        L.oracle_f2h_sr.argtypes = [ctypes.c_float, ctypes.c_uint32]
        L.oracle_f2h_sr.restype = ctypes.c_uint16
        L.oracle_f2bf_sr.argtypes = [ctypes.c_float, ctypes.c_uint32]
        L.oracle_f2bf_sr.restype = ctypes.c_uint16
        L.oracle_mix32.argtypes = [ctypes.c_uint32]
        L.oracle_mix32.restype = ctypes.c_uint32
        L.oracle_fields.argtypes = [dsc, u8p, dp, dp, dp, dp, dp]
        L.oracle_time_steps.argtypes = [dsc, u8p, ctypes.c_int64, ctypes.c_int64]
        L.oracle_time_steps.restype = ctypes.c_double
        L.oracle_tables.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int32), dp,
                                    ctypes.POINTER(ctypes.c_int32)]
        L.oracle_mrt_matrices.argtypes = [ctypes.c_int, dp, dp]
        L.oracle_f2h.argtypes = [ctypes.c_float]
        L.oracle_f2h.restype = ctypes.c_uint16
        L.oracle_h2f.argtypes = [ctypes.c_uint16]
        L.oracle_h2f.restype = ctypes.c_float
        L.oracle_f2bf.argtypes = [ctypes.c_float]
        L.oracle_f2bf.restype = ctypes.c_uint16
        L.oracle_bf2f.argtypes = [ctypes.c_uint16]
        L.oracle_bf2f.restype = ctypes.c_float
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_max_threads.restype = ctypes.c_int
        L.oracle_sc_run.argtypes = [ctypes.POINTER(ScDesc), dp, ctypes.c_int64]
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def make_desc(cfg):
    """cfg: dict with the lbm_desc field names (see include/lbm_capi.h)."""
    d = OracleDesc()
    d.lattice = cfg["lattice"]
    d.nx, d.ny, d.nz = cfg["nx"], cfg["ny"], cfg.get("nz", 1)
    d.collision = cfg.get("collision", 0)
    d.storage = cfg.get("storage", 0)
    d.bc_mode = cfg.get("bc_mode", 0)
    d.mrt_engine = cfg.get("mrt_engine", 0)
    d.meq_mode = cfg.get("meq_mode", 0)
    d.inlet_profile = cfg.get("inlet_profile", 0)
    d.inlet_closure = cfg.get("inlet_closure", 0)
    # test switch of the oracle alone: shifted storages round to nearest in the steps
    d.reserved = 1 if cfg.get("oracle_round_nearest", False) else 0
    d.omega = cfg.get("omega", 1.0)
    s = cfg.get("s", [1.0] * 27)
    for k in range(27):
        d.s[k] = s[k]
    uw = cfg.get("u_wall", (0.0, 0.0, 0.0))
    for k in range(3):
        d.u_wall[k] = uw[k]
    d.u_in = cfg.get("u_in", 0.0)
    d.rho_bc = cfg.get("rho_bc", 1.0)
    return d


def _q(cfg):
    return cfg["lattice"]


def _n(cfg):
    return cfg["nx"] * cfg["ny"] * cfg.get("nz", 1)


def init_fields(cfg, types, rho=None, ux=None, uy=None, uz=None):
    d = make_desc(cfg)
    f = np.zeros(_n(cfg) * _q(cfg), dtype=np.float64)
    args = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (rho, ux, uy, uz)]
    rc = lib().oracle_init_fields(ctypes.byref(d), _u8(types), *[_dp(a) for a in args], _dp(f))
    assert rc == 0, rc
    return f


def init_reference(cfg, types, rho0=1.0):
    d = make_desc(cfg)
    f = np.zeros(_n(cfg) * _q(cfg), dtype=np.float64)
    rc = lib().oracle_init_reference(ctypes.byref(d), _u8(types), rho0, _dp(f))
    assert rc == 0, rc
    return f


def run(cfg, types, f, nsteps, f32_index_compat=False, step0=0):
    """nsteps steps from populations f.  step0: the lattice's step count at f (the shifted
    storages' stochastic rounding draws on it; 0 after an initial state)."""
    d = make_desc(cfg)
    out = np.array(f, dtype=np.float64, copy=True)
    rc = lib().oracle_run_from(ctypes.byref(d), _u8(types), _dp(out), int(nsteps), int(f32_index_compat),
                               int(step0))
    if rc == 3:
        raise ValueError("f32_index_compat: a pull index rounded through float falls outside the lattice")
    assert rc == 0, rc
    return out


def fields(cfg, types, f):
    d = make_desc(cfg)
    n = _n(cfg)
    rho, ux, uy, uz = (np.zeros(n) for _ in range(4))
    f = np.ascontiguousarray(f, dtype=np.float64)
    rc = lib().oracle_fields(ctypes.byref(d), _u8(types), _dp(f), _dp(rho), _dp(ux), _dp(uy), _dp(uz))
    assert rc == 0, rc
    return rho, ux, uy, uz


def time_steps(cfg, types, nsteps, threads=0, warmup=1):
    """Seconds the oracle takes for nsteps steps after `warmup` untimed ones."""
    if threads:
        lib().oracle_set_threads(int(threads))
    d = make_desc(cfg)
    return lib().oracle_time_steps(ctypes.byref(d), _u8(types), int(nsteps), int(warmup))


def max_threads():
    return lib().oracle_max_threads()


def tables(Q):
    c = np.zeros(3 * Q, dtype=np.int32)
    w = np.zeros(Q)
    opp = np.zeros(Q, dtype=np.int32)
    rc = lib().oracle_tables(Q, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _dp(w),
                             opp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert rc == 0
    return c.reshape(3, Q), w, opp


def mrt_matrices(basis):
    """basis 1 raw, 2 Gram-Schmidt (27 x 27); 3 the D3Q19 orthogonal basis (19 x 19)."""
    n = 19 if basis == 3 else 27
    M = np.zeros(729)
    Mi = np.zeros(729)
    rc = lib().oracle_mrt_matrices(int(basis), _dp(M), _dp(Mi))
    assert rc == 0, rc
    return M[:n * n].reshape(n, n), Mi[:n * n].reshape(n, n)


def f2h(x):
    return lib().oracle_f2h(float(x))


def h2f(h):
    return lib().oracle_h2f(int(h))


def f2h_sr(x, r):
    return lib().oracle_f2h_sr(float(x), int(r))


def f2bf_sr(x, r):
    return lib().oracle_f2bf_sr(float(x), int(r))


def mix32(x):
    return lib().oracle_mix32(int(x))


def f2bf(x):
    return lib().oracle_f2bf(float(x))


def bf2f(b):
    return lib().oracle_bf2f(int(b))


def sc_run(cfg, f, nsteps):
    """Two-component Shan-Chen D2Q9: 18 N populations advanced nsteps (cfg: nx, ny, storage,
    omega1, omega2, g)."""
    d = ScDesc(cfg["nx"], cfg["ny"], cfg.get("storage", 0), 0, cfg["omega1"], cfg["omega2"], cfg["g"])
    out = np.array(f, dtype=np.float64, copy=True)
    rc = lib().oracle_sc_run(ctypes.byref(d), _dp(out), int(nsteps))
    assert rc == 0, rc
    return out
