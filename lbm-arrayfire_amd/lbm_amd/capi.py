"""ctypes binding of liblbm_amd.so (the C ABI declared in include/lbm_capi.h).

``Lattice`` mirrors the shape of the reference examples' ``lbm()`` (cpp/d2q9_bgk_lid.cpp:30-207):
descriptor + node-type mask in, ``step`` for the time loop, rho/u fields out.  Every call goes to
the HIP library; there is no CPU fallback — if the library is missing, ``load()`` raises.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "liblbm_amd.so")

# enums of lbm_capi.h
D2Q9, D3Q19, D3Q27 = 9, 19, 27
BGK, MRT_RAW, MRT_GS = 0, 1, 2
F64, F32, F16S, BF16S = 0, 1, 2, 3
BC_REFERENCE, BC_PHYSICAL = 0, 1
MRT_AUTO, MRT_SCALAR, MRT_MFMA_F32, MRT_MFMA_BF16X3, MRT_MFMA_FUSED = 0, 1, 2, 3, 4
MEQ_REFERENCE, MEQ_PROJECTED = 0, 1
INLET_UNIFORM, INLET_PARABOLIC = 0, 1
INLET_ZOU_HE, INLET_EXTRAPOLATED = 0, 1
NODE_FLUID, NODE_SOLID, NODE_MOVING, NODE_INLET, NODE_OUTLET, NODE_RHO_FIX = 0, 1, 2, 4, 8, 16

STATUS = {0: "LBM_OK", 1: "LBM_EINVAL", 2: "LBM_ENOMEM", 3: "LBM_EDEVICE", 4: "LBM_EUNSTABLE",
          5: "LBM_EINTERNAL", 6: "LBM_EAGAIN"}
STORAGE_BYTES = {F64: 8, F32: 4, F16S: 2, BF16S: 2}
NOALLOC_MAX_NODES = 16384  # LBM_NOALLOC_MAX_NODES

# Every symbol include/lbm_capi.h declares (checked by tests/test_capi_exports.py).
EXPORTS = [
    "lbm_abi_version", "lbm_set_device", "lbm_last_error", "lbm_footprint_bytes", "lbm_create",
    "lbm_destroy", "lbm_init_fields", "lbm_init_equilibrium", "lbm_init_reference", "lbm_step",
    "lbm_step_timed", "lbm_sync", "lbm_check_reciprocal", "lbm_steps_done", "lbm_steps_per_launch",
    "lbm_get_fields",
    "lbm_get_fields_at", "lbm_get_fields_at_noalloc",
    "lbm_get_populations",
    "lbm_set_populations", "lbm_state_bytes", "lbm_get_state", "lbm_set_state", "lbm_monitors",
    "lbm_force",
    "lbm_batch_create", "lbm_batch_add", "lbm_batch_remove", "lbm_batch_step", "lbm_batch_step_timed",
    "lbm_batch_sync", "lbm_batch_tune", "lbm_batch_size", "lbm_batch_destroy",
    "lbm_server_create", "lbm_server_submit", "lbm_server_tick", "lbm_server_poll", "lbm_server_take",
    "lbm_server_cancel", "lbm_server_set_check_every", "lbm_server_used_bytes", "lbm_server_clock",
    "lbm_server_get_stats", "lbm_server_destroy",
    "lbm_sc_create", "lbm_sc_init", "lbm_sc_step", "lbm_sc_step_timed", "lbm_sc_get_fields",
    "lbm_sc_get_populations", "lbm_sc_set_populations", "lbm_sc_destroy",
]


class LbmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class LbmDesc(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32), ("collision", ctypes.c_int32), ("storage", ctypes.c_int32),
        ("bc_mode", ctypes.c_int32), ("mrt_engine", ctypes.c_int32), ("meq_mode", ctypes.c_int32),
        ("inlet_profile", ctypes.c_int32), ("inlet_closure", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("omega", ctypes.c_double),
        ("s", ctypes.c_double * 27), ("u_wall", ctypes.c_double * 3), ("u_in", ctypes.c_double),
        ("rho_bc", ctypes.c_double),
    ]


class JobSpec(ctypes.Structure):
    _fields_ = [("desc", LbmDesc), ("node_type", ctypes.POINTER(ctypes.c_uint8)),
                ("rho0", ctypes.c_double), ("u0", ctypes.c_double * 3), ("seed", ctypes.c_uint64),
                ("noise_amp", ctypes.c_double), ("nsteps", ctypes.c_int64),
                ("want_fields", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class JobResult(ctypes.Structure):
    _fields_ = [("job_id", ctypes.c_int64), ("status", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("finish_s", ctypes.c_double),
                ("mass", ctypes.c_double), ("kinetic_energy", ctypes.c_double),
                ("nodes", ctypes.c_int64), ("steps", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("ticks", ctypes.c_int64), ("steps", ctypes.c_int64), ("admitted", ctypes.c_int64),
                ("retired", ctypes.c_int64), ("unstable", ctypes.c_int64),
                ("batch_rebuilds", ctypes.c_int64), ("graph_captures", ctypes.c_int64),
                ("admit_s", ctypes.c_double), ("retire_s", ctypes.c_double), ("launch_s", ctypes.c_double),
                ("wait_s", ctypes.c_double), ("check_s", ctypes.c_double)]


class ScDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("omega1", ctypes.c_double), ("omega2", ctypes.c_double),
                ("g", ctypes.c_double)]


def make_desc(cfg):
    """Builds an LbmDesc from a dict with the lbm_desc field names."""
    d = LbmDesc()
    d.lattice = cfg["lattice"]
    d.nx, d.ny, d.nz = cfg["nx"], cfg["ny"], cfg.get("nz", 1)
    d.collision = cfg.get("collision", BGK)
    d.storage = cfg.get("storage", F32)
    d.bc_mode = cfg.get("bc_mode", BC_REFERENCE)
    d.mrt_engine = cfg.get("mrt_engine", MRT_AUTO)
    d.meq_mode = cfg.get("meq_mode", MEQ_REFERENCE)
    d.inlet_profile = cfg.get("inlet_profile", INLET_UNIFORM)
    d.inlet_closure = cfg.get("inlet_closure", INLET_ZOU_HE)
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


_lib = None


def load(path=None):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("LBM_AMD_LIB", LIB_PATH)
    if not os.path.exists(p):
        raise RuntimeError(f"liblbm_amd.so not found at {p}: build it with __graft_entry__.build()")
    L = ctypes.CDLL(p)
    H = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    L.lbm_abi_version.restype = ctypes.c_int
    L.lbm_last_error.restype = ctypes.c_char_p
    L.lbm_footprint_bytes.argtypes = [ctypes.POINTER(LbmDesc)]
    L.lbm_footprint_bytes.restype = ctypes.c_size_t
    L.lbm_set_device.argtypes = [ctypes.c_int]
    L.lbm_create.argtypes = [ctypes.POINTER(LbmDesc), ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(H)]
    L.lbm_destroy.argtypes = [H]
    L.lbm_init_fields.argtypes = [H, dp, dp, dp, dp]
    L.lbm_init_equilibrium.argtypes = [H, ctypes.c_double, dp, ctypes.c_uint64, ctypes.c_double]
    L.lbm_init_reference.argtypes = [H, ctypes.c_double]
    L.lbm_step.argtypes = [H, ctypes.c_int64]
    L.lbm_step_timed.argtypes = [H, ctypes.c_int64, dp]
    L.lbm_sync.argtypes = [H]
    L.lbm_steps_done.argtypes = [H]
    L.lbm_steps_done.restype = ctypes.c_int64
    L.lbm_check_reciprocal.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]
    L.lbm_steps_per_launch.argtypes = [H]
    L.lbm_steps_per_launch.restype = ctypes.c_int32
    L.lbm_get_fields.argtypes = [H, dp, dp, dp, dp]
    L.lbm_get_fields_at.argtypes = [H, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32, dp, dp, dp, dp]
    L.lbm_get_fields_at_noalloc.argtypes = [H, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32, dp, dp, dp, dp]
    L.lbm_get_populations.argtypes = [H, dp]
    L.lbm_set_populations.argtypes = [H, dp]
    L.lbm_state_bytes.argtypes = [H]
    L.lbm_state_bytes.restype = ctypes.c_size_t
    L.lbm_get_state.argtypes = [H, ctypes.c_void_p]
    L.lbm_set_state.argtypes = [H, ctypes.c_void_p]
    L.lbm_monitors.argtypes = [H, dp, dp, ctypes.POINTER(ctypes.c_int64)]
    L.lbm_force.argtypes = [H, ctypes.POINTER(ctypes.c_int32), dp]
    L.lbm_batch_create.argtypes = [ctypes.POINTER(H)]
    L.lbm_batch_add.argtypes = [H, H]
    L.lbm_batch_remove.argtypes = [H, H]
    L.lbm_batch_step.argtypes = [H, ctypes.c_int64]
    L.lbm_batch_step_timed.argtypes = [H, ctypes.c_int64, dp]
    L.lbm_batch_sync.argtypes = [H]
    L.lbm_batch_tune.argtypes = [H, ctypes.c_int32]
    L.lbm_batch_size.argtypes = [H]
    L.lbm_batch_size.restype = ctypes.c_int32
    L.lbm_batch_destroy.argtypes = [H]
    L.lbm_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int32, ctypes.POINTER(H)]
    L.lbm_server_submit.argtypes = [H, ctypes.POINTER(JobSpec), ctypes.POINTER(ctypes.c_int64)]
    L.lbm_server_tick.argtypes = [H, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.lbm_server_poll.argtypes = [H, ctypes.POINTER(JobResult), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.lbm_server_take.argtypes = [H, ctypes.c_int64, dp, dp, dp, dp]
    L.lbm_server_cancel.argtypes = [H, ctypes.c_int64]
    L.lbm_server_set_check_every.argtypes = [H, ctypes.c_int32]
    L.lbm_server_used_bytes.argtypes = [H]
    L.lbm_server_used_bytes.restype = ctypes.c_size_t
    L.lbm_server_clock.argtypes = [H]
    L.lbm_server_clock.restype = ctypes.c_double
    L.lbm_server_get_stats.argtypes = [H, ctypes.POINTER(ServerStats)]
    L.lbm_server_destroy.argtypes = [H]
    L.lbm_sc_create.argtypes = [ctypes.POINTER(ScDesc), ctypes.POINTER(H)]
    L.lbm_sc_init.argtypes = [H, dp, dp]
    L.lbm_sc_step.argtypes = [H, ctypes.c_int64]
    L.lbm_sc_step_timed.argtypes = [H, ctypes.c_int64, dp]
    L.lbm_sc_get_fields.argtypes = [H, dp, dp, dp, dp]
    L.lbm_sc_get_populations.argtypes = [H, dp]
    L.lbm_sc_set_populations.argtypes = [H, dp]
    L.lbm_sc_destroy.argtypes = [H]
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _sized(a, n, name):
    """a as a contiguous float64 vector of exactly n entries (None stays None); the C ABI reads n
    values from it, so a wrong size raises here instead of reading past the buffer."""
    if a is None:
        return None
    v = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if v.size != n:
        raise ValueError(f"{name} has {v.size} entries, expected {n}")
    return v


def check(status):
    if status != 0:
        raise LbmError(status, load().lbm_last_error().decode())


def check_reciprocal(lo, hi):
    """Float bit patterns x in [lo, hi) whose FP32-kernel reciprocal differs from 1.0f / x."""
    n = ctypes.c_uint64(0)
    check(load().lbm_check_reciprocal(lo, hi, ctypes.byref(n)))
    return n.value


def footprint_bytes(cfg):
    d = make_desc(cfg)
    return load().lbm_footprint_bytes(ctypes.byref(d))


class Lattice:
    """One lattice on the GPU: a handle of the C ABI."""

    def __init__(self, cfg, node_type):
        self.cfg = dict(cfg)
        self.lib = load()
        self.desc = make_desc(cfg)
        self.Q = cfg["lattice"]
        self.shape = (cfg["nx"], cfg["ny"], cfg.get("nz", 1))
        self.N = self.shape[0] * self.shape[1] * self.shape[2]
        self.node_type = np.ascontiguousarray(node_type, dtype=np.uint8).reshape(-1)
        if self.node_type.size != self.N:
            raise ValueError(f"node_type has {self.node_type.size} entries, lattice has {self.N} nodes")
        self.h = ctypes.c_void_p()
        check(self.lib.lbm_create(ctypes.byref(self.desc),
                                  self.node_type.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                  ctypes.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.lbm_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init_fields(self, rho=None, ux=None, uy=None, uz=None):
        arrs = [_sized(a, self.N, name) for a, name in zip((rho, ux, uy, uz), ("rho", "ux", "uy", "uz"))]
        check(self.lib.lbm_init_fields(self.h, *[_dp(a) for a in arrs]))

    def init_equilibrium(self, rho0=1.0, u0=(0.0, 0.0, 0.0), seed=0, noise_amp=0.0):
        u = (ctypes.c_double * 3)(*u0)
        check(self.lib.lbm_init_equilibrium(self.h, rho0, u, seed, noise_amp))

    def init_reference(self, rho0=1.0):
        check(self.lib.lbm_init_reference(self.h, rho0))

    def step(self, n=1):
        check(self.lib.lbm_step(self.h, int(n)))

    def step_timed(self, n):
        ms = ctypes.c_double()
        check(self.lib.lbm_step_timed(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def sync(self):
        check(self.lib.lbm_sync(self.h))

    @property
    def steps_done(self):
        return self.lib.lbm_steps_done(self.h)

    @property
    def steps_per_launch(self):
        return self.lib.lbm_steps_per_launch(self.h)

    def fields(self):
        out = [np.zeros(self.N) for _ in range(4)]
        check(self.lib.lbm_get_fields(self.h, *[_dp(a) for a in out]))
        return out

    def fields_at(self, nodes):
        """rho, ux, uy, uz at the given node indices (x + nx*(y + ny*z))."""
        idx = np.ascontiguousarray(nodes, dtype=np.int64).reshape(-1)
        out = [np.zeros(idx.size) for _ in range(4)]
        check(self.lib.lbm_get_fields_at(self.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                         int(idx.size), *[_dp(a) for a in out]))
        return out

    def fields_at_noalloc(self, nodes):
        """fields_at through lbm_get_fields_at_noalloc (no device allocation; at most
        NOALLOC_MAX_NODES nodes)."""
        idx = np.ascontiguousarray(nodes, dtype=np.int64).reshape(-1)
        out = [np.zeros(idx.size) for _ in range(4)]
        check(self.lib.lbm_get_fields_at_noalloc(self.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                 int(idx.size), *[_dp(a) for a in out]))
        return out

    def populations(self):
        f = np.zeros(self.N * self.Q)
        check(self.lib.lbm_get_populations(self.h, _dp(f)))
        return f

    def set_populations(self, f):
        f = _sized(f, self.N * self.Q, "populations")
        check(self.lib.lbm_set_populations(self.h, _dp(f)))

    def state(self):
        b = np.zeros(self.lib.lbm_state_bytes(self.h), dtype=np.uint8)
        check(self.lib.lbm_get_state(self.h, b.ctypes.data_as(ctypes.c_void_p)))
        return b

    def set_state(self, b):
        b = np.ascontiguousarray(b, dtype=np.uint8).reshape(-1)
        want = self.N * self.Q * STORAGE_BYTES[self.cfg.get("storage", F32)] + 8  # lbm_state_bytes
        if b.size != want:
            raise ValueError(f"state has {b.size} bytes, the lattice holds {want}")
        check(self.lib.lbm_set_state(self.h, b.ctypes.data_as(ctypes.c_void_p)))

    def monitors(self):
        """(mass, kinetic_energy, nonfinite); does not raise on non-finite populations."""
        m, k, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        st = self.lib.lbm_monitors(self.h, ctypes.byref(m), ctypes.byref(k), ctypes.byref(n))
        if st not in (0, 4):
            check(st)
        return m.value, k.value, n.value

    def force(self, box=None):
        """Momentum-exchange force on the solid nodes inside box = (x0, x1, y0, y1, z0, z1)
        (half-open; None = all solid nodes)."""
        out = np.zeros(3)
        b = None if box is None else (ctypes.c_int32 * 6)(*[int(v) for v in box])
        check(self.lib.lbm_force(self.h, b, _dp(out)))
        return out


class Batch:
    """Lattices stepped together, one persistent launch per configuration group per step."""

    def __init__(self, lattices=()):
        self.lib = load()
        self.h = ctypes.c_void_p()
        check(self.lib.lbm_batch_create(ctypes.byref(self.h)))
        self.members = []
        for lat in lattices:
            self.add(lat)

    def add(self, lat):
        check(self.lib.lbm_batch_add(self.h, lat.h))
        self.members.append(lat)

    def remove(self, lat):
        check(self.lib.lbm_batch_remove(self.h, lat.h))
        self.members.remove(lat)

    def step(self, n=1):
        check(self.lib.lbm_batch_step(self.h, int(n)))

    def step_timed(self, n):
        ms = ctypes.c_double()
        check(self.lib.lbm_batch_step_timed(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def sync(self):
        check(self.lib.lbm_batch_sync(self.h))

    def tune(self, trials=5):
        check(self.lib.lbm_batch_tune(self.h, int(trials)))

    def __len__(self):
        return self.lib.lbm_batch_size(self.h)

    def close(self):
        if self.h:
            self.lib.lbm_batch_destroy(self.h)
            self.h = ctypes.c_void_p()
            self.members = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching job server (lbm_server_*)."""

    def __init__(self, budget_bytes, steps_per_tick=100, check_every=16):
        self.lib = load()
        self.h = ctypes.c_void_p()
        check(self.lib.lbm_server_create(int(budget_bytes), int(steps_per_tick), ctypes.byref(self.h)))
        check(self.lib.lbm_server_set_check_every(self.h, int(check_every)))
        self._keep = {}

    def submit(self, cfg, node_type, nsteps, rho0=1.0, u0=(0.0, 0.0, 0.0), seed=0, noise_amp=0.0,
               want_fields=False):
        types = np.ascontiguousarray(node_type, dtype=np.uint8).reshape(-1)
        n = int(np.prod([cfg["nx"], cfg["ny"], cfg.get("nz", 1)]))
        if types.size != n:
            raise ValueError(f"node_type has {types.size} entries, the job's lattice has {n} nodes")
        spec = JobSpec()
        spec.desc = make_desc(cfg)
        spec.node_type = types.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        spec.rho0 = rho0
        for k in range(3):
            spec.u0[k] = u0[k]
        spec.seed = seed
        spec.noise_amp = noise_amp
        spec.nsteps = int(nsteps)
        spec.want_fields = int(bool(want_fields))
        jid = ctypes.c_int64()
        check(self.lib.lbm_server_submit(self.h, ctypes.byref(spec), ctypes.byref(jid)))
        self._keep[jid.value] = int(np.prod([cfg["nx"], cfg["ny"], cfg.get("nz", 1)]))
        return jid.value

    def tick(self):
        a, q = ctypes.c_int32(), ctypes.c_int32()
        check(self.lib.lbm_server_tick(self.h, ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def poll(self, cap=4096):
        buf = (JobResult * cap)()
        n = ctypes.c_int32()
        check(self.lib.lbm_server_poll(self.h, buf, cap, ctypes.byref(n)))
        return [dict(job_id=r.job_id, status=r.status, submit_s=r.submit_s, admit_s=r.admit_s,
                     finish_s=r.finish_s, mass=r.mass, kinetic_energy=r.kinetic_energy, nodes=r.nodes,
                     steps=r.steps) for r in buf[:n.value]]

    def take(self, job_id):
        N = self._keep[job_id]
        out = [np.zeros(N) for _ in range(4)]
        check(self.lib.lbm_server_take(self.h, int(job_id), *[_dp(a) for a in out]))
        return out

    def cancel(self, job_id):
        check(self.lib.lbm_server_cancel(self.h, int(job_id)))

    @property
    def used_bytes(self):
        return self.lib.lbm_server_used_bytes(self.h)

    def stats(self):
        st = ServerStats()
        check(self.lib.lbm_server_get_stats(self.h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def clock(self):
        return self.lib.lbm_server_clock(self.h)

    def close(self):
        if self.h:
            self.lib.lbm_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ShanChen:
    """Two-component Shan-Chen D2Q9 on a periodic lattice (lbm_sc_*), the reference's
    cpp/d2q9_shanchen.cpp:39-181: fluids of relaxation rates omega1, omega2 coupled through
    G sum_i w_i rho_other(x - c_i) c_i in each fluid's equilibrium velocity."""

    def __init__(self, nx, ny, omega1=1.0, omega2=1.0, g=-1.2, storage=F64):
        self.lib = load()
        self.nx, self.ny, self.N = int(nx), int(ny), int(nx) * int(ny)
        d = ScDesc(int(nx), int(ny), int(storage), 0, float(omega1), float(omega2), float(g))
        self.h = ctypes.c_void_p()
        check(self.lib.lbm_sc_create(ctypes.byref(d), ctypes.byref(self.h)))

    def init(self, rho1, rho2):
        """Both fluids at rest: f_i = w_i rho_s."""
        r1, r2 = _sized(rho1, self.N, "rho1"), _sized(rho2, self.N, "rho2")
        check(self.lib.lbm_sc_init(self.h, _dp(r1), _dp(r2)))

    def step(self, n=1):
        check(self.lib.lbm_sc_step(self.h, int(n)))

    def step_timed(self, n):
        ms = ctypes.c_double()
        check(self.lib.lbm_sc_step_timed(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def fields(self):
        """(rho1, rho2, ux, uy) with u the barycentric velocity (j1 + j2) / (rho1 + rho2)."""
        out = [np.zeros(self.N) for _ in range(4)]
        check(self.lib.lbm_sc_get_fields(self.h, *[_dp(a) for a in out]))
        return tuple(out)

    def populations(self):
        """18 N values: fluid 1's populations 0..8 then fluid 2's, population i of node n at
        n + N i."""
        f = np.zeros(18 * self.N)
        check(self.lib.lbm_sc_get_populations(self.h, _dp(f)))
        return f

    def set_populations(self, f):
        f = _sized(f, 18 * self.N, "populations")
        check(self.lib.lbm_sc_set_populations(self.h, _dp(f)))

    def close(self):
        if self.h:
            self.lib.lbm_sc_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
