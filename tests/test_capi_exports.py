"""The C-ABI library loads and exports every symbol include/lbm_capi.h declares; descriptor
validation (pure host code) rejects bad input.  No GPU work is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from lbm_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lbm_capi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol(lbm_lib):
    missing = [s for s in declared_functions() if not hasattr(lbm_lib, s)]
    assert not missing, missing
    assert lbm_lib.lbm_abi_version() == 3


def test_desc_layout_matches_header():
    # 12 int32 fields, then 1 + 27 + 3 + 1 + 1 doubles
    assert ctypes.sizeof(capi.LbmDesc) == 48 + 8 * 33
    assert capi.LbmDesc.omega.offset == 48
    assert capi.LbmDesc.inlet_closure.offset == 40


@pytest.mark.parametrize("bad", [
    dict(lattice=7),
    dict(nz=2),
    dict(collision=1),
    dict(storage=9),
    dict(bc_mode=3),
    dict(nx=0),
    dict(omega=float("nan")),
    dict(inlet_closure=2),
    dict(storage=2, nx=65536, ny=32768),  # F16S: 2^31 nodes
    dict(storage=3, nx=65536, ny=32768),  # BF16S: 2^31 nodes
])
def test_create_rejects_bad_descriptors(lbm_lib, bad):
    cfg = dict(lattice=9, nx=8, ny=8, nz=1, omega=1.0)
    cfg.update(bad)
    d = capi.make_desc(cfg)
    types = np.zeros(8 * 8 * max(cfg["nz"], 1), np.uint8)
    h = ctypes.c_void_p()
    st = lbm_lib.lbm_create(ctypes.byref(d), types.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                            ctypes.byref(h))
    assert st == 1  # LBM_EINVAL, before any device call
    assert lbm_lib.lbm_last_error()


def test_footprint_bytes(lbm_lib):
    cfg = dict(lattice=27, nx=64, ny=32, nz=16, collision=0, storage=capi.F16S)
    N = 64 * 32 * 16
    fp = capi.footprint_bytes(cfg)
    assert fp >= 2 * N * 27 * 2 + N
    assert fp < 2 * (N * 2 + 8448) * 27 + N + 64 * 1024  # population arrays may be padded
    assert capi.footprint_bytes(dict(cfg, lattice=5)) == 0


class _NoCall:
    """Stands in for the library: any C call made by a wrapper under test is a failure."""

    def __getattr__(self, name):
        raise AssertionError(f"{name} called although the arguments were invalid")


def _unbound(cls, **attrs):
    obj = cls.__new__(cls)
    obj.lib = _NoCall()
    obj.h = ctypes.c_void_p()
    for k, v in attrs.items():
        setattr(obj, k, v)
    return obj


def test_wrappers_reject_wrongly_sized_buffers():
    """A buffer of the wrong length raises ValueError before the C ABI (which reads N, N*Q or
    lbm_state_bytes elements from it) is called."""
    lat = _unbound(capi.Lattice, N=48, Q=9, cfg=dict(lattice=9, nx=8, ny=6, storage=capi.F32))
    with pytest.raises(ValueError):
        lat.init_fields(np.ones(47))
    with pytest.raises(ValueError):
        lat.init_fields(None, np.zeros(49))
    with pytest.raises(ValueError):
        lat.set_populations(np.zeros(48 * 9 - 1))
    with pytest.raises(ValueError):
        lat.set_state(np.zeros(48 * 9 * 4 + 1, np.uint8))
    srv = _unbound(capi.Server, _keep={})
    with pytest.raises(ValueError):
        srv.submit(dict(lattice=9, nx=8, ny=6), np.zeros(47, np.uint8), 10)
    sc = _unbound(capi.ShanChen, N=20, nx=5, ny=4)
    with pytest.raises(ValueError):
        sc.init(np.ones(20), np.ones(19))
    with pytest.raises(ValueError):
        sc.set_populations(np.ones(18 * 20 + 1))
