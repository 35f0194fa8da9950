#!/usr/bin/env python3
"""Writes the small FP64 golden vectors in this directory from the CPU oracle.

Each fixture holds the case descriptor, the node-type mask, an initial state and the state after
a few steps.  The oracle itself is pinned by tests/test_oracle.py (closed forms, brute force,
independent numpy restatement); these fixtures then pin the GPU kernels to it bit for bit.
Run from the repository root:  python3 tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "lbm-arrayfire_amd")]

import oracle as orc  # noqa: E402
from lbm_amd import cases, geometry as geo  # noqa: E402
from lbm_amd.capi import F64, MRT_GS  # noqa: E402


def fixtures():
    out = {}
    cfg, t = cases.d2q9_lid_reference(8, re=5.0, storage=F64)
    t = geo.lid_cavity_2d_reference(8, 6)
    out["d2q9_lid_ref_8x6"] = (dict(cfg, ny=6), t, 3)
    cfg, t = cases.d2q9_channel_reference(14, 10, storage=F64)
    out["d2q9_channel_ref_14x10"] = (cfg, t, 3)
    cfg, t = cases.d2q9_karman(24, 12, 7, 6, 2, storage=F64)
    out["d2q9_karman_phys_24x12"] = (cfg, t, 3)
    cfg, t = cases.d2q9_lid_physical(9, re=5.0, storage=F64)
    t = geo.lid_cavity_2d_physical(9, 7)
    out["d2q9_lid_phys_9x7"] = (dict(cfg, ny=7), t, 3)
    cfg, t = cases.d3q27_bgk_lid_reference(5, nu=0.05, storage=F64)
    out["d3q27_bgk_lid_ref_5"] = (cfg, t, 2)
    cfg, t = cases.d3q27_mrt_lid_reference(5, storage=F64)
    out["d3q27_mrt_raw_lid_ref_5"] = (cfg, t, 2)
    cfg, t = cases.d3q27_mrt_lid_gs(5, re=10.0, storage=F64)
    out["d3q27_mrt_gs_lid_ref_5"] = (cfg, t, 2)
    cfg, t = cases.d3q27_sphere_channel(12, 7, 7, (4, 3, 3), 1, storage=F64, collision=MRT_GS)
    out["d3q27_sphere_mrt_phys_12x7x7"] = (dict(cfg, inlet_closure=0), t, 2)  # Zou-He inlet
    return out


def initial_state(cfg, types, seed):
    rng = np.random.default_rng(seed)
    N = types.size
    u = np.zeros((3, N))
    u[0] = 0.03 + 0.01 * rng.standard_normal(N)
    u[1] = 0.01 * rng.standard_normal(N)
    if cfg["lattice"] == 27:
        u[2] = 0.01 * rng.standard_normal(N)
    rho = 1 + 0.005 * rng.standard_normal(N)
    return orc.init_fields(cfg, types, rho, *u)


def main():
    for k, (name, (cfg, types, steps)) in enumerate(sorted(fixtures().items())):
        types = np.ascontiguousarray(types, dtype=np.uint8).reshape(-1)
        f0 = initial_state(cfg, types, 1000 + k)
        f1 = orc.run(cfg, types, f0, steps)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), types=types, f0=f0, f1=f1,
                            steps=np.int64(steps), cfg=json.dumps(cfg))
        print(name, f0.size, steps)


if __name__ == "__main__":
    main()
