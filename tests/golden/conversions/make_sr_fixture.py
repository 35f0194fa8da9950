"""Writes tests/golden/conversions/sr_conversions.npz: outputs of MI355X's stochastic-rounding
conversions (v_cvt_sr_f16_f32, v_cvt_sr_bf16_f32) on (x, r) pairs, the fixture the oracle's
restatement of them (f2h_sr, f2bf_sr) is pinned against.

Made in two steps: tests/golden/conversions/sr_probe.hip, built with
    hipcc --offload-arch=gfx950 -O2 tests/golden/conversions/sr_probe.hip -o sr_probe
and run on the GPU as `sr_probe out.bin`, converts 2^20 pairs (half-normal, half-subnormal, tiny
and 2^-24-multiple magnitudes of both signs, zeros, the largest halves and an overflow) and
writes x, r and both results; this script keeps the 6 special pairs and a seeded random sample:
    python3 tests/golden/conversions/make_sr_fixture.py out.bin
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
n = 1 << 20
raw = open(sys.argv[1], "rb").read()
x = np.frombuffer(raw, np.float32, n, 0)
r = np.frombuffer(raw, np.uint32, n, 4 * n)
h = np.frombuffer(raw, np.uint16, n, 8 * n)
b = np.frombuffer(raw, np.uint16, n, 10 * n)
rng = np.random.default_rng(0x5EED5A)
keep = np.concatenate([np.arange(6), np.sort(rng.choice(np.arange(6, n), 16384 - 6, replace=False))])
np.savez_compressed(os.path.join(HERE, "sr_conversions.npz"), x=x[keep], r=r[keep], f16=h[keep], bf16=b[keep])
print(f"kept {keep.size} of {n}")
