"""Exact division of node indices, as the flat-tile kernels do it.

Lattices whose rows are not a multiple of 64 nodes wide are stepped in tiles of 256 consecutive
node indices; each lane recovers (x, y, z) from its index n by two divisions through one high
multiply each (`lbm::fastdiv`, Granlund & Montgomery 1994, fig. 4.1, constants from
`lbm::make_fastdiv`).  The header is shared by the host and the kernels, so the routine compiled
here is the one the kernels inline.  Every divisor up to 4096 is checked over the first 2^18
indices (against a quotient counted up one remainder at a time) and over hashed and top-of-range
indices (against the hardware divide); two divisors (a width of 193 nodes, and 2^31 - 1 for the
largest shift) over every 32-bit index.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lbm-arrayfire_amd", "csrc")


def test_fastdiv_matches_the_hardware_divide(tmp_path):
    exe = str(tmp_path / "fastdiv_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", exe,
                    os.path.join(ROOT, "tests", "fastdiv_check.cpp")], check=True)
    out = subprocess.run([exe, "193", str(2**31 - 1)], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
