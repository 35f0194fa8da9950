"""lbm_get_fields_at_noalloc (no device allocation) against lbm_get_fields_at: the same values bit
for bit, on two calls in a row (the second with fewer nodes, so stale entries of the shared device
array from the first call would show), on lattices of every boundary mode and 2- and 4-byte
storage; and the size limit."""
import numpy as np
import pytest

from lbm_amd import Lattice, LbmError, cases
from lbm_amd.capi import F16S, F32, F64, NOALLOC_MAX_NODES

pytestmark = pytest.mark.gpu

CASES = [
    ("d2q9_karman_phys_f32", lambda: cases.d2q9_karman(256, 64, 64, 32, 6, re=30.0, storage=F32), (0.1, 0, 0)),
    ("d3q27_mrt_gs_lid_ref_f16s", lambda: cases.d3q27_mrt_lid_gs(20, re=100.0, storage=F16S), (0, 0, 0)),
    ("d3q27_sphere_phys_f64", lambda: cases.d3q27_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0, storage=F64),
     (0.05, 0, 0)),
]


@pytest.mark.parametrize("name,make,u0", CASES, ids=[c[0] for c in CASES])
def test_noalloc_fields_equal_fields_at(name, make, u0):
    cfg, types = make()
    rng = np.random.default_rng(7)
    with Lattice(cfg, types) as lat:
        lat.init_equilibrium(1.0, u0, 3, 1e-3)
        lat.step(37)
        for count in (min(NOALLOC_MAX_NODES, lat.N), 97):  # the second call reads fewer nodes
            nodes = rng.integers(0, lat.N, count)
            nodes[:3] = (0, lat.N - 1, lat.N // 2)
            want = lat.fields_at(nodes)
            got = lat.fields_at_noalloc(nodes)
            for a, (g, w) in enumerate(zip(got, want)):
                assert np.array_equal(g, w), (name, count, a)
            lat.step(5)  # and the state moves on between the calls


def test_noalloc_rejects_more_nodes_than_its_array_holds():
    cfg, types = cases.d2q9_lid_reference(160, storage=F32)  # 25,600 nodes
    with Lattice(cfg, types) as lat:
        lat.init_reference(1.0)
        with pytest.raises(LbmError, match="LBM_NOALLOC_MAX_NODES"):
            lat.fields_at_noalloc(np.arange(NOALLOC_MAX_NODES + 1))
        with pytest.raises(LbmError):
            lat.fields_at_noalloc(np.array([lat.N]))
        assert all(a.size == 0 for a in lat.fields_at_noalloc(np.zeros(0, np.int64)))
        r = lat.fields_at_noalloc(np.arange(NOALLOC_MAX_NODES))  # the limit itself is accepted
        assert np.array_equal(r[0], lat.fields()[0][:NOALLOC_MAX_NODES])
