"""Batched launches and the continuous-batching server: every lattice/job ends with exactly the
populations it reaches when stepped alone (same per-node arithmetic), the server keeps resident
jobs within its HBM budget, and a diverging job is retired alone."""
import numpy as np
import pytest

from lbm_amd import Batch, Lattice, LbmError, Server, cases, geometry as geo
from lbm_amd.capi import (BC_PHYSICAL, BF16S, BGK, D3Q27, F16S, F32, F64, MRT_GS, MRT_MFMA_BF16X3, MRT_MFMA_F32,
                          MRT_MFMA_FUSED, MRT_SCALAR)

pytestmark = pytest.mark.gpu


def member_specs():
    """(cfg, types, init) of a mixed set: several kernel configuration groups, uneven sizes."""
    specs = []
    cfg, t = cases.d2q9_lid_reference(64, storage=F32)
    specs.append((cfg, t, ("ref",)))
    cfg, t = cases.d2q9_lid_reference(32, storage=F32)
    specs.append((cfg, t, ("ref",)))
    cfg, t = cases.d2q9_karman(128, 48, 32, 24, 5, re=30.0, storage=F16S)
    specs.append((cfg, t, ("eq", (0.1, 0, 0), 5, 1e-4)))
    cfg, t = cases.d3q27_mrt_lid_gs(24, re=100.0, storage=F16S, engine=MRT_MFMA_F32)  # two-stage MFMA
    specs.append((cfg, t, ("ref",)))
    cfg, t = cases.d3q27_mrt_lid_gs(20, re=100.0, storage=F16S)  # the scalar engine (AUTO)
    specs.append((cfg, t, ("ref",)))
    cfg, t = cases.d3q27_mrt_lid_reference(18, storage=BF16S, engine=MRT_MFMA_BF16X3)  # BF16x3, raw basis
    specs.append((cfg, t, ("ref",)))
    cfg, t = cases.d3q27_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0, storage=F32)
    specs.append((cfg, t, ("eq", (0.05, 0, 0), 9, 1e-4)))
    cfg, t = cases.d3q27_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0, storage=F32, collision=MRT_GS)
    specs.append((dict(cfg, mrt_engine=MRT_MFMA_FUSED), t, ("eq", (0.05, 0, 0), 11, 1e-4)))  # fused MFMA
    cfg, t = cases.d3q27_mrt_lid_reference(12, storage=F64, engine=MRT_SCALAR)
    specs.append((cfg, t, ("ref",)))
    cfg, t = cases.d3q27_bgk_lid_reference(20, u_lid=0.05, re=50.0, storage=F32)
    cfg = dict(cfg, nx=70)  # 70 x 20 x 20: ragged tile row
    t = geo.lid_cavity_3d_reference(70, 20, 20)
    specs.append((cfg, t, ("ref",)))
    cfg, t = cases.d3q19_sphere_channel(40, 20, 20, (12, 10, 10), 4, re=10.0, storage=F16S, collision=MRT_GS)
    specs.append((cfg, t, ("eq", (0.05, 0, 0), 13, 1e-4)))
    # PHYSICAL lid cavities 64 nodes wide (64 x 4 tiles, wavefronts of solid and fluid lanes, and
    # of solid lanes only): BGK FP32, and MRT-GS F16S on the fused MFMA engine
    om = cases.omega_from_nu(0.05 * 62 / 100.0)
    lid = geo.lid_cavity_3d_physical(64, 20, 12)
    base = dict(lattice=D3Q27, nx=64, ny=20, nz=12, bc_mode=BC_PHYSICAL, omega=om, u_wall=(0.05, 0.0, 0.0))
    specs.append((dict(base, collision=BGK, storage=F32), lid, ("eq", (0.0, 0, 0), 17, 1e-4)))
    specs.append((dict(base, collision=MRT_GS, storage=F16S, s=cases.shear_mrt_rates(om), mrt_engine=MRT_MFMA_FUSED),
                  lid, ("eq", (0.0, 0, 0), 19, 1e-4)))
    return specs


def make(spec):
    cfg, t, init = spec
    lat = Lattice(cfg, t)
    if init[0] == "ref":
        lat.init_reference(1.0)
    else:
        lat.init_equilibrium(1.0, init[1], init[2], init[3])
    return lat


@pytest.mark.parametrize("graphs,concurrent", [("1", "1"), ("0", "1"), ("1", "0")])
def test_batch_matches_individual_stepping(monkeypatch, graphs, concurrent):
    """Replayed two-step graphs (default) and direct launches (LBM_BATCH_GRAPH=0) alike, with the
    configuration groups on streams of their own (default) or all on the batch's stream
    (LBM_BATCH_CONCURRENT=0)."""
    monkeypatch.setenv("LBM_BATCH_GRAPH", graphs)
    monkeypatch.setenv("LBM_BATCH_CONCURRENT", concurrent)
    specs = member_specs()
    alone = []
    for sp in specs:
        with make(sp) as lat:
            lat.step(17)
            lat.step(10)
            alone.append(lat.state())
    lats = [make(sp) for sp in specs]
    with Batch(lats) as b:
        assert len(b) == len(lats)
        b.step(7)
        b.tune(2)  # timing launches must leave every state as it is
        b.step(19)
        b.step(1)
        got = [lat.state() for lat in lats]
        steps = [lat.steps_done for lat in lats]
    for k, (a, g) in enumerate(zip(alone, got)):
        assert steps[k] == 27
        assert np.array_equal(a, g), f"member {k} differs"
    for lat in lats:
        lat.close()


def test_batch_membership_changes():
    specs = member_specs()
    ref = []
    for sp in specs[:3]:
        with make(sp) as lat:
            lat.step(30)
            ref.append(lat.state())
    a, b_, c = (make(sp) for sp in specs[:3])
    a.step(11)  # a joins at step 11, c joins late, b leaves early
    with Batch([b_]) as bt:
        bt.step(11)
        bt.add(a)
        bt.add(c)
        bt.step(5)
        bt.remove(b_)
        b_.step(14)
        bt.step(14)
        c_missing = 30 - c.steps_done
        bt.remove(c)
        c.step(c_missing)
        assert np.array_equal(a.state(), ref[0])
    assert np.array_equal(b_.state(), ref[1])
    assert np.array_equal(c.state(), ref[2])
    with pytest.raises(LbmError):
        with Batch([a]):
            a.step(1)  # a member is stepped through its batch only
    for lat in (a, b_, c):
        lat.close()


def test_batch_grows_past_its_table_after_capture():
    """A group whose step graph was captured with room for 8 members grows to 17 and shrinks
    again: every member still matches stepping alone."""
    def lid(n):
        cfg, t = cases.d2q9_lid_reference(n, re=30.0, storage=F32)
        return cfg, t

    sizes = [16 + 2 * k for k in range(17)]
    first, late = sizes[:3], sizes[3:]
    # steps each member takes: first three 6 + 7 + 5, late ones 7 + 5, the last four late ones 7
    want = {n: (18 if n in first else (7 if n in late[-4:] else 12)) for n in sizes}
    alone = {}
    for n in sizes:
        cfg, t = lid(n)
        with Lattice(cfg, t) as lat:
            lat.init_reference(1.0)
            lat.step(want[n])
            alone[n] = lat.state()
    lats = {}
    for n in sizes:
        cfg, t = lid(n)
        lats[n] = Lattice(cfg, t)
        lats[n].init_reference(1.0)
    with Batch([lats[n] for n in first]) as b:
        b.step(6)  # captures the two-step graph
        for n in late:
            b.add(lats[n])
        b.step(7)
        for n in late[-4:]:
            b.remove(lats[n])
        b.step(5)
        got = {n: lats[n].state() for n in sizes}
    for n in sizes:
        assert lats[n].steps_done == want[n]
        assert np.array_equal(got[n], alone[n]), f"member {n} differs"
        lats[n].close()


def job_list():
    jobs = []
    rng = np.random.default_rng(5)
    for k in range(12):
        if k % 2 == 0:
            n = int(rng.integers(24, 96))
            cfg, t = cases.d2q9_lid_reference(n, re=float(rng.uniform(20, 100)), storage=F16S)
        else:
            n = int(rng.integers(12, 28))
            cfg, t = cases.d3q27_mrt_lid_gs(n, re=float(rng.uniform(20, 100)), storage=F16S)
        jobs.append((cfg, t, int(rng.integers(20, 400)), int(rng.integers(1 << 30))))
    return jobs


def test_server_alone_equals_mix_and_budget():
    jobs = job_list()
    alone = []
    for cfg, t, steps, seed in jobs:
        with Lattice(cfg, t) as lat:
            lat.init_equilibrium(1.0, (0, 0, 0), seed, 1e-4)
            lat.step(steps)
            alone.append(lat.fields())
    from lbm_amd.capi import footprint_bytes
    fps = [footprint_bytes(cfg) for cfg, _, _, _ in jobs]
    budget = 3 * max(fps)  # forces queueing and admission as memory frees up
    with Server(budget, steps_per_tick=37) as srv:
        ids = []
        for i, (cfg, t, steps, seed) in enumerate(jobs[:6]):
            ids.append(srv.submit(cfg, t, steps, 1.0, (0, 0, 0), seed, 1e-4, want_fields=True))
        results = []
        max_used = 0
        late = 6
        while len(results) < len(jobs):
            active, queued = srv.tick()
            max_used = max(max_used, srv.used_bytes)
            if late < len(jobs):  # later arrivals join while others are resident
                cfg, t, steps, seed = jobs[late]
                ids.append(srv.submit(cfg, t, steps, 1.0, (0, 0, 0), seed, 1e-4, want_fields=True))
                late += 1
            results += srv.poll()
        assert max_used <= budget
        by_id = {r["job_id"]: r for r in results}
        for k, jid in enumerate(ids):
            r = by_id[jid]
            assert r["status"] == 0 and r["steps"] == jobs[k][2]
            assert r["submit_s"] <= r["admit_s"] <= r["finish_s"]
            got = srv.take(jid)
            for a, b in zip(got, alone[k]):
                assert np.array_equal(a, b), f"job {k}"


def test_server_isolates_a_diverging_job():
    good_cfg, good_t = cases.d2q9_lid_reference(48, storage=F32)
    bad_cfg = dict(good_cfg, omega=1.99, u_wall=(0.4, 0.0, 0.0))
    with Lattice(good_cfg, good_t) as lat:
        lat.init_equilibrium(1.0, (0, 0, 0), 1, 0.0)
        lat.step(3000)
        want = lat.fields()
    with Server(1 << 30, steps_per_tick=100, check_every=2) as srv:
        g = srv.submit(good_cfg, good_t, 3000, seed=1, want_fields=True)
        b = srv.submit(bad_cfg, good_t, 3000, seed=2, noise_amp=2.0)
        done = []
        while len(done) < 2:
            srv.tick()
            done += srv.poll()
        st = {r["job_id"]: r for r in done}
        assert st[b]["status"] == 4  # LBM_EUNSTABLE, retired early
        assert st[b]["steps"] < 3000
        assert st[g]["status"] == 0 and st[g]["steps"] == 3000
        for a, w in zip(srv.take(g), want):
            assert np.array_equal(a, w)


def test_server_rejects_and_cancels():
    cfg, t = cases.d2q9_lid_reference(64, storage=F32)
    with Server(1 << 16, steps_per_tick=10) as srv:  # 64 KiB: too small for the job
        with pytest.raises(LbmError):
            srv.submit(cfg, t, 10)
    with Server(1 << 30, steps_per_tick=10) as srv:
        a = srv.submit(cfg, t, 100)
        b = srv.submit(cfg, t, 100)
        srv.tick()
        srv.cancel(a)
        assert srv.used_bytes > 0
        while True:
            active, queued = srv.tick()
            if active == 0 and queued == 0:
                break
        ids = [r["job_id"] for r in srv.poll()]
        assert ids == [b]
        assert srv.used_bytes == 0


def test_server_destroyed_with_resident_jobs():
    """Destroying a server while jobs are resident releases their lattices before the batch whose
    stream they share; lattices created afterwards step as before."""
    cfg, t = cases.d2q9_lid_reference(64, storage=F32)
    with Lattice(cfg, t) as lat:
        lat.init_equilibrium(1.0, (0, 0, 0), 3, 1e-4)
        lat.step(50)
        want = lat.fields()
    for _ in range(2):
        with Server(1 << 30, steps_per_tick=10) as srv:
            for seed in range(3):
                srv.submit(cfg, t, 1000, seed=seed, noise_amp=1e-4)
            active, queued = srv.tick()
            assert active == 3 and queued == 0
    with Lattice(cfg, t) as lat:
        lat.init_equilibrium(1.0, (0, 0, 0), 3, 1e-4)
        lat.step(50)
        for a, b in zip(lat.fields(), want):
            assert np.array_equal(a, b)


def test_server_reuses_buffers_of_finished_jobs():
    """A budget that holds one job at a time makes each job run on the device buffers the previous
    one left (64^2 and 62^2 D2Q9 lattices share every size class): each job still equals its run
    alone bit for bit, whatever the reused buffers held."""
    from lbm_amd.capi import footprint_bytes
    specs = [(cases.d2q9_lid_reference(64, storage=F32), 1), (cases.d2q9_lid_reference(62, storage=F32), 2),
             (cases.d2q9_lid_reference(64, storage=F32), 3), (cases.d2q9_lid_reference(62, storage=F32), 4)]
    alone = []
    for (cfg, t), seed in specs:
        with Lattice(cfg, t) as lat:
            lat.init_equilibrium(1.0, (0, 0, 0), seed, 1e-3)
            lat.step(300)
            alone.append(lat.fields())
    # a job that keeps its fields also holds their readout scratch, 4 doubles per node
    budget = int(1.5 * (footprint_bytes(specs[0][0][0]) + 4 * 8 * 64 * 64))
    with Server(budget, steps_per_tick=100) as srv:
        ids = [srv.submit(cfg, t, 300, seed=seed, noise_amp=1e-3, want_fields=True) for (cfg, t), seed in specs]
        done = []
        while len(done) < len(ids):
            active, queued = srv.tick()
            assert active <= 1 and srv.used_bytes <= budget
            done += srv.poll()
        assert all(r["status"] == 0 for r in done)
        for jid, want in zip(ids, alone):
            for a, b in zip(srv.take(jid), want):
                assert np.array_equal(a, b)


def test_batch_group_switching_to_single_steps_recaptures(monkeypatch):
    """One membership change swaps a D2Q9 lattice that advances several steps per launch for one of
    the same shape created with LBM_MULTISTEP=off: its group leaves the multi-step launches and
    joins the replayed two-step graph (and back again), and every member still equals its run
    alone."""
    cfg2, t2 = cases.d2q9_lid_reference(64, storage=F32)
    cfg3, t3 = cases.d3q27_bgk_lid_reference(12, u_lid=0.05, re=50.0, storage=F32)

    def lattice(cfg, t, single=False):
        if single:
            monkeypatch.setenv("LBM_MULTISTEP", "off")
        try:
            lat = Lattice(cfg, t)
        finally:
            monkeypatch.delenv("LBM_MULTISTEP", raising=False)
        lat.init_reference(1.0)
        return lat

    want = {}
    for key, cfg, t, single, n in (("a", cfg2, t2, False, 20), ("b", cfg2, t2, True, 10), ("c", cfg3, t3, False, 30)):
        with lattice(cfg, t, single) as lat:
            lat.step(n)
            want[key] = lat.state()
    a, b, c = lattice(cfg2, t2), lattice(cfg2, t2, True), lattice(cfg3, t3)
    assert a.steps_per_launch > 1 and b.steps_per_launch == 1
    with Batch([c, a]) as bt:
        bt.step(10)  # c replays the captured graph, a runs multi-step launches
        bt.remove(a)
        bt.add(b)  # same group, same tiles: now single steps, inside the graph
        bt.step(10)
        bt.remove(b)
        bt.add(a)  # and out of it again
        bt.step(10)
        got = {"a": a.state(), "b": b.state(), "c": c.state()}
    for key in want:
        assert np.array_equal(got[key], want[key]), key
    for lat in (a, b, c):
        lat.close()


def equilibrium_start(cfg, types, seed, u0, amp=1e-4):
    """rho, ux, uy, uz that lbm_init_equilibrium(1, u0, seed, amp) builds (its documented SplitMix64
    draws, vectorised): u0 plus amp * (2 r - 1) on every component of every non-solid node."""
    types = np.asarray(types).reshape(-1)
    N, D = types.size, (2 if cfg["lattice"] == 9 else 3)
    k = np.arange(N * D, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (k + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    r = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    u = np.zeros((3, N))
    for a in range(3):
        u[a] = u0[a]
    fluid = (types & 1) == 0
    noise = (amp * (2.0 * r - 1.0)).reshape(N, D)
    for a in range(D):
        u[a, fluid] += noise[fluid, a]
    return np.ones(N), u[0], u[1], u[2]


def test_equilibrium_start_restates_the_device_generator():
    cfg, t = cases.d2q9_karman(40, 24, 10, 12, 3, storage=F64)
    t = t.reshape(-1)
    start = equilibrium_start(cfg, t, 0x5EED0005, (0.05, 0, 0))
    with Lattice(cfg, t) as a, Lattice(cfg, t) as b:
        a.init_equilibrium(1.0, (0.05, 0.0, 0.0), 0x5EED0005, 1e-4)
        b.init_fields(*start)
        assert np.array_equal(a.populations(), b.populations())


def field_errors(a, b):
    e_rho = float(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0]))
    ua, ub = np.stack(a[1:4]), np.stack(b[1:4])
    return e_rho, float(np.linalg.norm(ua - ub) / max(np.linalg.norm(ub), 1e-300))


def tiny_mix():
    """Jobs of the serving mix's generator at sizes the CPU oracle runs in seconds: every
    combination of lattice (D2Q9 BGK, D3Q27 MRT-GS), geometry (lid cavity, channel with an
    obstacle) and storage (F32, F16S), two of each, 20-400 steps."""
    from lbm_amd import serving
    picked = []
    for storage in (F32, F16S):
        seen = {}
        for job in serving.make_mix(80, seed=0x5EED0005 + storage, step_scale=0.02, sides="tiny", storage=storage):
            key = (job["cfg"]["lattice"], "u_wall" in job["cfg"])
            if seen.get(key, 0) < 2:
                seen[key] = seen.get(key, 0) + 1
                picked.append(job)
        assert len(seen) == 4
    return picked


def test_server_mix_matches_oracle():
    """A tiny serving mix through the continuous-batching server (jobs joining and leaving between
    ticks of 37 steps): each job's rho/u fields against the CPU oracle running the same job from
    the same initial state.  F32 jobs: within 2e-5 rel-L2 of the FP64 oracle and 5e-5 of the
    oracle's own FP32 run (two FP32 evaluations in different summation orders, each about 1e-5
    from FP64 after a few hundred steps).  F16S jobs: within 1e-3 of the FP64 oracle (the accuracy
    the serving figure is quoted at) and 5e-4 of the oracle's run through the same codec (the
    engines' FP32 results differ from the oracle's in the last bits, so a stored half may round
    the other way; such differences are of the size of the storage rounding itself, as large as
    the distance to FP64)."""
    import oracle as orc
    from lbm_amd import serving
    jobs = tiny_mix()
    res = serving.serve(jobs, 1 << 30, steps_per_tick=37, realtime=False, keep=range(len(jobs)))
    assert res["failed"] == 0 and len(res["fields"]) == len(jobs)
    worst, bad = {}, []
    for k, job in enumerate(jobs):
        cfg, t = job["cfg"], job["types"]
        start = equilibrium_start(cfg, t, job["seed"], job["u0"])
        got = res["fields"][k]
        fo = orc.run(cfg, t, orc.init_fields(cfg, t, *start), job["steps"])
        e_o = field_errors(got, orc.fields(cfg, t, fo))
        cfg64 = dict(cfg, storage=F64)
        f64 = orc.run(cfg64, t, orc.init_fields(cfg64, t, *start), job["steps"])
        e_64 = field_errors(got, orc.fields(cfg64, t, f64))
        kind = f"{'D2Q9' if cfg['lattice'] == 9 else 'D3Q27'} {'lid' if 'u_wall' in cfg else 'channel'} " \
               f"{'F32' if cfg['storage'] == F32 else 'F16S'}"
        print(f"job {k} {kind} side {cfg['nx']} steps {job['steps']}: vs oracle rho {e_o[0]:.2e} u {e_o[1]:.2e}; "
              f"vs FP64 rho {e_64[0]:.2e} u {e_64[1]:.2e}")
        worst[kind] = max(worst.get(kind, (0.0, 0.0)), (max(e_o), max(e_64)))
        bad.append(k) if (max(e_o) > (5e-5 if cfg["storage"] == F32 else 5e-4) or
                          max(e_64) > (2e-5 if cfg["storage"] == F32 else 1e-3)) else None
    print("largest rel-L2 per kind (vs oracle in the job's storage, vs FP64 oracle):", worst)
    assert not bad, bad


def test_batch_group_larger_than_the_caches_matches_alone():
    """Two D3Q27 lattices whose buffers together exceed 256 MiB: the group takes its tiles in
    stream order (and a grid with spare workgroups); each member still equals its run alone, also
    after a third member joins and one leaves."""
    specs = [cases.d3q27_mrt_lid_gs(96, re=200.0, storage=F32), cases.d3q27_mrt_lid_gs(100, re=200.0, storage=F32),
             cases.d3q27_mrt_lid_gs(44, re=100.0, storage=F32)]  # one group: 407 MB, later 209 MB
    want = []
    for (cfg, t), n in zip(specs, (12, 7, 10)):
        with Lattice(cfg, t) as lat:
            lat.init_reference(1.0)
            lat.step(n)
            want.append(lat.state())
    lats = []
    for cfg, t in specs:
        lat = Lattice(cfg, t)
        lat.init_reference(1.0)
        lats.append(lat)
    with Batch(lats[:2]) as b:
        b.step(4)
        b.add(lats[2])
        b.step(3)
        b.remove(lats[1])
        b.step(5)
        b.remove(lats[2])
        lats[2].step(2)
        got = [lat.state() for lat in lats]
    for k in range(3):
        assert np.array_equal(got[k], want[k]), k
    for lat in lats:
        lat.close()
