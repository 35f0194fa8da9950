#!/usr/bin/env python3
"""Throughput benchmark of the fused LBM time step on one MI355X.

Default workload (the largest single-GPU configuration): D3Q27 BGK duct with a sphere,
512x256x256 nodes, FP32, sphere r = 24 at (128,128,128), uniform Zou-He inlet u = 0.05,
zero-gradient outlet, no-slip walls, Re = 200, 1e-4 seeded noise (seed 0x5EED0004).  A "step" is
one lattice time step over the whole lattice; MLUPS counts every node (solid included), 1e6
lattice updates/s.  ``value`` is the median of five timed regions of ``--steps`` steps each.

Prints ONE JSON line.  ``roofline.achieved`` is the algorithmic bytes of the dominant kernel
(2 * Q * bytes-per-population per node update, times the node updates one launch performs) over
its average launch duration, measured with HIP events on the lattice's stream; ``traffic`` is the
HBM bytes per launch that rocprofv3 counted for the same kernel (profiles/traffic.json) and
``traffic_frac`` those bytes over the same launch duration against the 8 TB/s peak.
``cpu_baseline`` times the CPU oracle (a restatement of the reference step, OpenMP) on a bounded
sample of the same workload on this host, plus the reference's multi-pass step order on the
128^2 D2Q9 lid and 128^3 D3Q27 MRT lid shapes.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "lbm-arrayfire_amd")]

METRIC = "MLUPS & % HBM roofline; batched jobs/s + p99 at ≤1e-3 rel-L2 vs FP64"
HBM_PEAK_GBS = 8000.0


def workloads():
    from lbm_amd import cases
    from lbm_amd.capi import BF16S, F16S, F32, F64, MRT_GS, MRT_MFMA_BF16X3, MRT_MFMA_F32, MRT_MFMA_FUSED, MRT_SCALAR

    W = {
        "d2q9_bgk_karman_1024x256_fp32": lambda: cases.d2q9_karman(1024, 256, 256, 128, 20, storage=F32),
        "d2q9_bgk_karman_1024x256_f16s": lambda: cases.d2q9_karman(1024, 256, 256, 128, 20, storage=F16S),
        # the same lattices with one launch per step (the multi-step launches switched off)
        "d2q9_bgk_karman_1024x256_fp32_1step": lambda: cases.d2q9_karman(1024, 256, 256, 128, 20, storage=F32),
        "d2q9_bgk_karman_1024x256_f16s_1step": lambda: cases.d2q9_karman(1024, 256, 256, 128, 20, storage=F16S),
        "d2q9_bgk_karman_1024x256_fp64": lambda: cases.d2q9_karman(1024, 256, 256, 128, 20, storage=F64),
        "d2q9_bgk_karman_1024x256_fp64_1step": lambda: cases.d2q9_karman(1024, 256, 256, 128, 20, storage=F64),
        "d3q27_mrt_gs_lid_256_fp32": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F32),
        "d3q27_mrt_gs_lid_256_f16s": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F16S),
        "d3q27_mrt_gs_lid_256_fp32_mfma2": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F32, engine=MRT_MFMA_F32),
        "d3q27_mrt_gs_lid_256_fp32_fused": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F32, engine=MRT_MFMA_FUSED),
        "d3q27_mrt_gs_lid_256_fp32_scalar": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F32, engine=MRT_SCALAR),
        "d3q27_mrt_gs_lid_256_f16s_mfma2": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F16S, engine=MRT_MFMA_F32),
        "d3q27_mrt_gs_lid_256_f16s_fused": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F16S, engine=MRT_MFMA_FUSED),
        "d3q27_mrt_gs_lid_256_f16s_scalar": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F16S, engine=MRT_SCALAR),
        "d3q27_mrt_gs_lid_256_fp32_bf16x3": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F32, engine=MRT_MFMA_BF16X3),
        "d3q27_mrt_gs_lid_256_f16s_bf16x3": lambda: cases.d3q27_mrt_lid_gs(256, re=400.0, storage=F16S, engine=MRT_MFMA_BF16X3),
        "d3q27_bgk_sphere_512x256x256_fp32": lambda: cases.d3q27_sphere_channel(storage=F32),
        "d3q27_bgk_sphere_512x256x256_f16s": lambda: cases.d3q27_sphere_channel(storage=F16S),
        "d3q27_bgk_sphere_512x256x256_bf16s": lambda: cases.d3q27_sphere_channel(storage=BF16S),
        "d3q27_mrt_gs_sphere_512x256x256_f16s": lambda: cases.d3q27_sphere_channel(storage=F16S, collision=MRT_GS),
        # the reference's work-in-progress MRT channel: 27 velocities, Suga's rates, 320x80x80
        "d3q27_mrt_gs_suga_channel_320x80x80_fp32": lambda: cases.d3q27_mrt_channel_suga(storage=F32),
        "d3q19_bgk_sphere_512x256x256_fp32": lambda: cases.d3q19_sphere_channel(storage=F32),
        "d3q19_bgk_sphere_512x256x256_f16s": lambda: cases.d3q19_sphere_channel(storage=F16S),
        "d3q19_mrt_lid_256_fp32": lambda: cases.d3q19_lid(256, re=400.0, storage=F32, collision=MRT_GS),
        "d3q19_mrt_lid_256_f16s": lambda: cases.d3q19_lid(256, re=400.0, storage=F16S, collision=MRT_GS),
    }
    return W


def seed_of(name):
    """The recorded noise seed of each configuration: 0x5EED0004 for the sphere ducts, 0x5EED0002
    otherwise."""
    return 0x5EED0004 if "sphere" in name else 0x5EED0002


def measure(name, steps, warmup, regions=1):
    """Times `regions` regions of `steps` steps each (HIP events on the lattice's stream around each
    region, the stream synchronised before and after); the figures are those of the median region."""
    from lbm_amd import Lattice
    from lbm_amd.capi import STORAGE_BYTES

    cfg, types = workloads()[name]()
    saved = os.environ.get("LBM_MULTISTEP")
    if name.endswith("_1step"):
        os.environ["LBM_MULTISTEP"] = "off"
    try:
        lat = Lattice(cfg, types)
    finally:
        if saved is None:
            os.environ.pop("LBM_MULTISTEP", None)
        else:
            os.environ["LBM_MULTISTEP"] = saved
    with lat:
        k = lat.steps_per_launch
        # whole launches only, so the timed region holds steps // k launches of one kernel
        steps, warmup = max(k, steps - steps % k), warmup - warmup % k
        u0 = (cfg.get("u_in", 0.0), 0.0, 0.0)
        lat.init_equilibrium(1.0, u0, seed_of(name), 1e-4)
        lat.step(warmup)
        lat.sync()
        region_ms, walls = [], []
        for _ in range(regions):
            lat.sync()
            t0 = time.perf_counter()
            region_ms.append(lat.step_timed(steps))  # returns after the stream has drained
            walls.append(time.perf_counter() - t0)
        _, _, bad = lat.monitors()
    N = lat.N
    Q = cfg["lattice"]
    order = sorted(range(regions), key=lambda i: region_ms[i])
    mid = order[regions // 2]
    ms = region_ms[mid]
    # algorithmic bytes: every population read and written once per node update, k updates of
    # every node per launch
    bytes_per_launch = 2 * Q * STORAGE_BYTES[cfg["storage"]] * N * k
    launch_s = ms * 1e-3 / (steps // k)
    achieved = bytes_per_launch / launch_s / 1e9
    return dict(cfg=cfg, N=N, Q=Q, steps=steps, warmup=warmup, ms=ms, wall_s=walls[mid],
                mlups=N * steps / (ms * 1e3), ms_per_step=ms / steps, steps_per_launch=k,
                launch_us=launch_s * 1e6, bytes_per_launch=bytes_per_launch, achieved_gbs=achieved,
                region_mlups=[N * steps / (t * 1e3) for t in region_ms], nonfinite=int(bad),
                nt=bool(Q != 9 and N * Q * STORAGE_BYTES[cfg["storage"]] > (256 << 20)))


def pcie_inclusive_leg(name, step_counts=None):
    """The workload as a host caller runs it end to end through the C ABI: create the lattice
    (node types uploaded), initialise it from host fields (rho, ux, uy uploaded), step, read rho
    and u back to host, destroy.  Host wall time around all of it; never the headline value, whose
    inputs are resident in HBM when its timed region starts."""
    import numpy as np
    from lbm_amd import Lattice

    cfg, types = workloads()[name]()
    N = cfg["nx"] * cfg["ny"] * cfg["nz"]
    if step_counts is None:  # about 0.02 s and 2 s of steps at 25k MLUPS
        step_counts = tuple(max(1, int(v * 25e9 / N)) for v in (0.02, 2.0))
    rho, ux, uy = np.ones(N), np.full(N, 0.05), np.zeros(N)
    up = N * (1 + 3 * 8)  # node types + three double fields
    down = N * 4 * 8  # rho, ux, uy, uz as doubles
    out = {"sequence": "create + upload types and fields + steps + download fields + destroy",
           "host_to_device_bytes": up, "device_to_host_bytes": down}
    for steps in step_counts:
        t0 = time.perf_counter()
        with Lattice(cfg, types) as lat:
            lat.init_fields(rho, ux, uy, None)
            lat.step(steps)
            lat.fields()
        dt = time.perf_counter() - t0
        out[f"steps_{steps}"] = {"wall_s": round(dt, 5), "mlups": round(N * steps / dt / 1e6, 1)}
    return out


def shanchen_leg(steps=200, warmup=20, n=1024):
    """Two-component Shan-Chen D2Q9 on an n x n periodic lattice (the reference's
    cpp/d2q9_shanchen.cpp at a GPU-sized lattice), G = -1.2, omega = 1, rho = 1 +- 0.001 noise.
    Algorithmic bytes per node update: both fluids' populations read and written once,
    2 x 2 x 9 s."""
    import numpy as np
    from lbm_amd.capi import F32, F64, ShanChen

    out = {}
    rng = np.random.default_rng(0)
    d = -0.001 * (1 - 2 * rng.random(n * n))
    for storage, tag, s in ((F32, "fp32", 4), (F64, "fp64", 8)):
        with ShanChen(n, n, 1.0, 1.0, -1.2, storage) as sc:
            sc.init(1 + d, 1 - d)
            sc.step(warmup)
            ms = sc.step_timed(steps)
            r1, r2 = sc.fields()[:2]
        gbs = 2 * 2 * 9 * s * n * n * steps / (ms * 1e-3) / 1e9
        out[f"d2q9_shanchen_{n}x{n}_{tag}"] = {
            "mlups": round(n * n * steps / (ms * 1e3), 1), "ms_per_step": round(ms / steps, 5),
            "kernels_per_step": 1, "achieved_gbs": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
            "steps": steps, "nonfinite": int((~np.isfinite(r1)).sum() + (~np.isfinite(r2)).sum())}
    return out


def kernel_name(cfg, k, nt=False):
    """The kernel the timed launches run (as rocprofv3 lists it, template arguments abridged)."""
    if k > 1:
        return f"lbm::k_step_multi<storage {cfg['storage']}, bc {cfg.get('bc_mode', 0)}, {k} steps per launch>"
    tail = ", true" if nt else ""
    return f"lbm::k_step<{cfg['lattice']}, {cfg['storage']}, {cfg.get('bc_mode', 0)}, {cfg.get('collision', 0)}{tail}>"


def traffic_for(name):
    """The counted-traffic record of a workload (profiles/traffic.json), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get(name)
    except (OSError, ValueError):
        return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def _oracle_rate(orc, cfg, types, threads, target_s, warmup=1):
    """MLUPS of the oracle over about target_s seconds (at least one timed step)."""
    N = types.size
    t1 = orc.time_steps(cfg, types, 1, threads=threads, warmup=warmup)
    n = max(1, int(target_s / max(t1, 1e-6)))
    t = t1 if n == 1 else orc.time_steps(cfg, types, n, threads=threads, warmup=0)
    return N * n / (t * 1e6), n, t


def cpu_baseline(name, target_s=10.0):
    """The CPU oracle on this host.  value: the headline workload itself (FP32, OpenMP on every
    core the process may use), a bounded number of steps.  Beside it, FP64 on all cores, one core
    on a 16-plane slab of the same duct (the per-node work of the full lattice, a step short
    enough for one core), and the reference's multi-pass step order (materialised pull index,
    gather, reductions, overrides, collision at every node, bounce-back scatter) on the 128^2
    D2Q9 lid cavity and the 128^3 D3Q27 MRT (Gram-Schmidt) lid cavity, FP32 and FP64, all cores
    and one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from lbm_amd import cases

    threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    cfg, types = workloads()[name]()
    types = types.reshape(-1)
    v, n, t = _oracle_rate(orc, cfg, types, threads, target_s)
    model = _cpu_model()
    out = {"value": v, "unit": "MLUPS", "cores": threads, "kind": "port",
           "sample": f"{name}: {n} steps of the whole lattice by the CPU oracle (restated reference step, same "
                     f"boundary rules, FP32, OpenMP {threads} threads) in {t:.1f} s after one untimed step, "
                     f"on {model}"}
    out["fp64_mlups"] = _oracle_rate(orc, dict(cfg, storage=0), types, threads, 3.0)[0]
    if cfg["lattice"] == 27 and cfg.get("nz", 1) > 16:
        nz = 16
        slab = (cfg["nx"] * cfg["ny"] * nz)
        cfg1, types1 = dict(cfg, nz=nz), types[:slab].copy()
        types1[-cfg["nx"] * cfg["ny"]:] = types.reshape(cfg["nz"], -1)[-1]  # keep the closing wall plane
        out["single_core_mlups"] = _oracle_rate(orc, cfg1, types1, 1, 3.0)[0]
        out["single_core_sample"] = f"planes z < {nz} of the duct (the wall plane z = nz-1 closing them)"
    else:
        out["single_core_mlups"] = _oracle_rate(orc, cfg, types, 1, 3.0)[0]
    ref = {}
    for key, make in (("d2q9_bgk_lid_128x128_reference_order", lambda st: cases.d2q9_lid_reference(128, storage=st)),
                      ("d3q27_mrt_gs_lid_128cubed_reference_order",
                       lambda st: cases.d3q27_mrt_lid_gs(128, re=400.0, storage=st))):
        row = {}
        for st, tag in ((1, "fp32"), (0, "fp64")):
            c, ty = make(st)
            ty = ty.reshape(-1)
            row[f"{tag}_mlups_{threads}_threads"] = _oracle_rate(orc, c, ty, threads, 2.0)[0]
            row[f"{tag}_mlups_1_thread"] = _oracle_rate(orc, c, ty, 1, 2.0)[0]
        ref[key] = row
    out["reference_order"] = ref
    out["cpu_model"] = model
    return out


def _rounded(d):
    """Floats of a (nested) result rounded to 4 decimals; other values as they are."""
    if isinstance(d, dict):
        return {k: _rounded(v) for k, v in d.items()}
    if isinstance(d, (list, tuple)):
        return [_rounded(v) for v in d]
    return round(d, 4) if isinstance(d, float) else d


def serving_leg(njobs, scale, sides="mixed", accuracy_every=10,
                modes=("realtime", "saturated", "sequential", "plain"), storage="auto"):
    """Serving mix (Poisson arrivals at 50 jobs/s, half D2Q9 BGK 64^2-2048^2, half D3Q27 MRT
    32^3-192^3, FP16-shifted storage; sides="small": 32^2-256^2 and 16^3-48^3) through the
    continuous-batching server, in real time and saturated, against the same jobs run one at a
    time: "sequential" with the solver the server runs (same storage and kernels), "plain" with the
    plain solver (FP32 storage, one fused launch per step).  Accuracy: every accuracy_every-th job
    is served again (alone in a server of its own kind of traffic: the sample, saturated) with its
    fields kept, and compared with the same job run in FP64 storage."""
    import numpy as np

    from lbm_amd import serving

    from lbm_amd.capi import F16S

    jobs = serving.make_mix(njobs, step_scale=scale, sides=sides, storage="auto" if storage == "auto" else F16S)
    # admission budget: 4 GiB of resident jobs measured best on the mixed mix (1,000 jobs, steps x 0.2:
    # 31.8k MLUPS and p50 36.7 s, against 30.5k and 53.4 s with 64 GiB, where hundreds of jobs share
    # every tick)
    budget = 4 << 30
    prog = 100 if njobs >= 500 else 0
    rt = serving.serve(jobs, budget, realtime=True, progress=prog) if "realtime" in modes else None
    sat = serving.serve(jobs, budget, realtime=False, progress=prog) if "saturated" in modes else None
    seq = serving.sequential(jobs, progress=prog) if "sequential" in modes else None
    plain = serving.sequential(jobs, progress=prog, plain=True) if "plain" in modes else None
    served = sat or rt
    out = {"mix": sides, "jobs": njobs, "arrival_rate_per_s": 50.0, "step_scale": scale, "budget_bytes": budget,
           "storage": "D2Q9 FP32, D3Q27 F16S" if storage == "auto" else "F16S",
           "realtime": rt and _rounded(rt), "saturated": sat and _rounded(sat),
           "one_at_a_time": seq and _rounded({k: v for k, v in seq.items() if not isinstance(v, str)}),
           "plain_one_at_a_time": plain and _rounded({k: v for k, v in plain.items() if not isinstance(v, str)}),
           "throughput_vs_one_at_a_time": round(served["mlups"] / seq["mlups"], 3) if served and seq else None,
           "throughput_vs_plain_solver": round(served["mlups"] / plain["mlups"], 3) if served and plain else None,
           "baselines": {"one_at_a_time": "each job alone with the server's solver: same storage, same kernels, "
                                          "multi-step launches where they apply",
                         "plain_solver": "each job alone, FP32 storage, one fused launch per step"},
           "results_equal_alone_and_in_mix": "bitwise (tests/test_gpu_batch_serve.py)",
           "total_lattice_updates": sum(j["nodes"] * j["steps"] for j in jobs)}
    if accuracy_every:
        sample = jobs[::accuracy_every]
        acc = serving.serve(sample, budget, realtime=False, keep=range(len(sample)))
        errs, worst = [], None
        for k, j in enumerate(sample):
            if k not in acc["fields"]:
                continue
            e = serving.fp64_errors(j, acc["fields"][k])
            errs.append(e)
            if prog:
                print(f"accuracy: job {k + 1}/{len(sample)} rel-L2 rho {e[0]:.2e} u {e[1]:.2e}", file=sys.stderr, flush=True)
            if worst is None or e[1] > worst["rel_l2_u"]:
                c = j["cfg"]
                worst = {"job": f"{'D2Q9 BGK' if c['lattice'] == 9 else 'D3Q27 MRT-GS'} "
                                f"{'lid' if 'u_wall' in c else 'channel'} side {c['nx']}, {j['steps']} steps, "
                                f"Re {j['re_run']:.0f}", "rel_l2_rho": e[0], "rel_l2_u": e[1]}
        out["accuracy_vs_fp64"] = {
            "sample": f"every {accuracy_every}th job ({len(sample)} jobs, {len(errs)} finished)",
            "max_rel_l2_rho": float(max(e[0] for e in errs)) if errs else None,
            "max_rel_l2_u": float(max(e[1] for e in errs)) if errs else None,
            "median_rel_l2_u": float(np.median([e[1] for e in errs])) if errs else None,
            "worst_job": worst}
    return out


def copy_bandwidth(mib=1024):
    """GB/s of a plain streaming copy (read + write) of mib MiB: the attainable HBM rate.  Measured
    in a child process (torch brings its own HIP runtime, which cannot share a process with the
    library's)."""
    import subprocess

    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-calibration", str(mib)],
                       capture_output=True, text=True, timeout=300)
    for line in r.stdout.splitlines():
        if line.startswith("{"):
            return json.loads(line).get("copy_gbs")
    return None


def pull_ceilings():
    """GB/s of the D3Q27 step's access pattern with no arithmetic (27 periodic pulls and stores per
    node of a 256^3 lattice in the library's padded SoA layout; 4- and 2-byte populations) and of a
    16-byte-per-lane copy of the same bytes: lib/pull_ceiling, built from csrc/pull_ceiling.hip."""
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lbm-arrayfire_amd", "lib", "pull_ceiling")
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    for line in r.stdout.splitlines():
        if line.startswith("{"):
            return json.loads(line)
    return None


def accuracy_run(kind, steps=None):
    """Accuracy of each storage (and MRT engine) against the FP64 kernel on one configuration at its
    stated size: rel-L2 of rho and of the velocity vector field after the stated step count.
    kind "sphere": the 512x256x256 D3Q27 BGK sphere duct (Re 200, 10,000 steps; F32, F16S, BF16S;
    drag coefficient on the sphere's box).  kind "lid128": the 128^3 D3Q27 MRT-GS lid cavity (Re
    400, 20,000 steps; FP32 scalar and both MFMA engines, F16S, BF16S).  Prints progress lines and
    returns a dict."""
    import numpy as np

    from lbm_amd import Lattice, cases
    from lbm_amd.capi import BF16S, F16S, F32, F64, MRT_MFMA_BF16X3, MRT_MFMA_F32, MRT_MFMA_FUSED, MRT_SCALAR

    def errs(a, b):
        ua, ub = np.stack(a[1:4]), np.stack(b[1:4])
        return (float(np.linalg.norm(a[0] - b[0]) / np.linalg.norm(b[0])),
                float(np.linalg.norm(ua - ub) / np.linalg.norm(ub)))

    out = {"kind": kind}
    if kind == "sphere":
        steps = steps or 10000
        c, r = (128, 128, 128), 24
        box = (c[0] - r - 2, c[0] + r + 3, c[1] - r - 2, c[1] + r + 3, c[2] - r - 2, c[2] + r + 3)
        runs = [("fp64", F64, None), ("fp32", F32, None), ("f16s", F16S, None), ("bf16s", BF16S, None)]
    else:
        steps = steps or 20000
        runs = [("fp64", F64, MRT_SCALAR), ("fp32_scalar", F32, MRT_SCALAR), ("fp32_mfma2", F32, MRT_MFMA_F32),
                ("fp32_fused", F32, MRT_MFMA_FUSED), ("fp32_bf16x3", F32, MRT_MFMA_BF16X3),
                ("f16s_scalar", F16S, MRT_SCALAR), ("f16s_bf16x3", F16S, MRT_MFMA_BF16X3),
                ("bf16s_scalar", BF16S, MRT_SCALAR)]
    ref = None
    for tag, st, en in runs:
        if kind == "sphere":
            cfg, types = cases.d3q27_sphere_channel(storage=st)
        else:
            cfg, types = cases.d3q27_mrt_lid_gs(128, re=400.0, storage=st, engine=en)
        t0 = time.perf_counter()
        with Lattice(cfg, types) as lat:
            if kind == "sphere":
                lat.init_equilibrium(1.0, (cfg["u_in"], 0.0, 0.0), 0x5EED0004, 1e-4)
            else:
                lat.init_fields()
            lat.step(steps)
            flds = lat.fields()
            row = {"nonfinite": int(lat.monitors()[2])}
            if kind == "sphere":
                row["c_d"] = float(2 * lat.force(box)[0] / (cfg["u_in"] ** 2 * np.pi * r * r))
        row["wall_s"] = round(time.perf_counter() - t0, 2)
        if ref is None:
            ref = (flds, row)
        else:
            row["rel_l2_rho"], row["rel_l2_u"] = errs(flds, ref[0])
            if "c_d" in row:
                row["c_d_rel_diff"] = abs(row["c_d"] / ref[1]["c_d"] - 1)
        out[tag] = row
        print(f"accuracy {kind} {tag}: {row}", flush=True)
    out["steps"] = steps
    return out


def mrt_flop_per_lup(Q):
    """The two dense Q x Q transforms (M f and M^-1 m*) per node update."""
    return 2 * (2 * Q * Q)


def copy_calibration(mib):
    import torch

    n = (mib << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.neg(a, out=b)  # first launch: one-time setup, not timed
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        torch.neg(a, out=b)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * n * 4 * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4, "launches": 6,
                      "copy_gbs": gbs}))


def wall_layouts(side, steps, storage_name):
    """The cost of walls by layout: one D3Q27 BGK PHYSICAL side^3 lattice per node-type layout (no
    walls; solid planes x = 0 and x = side-1; x = 0 only; y and z faces; the lid cavity), each
    stepped 20 times, then 3 x ``steps`` steps timed, best of three.  steps = 0: 10 steps per
    layout and no timing (the run rocprofv3 counts)."""
    import numpy as np

    from lbm_amd import Lattice, cases, geometry as geo
    from lbm_amd.capi import BC_PHYSICAL, BGK, D3Q27, F16S, F32

    S = side
    st = {"f16s": F16S, "fp32": F32}[storage_name]
    cfg = dict(lattice=D3Q27, nx=S, ny=S, nz=S, collision=BGK, storage=st, bc_mode=BC_PHYSICAL,
               omega=cases.omega_from_nu(0.05 * S / 500), u_wall=(0.05, 0.0, 0.0))
    z, y, x = np.meshgrid(np.arange(S), np.arange(S), np.arange(S), indexing="ij")
    layouts = {
        "no_walls": np.zeros((S, S, S), np.uint8),
        "x_walls": np.where((x == 0) | (x == S - 1), geo.NODE_SOLID, 0).astype(np.uint8),
        "one_x_wall": np.where(x == 0, geo.NODE_SOLID, 0).astype(np.uint8),
        "y_z_walls": np.where((y == 0) | (y == S - 1) | (z == 0) | (z == S - 1), geo.NODE_SOLID, 0).astype(np.uint8),
        "lid_cavity": geo.lid_cavity_3d_physical(S, S, S),
    }
    out = {}
    for name, t in layouts.items():
        print(f"wall layout {name}, {S}^3 {storage_name}", file=sys.stderr, flush=True)
        with Lattice(cfg, t.reshape(-1)) as lat:
            lat.init_equilibrium(1.0, (0.0, 0.0, 0.0), 3, 1e-4)
            if steps == 0:
                lat.step(10)
                lat.sync()
                continue
            lat.step(20)
            lat.sync()
            best = float("inf")
            for _ in range(3):
                t0 = time.perf_counter()
                lat.step(steps)
                lat.sync()
                best = min(best, time.perf_counter() - t0)
            out[name] = round(S ** 3 * steps / best / 1e6, 1)
    return {"wall_layouts_mlups": out, "side": S, "storage": storage_name, "steps": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="steps per timed region (five regions)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="d3q27_bgk_sphere_512x256x256_fp32")
    ap.add_argument("--also", default="default",
                    help="comma list of extra workloads timed briefly, 'none', or 'default'")
    ap.add_argument("--also-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pcie", action="store_true", help="skip the host-buffer (PCIe-inclusive) leg")
    ap.add_argument("--serve-only", action="store_true",
                    help="only run the serving leg of the mixed mix and print it (with --serve-jobs, "
                         "--serve-scale, --serve-modes)")
    ap.add_argument("--serve-modes", default="realtime,saturated,sequential,plain",
                    help="serving runs of --serve-only: realtime, saturated, sequential (same solver), plain "
                         "(FP32, one launch per step), accuracy (the FP64 comparison of a sample)")
    ap.add_argument("--accuracy", choices=["sphere", "lid128"], default=None,
                    help="only run the accuracy comparison of one configuration against FP64 and print it")
    ap.add_argument("--serve-storage", choices=["auto", "f16s"], default="auto",
                    help="storage of the serving mix's jobs: auto (D2Q9 FP32, D3Q27 FP16-shifted) or f16s (all)")
    ap.add_argument("--serve-jobs", type=int, default=120,
                    help="jobs of the serving mix (0 skips the serving leg)")
    ap.add_argument("--serve-scale", type=float, default=0.05,
                    help="scale of the mix's 1k-20k steps per job")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 6 plain streaming copies of MIB MiB (out = -in, float32): the "
                         "known-size kernel whose FETCH_SIZE / WRITE_SIZE counters scale the step "
                         "kernel's counters to bytes")
    ap.add_argument("--wall-layouts", type=int, default=0, metavar="SIDE",
                    help="only time D3Q27 BGK PHYSICAL SIDE^3 lattices with five wall layouts "
                         "(--wall-steps 0: ten untimed steps each, for counters)")
    ap.add_argument("--wall-steps", type=int, default=100)
    ap.add_argument("--wall-storage", choices=["f16s", "fp32"], default="f16s")
    args = ap.parse_args()
    if args.wall_layouts > 0:
        print(json.dumps(wall_layouts(args.wall_layouts, args.wall_steps, args.wall_storage)))
        return
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.accuracy:
        print(json.dumps(accuracy_run(args.accuracy)))
        return
    if args.serve_only:
        modes = tuple(args.serve_modes.split(","))
        print(json.dumps(serving_leg(args.serve_jobs, args.serve_scale, "mixed",
                                     accuracy_every=50 if "accuracy" in modes else 0, modes=modes,
                                     storage=args.serve_storage)))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the solver has no "
              "multi-GPU path in this build.", file=sys.stderr)
        sys.exit(2)

    copy_gbs = copy_bandwidth()
    names = [] if args.also == "none" else (
        [k for k in workloads() if k != args.workload] if args.also == "default" else args.also.split(","))
    ceilings = pull_ceilings() if any(k.startswith(("d3q27", "d3q19")) for k in names + [args.workload]) else None
    r = measure(args.workload, args.steps, args.warmup, regions=5)
    also = {}
    for k in names:
        try:
            m = measure(k, args.also_steps, 20)
            ceil = None
            if ceilings and m["cfg"]["lattice"] in (19, 27) and m["cfg"]["storage"] in (1, 2, 3):
                # 3D lattices over 256 MiB per buffer step with non-temporal stores
                c = m["cfg"]
                nt = c["nx"] * c["ny"] * c["nz"] * c["lattice"] * (4 if c["storage"] == 1 else 2) > (256 << 20)
                pre = "d3q19_" if c["lattice"] == 19 else ""
                ceil = ceilings.get(pre + ("pull_4byte" if c["storage"] == 1 else "pull_2byte") + ("_nt_gbs" if nt else "_gbs"))
            also[k] = {"mlups": round(m["mlups"], 1), "ms_per_step": round(m["ms_per_step"], 5),
                       "steps_per_launch": m["steps_per_launch"], "achieved_gbs": round(m["achieved_gbs"], 1),
                       **({"rate_kind": "algorithmic-equivalent: several steps per launch from an on-chip "
                                        "tile, HBM traffic about 0.24x the algorithmic bytes"}
                          if m["steps_per_launch"] > 1 else {}),
                       **({"frac_of_pull_ceiling": round(m["achieved_gbs"] / ceil, 4)} if ceil else {}),
                       **({"algorithmic_mrt_tflops": round(m["mlups"] * 1e6 * mrt_flop_per_lup(m["Q"]) / 1e12, 2),
                           "frac_fp32_matrix_peak": round(m["mlups"] * 1e6 * mrt_flop_per_lup(m["Q"]) / 157.3e12, 4)}
                          if m["cfg"].get("collision", 0) else {}),
                       "frac_hbm": round(m["achieved_gbs"] / HBM_PEAK_GBS, 4), "steps": m["steps"],
                       "nonfinite": m["nonfinite"]}
        except Exception as e:  # noqa: BLE001 — report, keep the headline line
            also[k] = {"error": str(e)}
    serving = serving_leg(args.serve_jobs, args.serve_scale) if args.serve_jobs > 0 else None
    serving_small = serving_leg(args.serve_jobs, args.serve_scale * 2, "small") if args.serve_jobs > 0 else None
    try:
        shanchen = None if args.also == "none" else shanchen_leg()
    except Exception as e:  # noqa: BLE001 — report, keep the headline line
        shanchen = {"error": str(e)}
    try:
        pcie = None if args.no_pcie else pcie_inclusive_leg(args.workload)
    except Exception as e:  # noqa: BLE001 — report, keep the headline line
        pcie = {"error": str(e)}
    cpu = None if args.no_cpu else cpu_baseline(args.workload)
    cfg = r["cfg"]
    tr = traffic_for(args.workload)
    traffic = tr.get("hbm_bytes_per_launch") if tr else None
    launch_s = r["launch_us"] * 1e-6
    # HBM-bound when the counted traffic is (nearly) the algorithmic traffic; a launch that moves
    # much less (several steps from on-chip tiles) is bound by issue and latency instead
    ratio = traffic / r["bytes_per_launch"] if traffic else None
    bound = "hbm" if ratio is None or ratio >= 0.9 else "issue/latency (temporal blocking)"
    out = {
        "metric": METRIC,
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": r["warmup"],
        "ms_per_step": round(r["ms_per_step"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if cfg["storage"] in (1, 2, 3) else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "nx": cfg["nx"], "ny": cfg["ny"], "nz": cfg.get("nz", 1),
                   "lattice": f"D{2 if cfg['lattice'] == 9 else 3}Q{cfg['lattice']}",
                   "collision": ["BGK", "MRT_RAW", "MRT_GS"][cfg.get("collision", 0)],
                   "storage": ["F64", "F32", "F16S", "BF16S"][cfg["storage"]],
                   "bc_mode": ["REFERENCE", "PHYSICAL"][cfg.get("bc_mode", 0)],
                   "parallelism": "single-gpu"},
        "roofline": {"bound": bound, "achieved": round(r["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_gbs": None if not traffic else round(traffic / launch_s / 1e9, 1),
                     "traffic_frac": None if not traffic else round(traffic / launch_s / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic_over_algorithmic": None if ratio is None else round(ratio, 4),
                     "traffic_source": None if not tr else ("profiles/traffic.json (rocprofv3 FETCH_SIZE, WRITE_SIZE: the L2's "
                                                             "fills from and writes to the memory fabric, Infinity "
                                                             "Cache hits included)"),
                     "kernel": kernel_name(cfg, r["steps_per_launch"], r["nt"]),
                     "steps_per_launch": r["steps_per_launch"], "avg_launch_us": round(r["launch_us"], 3),
                     "algorithmic_bytes_per_launch": r["bytes_per_launch"],
                     "attainable_copy_gbs": None if copy_gbs is None else round(copy_gbs, 1),
                     "frac_of_attainable_copy": None if not copy_gbs else round(r["achieved_gbs"] / copy_gbs, 4),
                     **({"pull_ceiling_gbs": ceilings.get("pull_4byte_nt_gbs" if cfg["storage"] == 1 else "pull_2byte_nt_gbs"),
                         "frac_of_pull_ceiling": round(r["achieved_gbs"] / ceilings.get(
                             "pull_4byte_nt_gbs" if cfg["storage"] == 1 else "pull_2byte_nt_gbs"), 4),
                         "device_copy_gbs": ceilings.get("copy_nt_gbs")}
                        if ceilings and cfg["lattice"] == 27 and cfg["storage"] in (1, 2, 3) and r["nt"] else {})},
        "regions": {"mlups": [round(v, 1) for v in r["region_mlups"]], "value_is": "median of the regions"},
        "cpu_baseline": cpu,
        "wall_s_timed": round(r["wall_s"], 4),
        "nonfinite": r["nonfinite"],
        "also": also,
        "ceilings_3d": ceilings,
        "serving": serving,
        "serving_small_jobs": serving_small,
        "shanchen": shanchen,
        "pcie_inclusive": pcie,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
