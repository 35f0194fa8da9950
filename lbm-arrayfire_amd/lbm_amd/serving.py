"""Serving-mix job generator and drivers for the continuous-batching server.

The mix: jobs arrive as a Poisson process; half are D2Q9 BGK square lattices with side
log-uniform in [64, 2048], half D3Q27 MRT (Gram-Schmidt basis) cubes with side log-uniform in
[32, 192] (the "small" mix: [32, 256] and [16, 48]; the "tiny" mix, small enough for the CPU
oracle: [24, 64] and [10, 20]); Re uniform in [50, 1000]; geometry is a lid cavity or a channel
with a cylinder/sphere of radius 5-15 % of the height; steps uniform in [1000, 20000] times
``step_scale``; FP32 arithmetic with FP16-shifted storage (``storage`` selects another, or "auto":
FP32 storage for the D2Q9 jobs, FP16-shifted for the D3Q27 ones.  The D2Q9 jobs run on the
multi-step kernel, whose HBM traffic is about a quarter of its updates' bytes, so 16-bit storage
buys them no speed (the 2D half of the mix served at 103k MLUPS in FP32 against 99k in F16S) while
on long unsteady channels it costs accuracy: a cylinder channel at Re 366 over 19,198 steps ends
5.8e-3 rel-L2 in u from FP64 in F16S, 4.1e-5 in FP32).

Lattice-velocity policy (the reference fixes u and derives nu from Re): u = 0.05 for lids and
inlets, nu = u L / Re, clamped to nu >= 1/150 (tau >= 0.52), and for the D2Q9 BGK channels with a
cylinder to nu >= 1/60 (tau >= 0.55): of the 1,000-job mix, all 52 such channels with tau <= 0.535
diverged within 2,200 steps, in FP32 and FP16-shifted storage alike, while no lid and no D3Q27 MRT
job did.  The Reynolds number actually run is recorded with the job.
"""
import math
import sys
import time

import numpy as np

from . import cases, geometry as geo
from .capi import (BC_PHYSICAL, BGK, D2Q9, D3Q27, F16S, F32, INLET_EXTRAPOLATED, INLET_UNIFORM, MRT_GS, Lattice,
                   Server)

NU_MIN = 1.0 / 150.0
NU_MIN_BGK_CHANNEL = 1.0 / 60.0
U_LAT = 0.05


# side ranges (log-uniform) of the 3D and 2D jobs and the viscosity floor: the mixed workload, and
# a mix of small jobs (each launch-latency-bound when run alone), where batching their steps pays
# most; on lattices of 16-48 nodes the floor is raised to nu >= 1/30 (tau >= 0.6) to keep every
# job stable
SIDES = {"mixed": ((32, 192), (64, 2048), NU_MIN), "small": ((16, 48), (32, 256), 1.0 / 30.0),
         "tiny": ((10, 20), (24, 64), 1.0 / 30.0)}


def make_job(rng, step_scale=1.0, sides="mixed", storage=F16S):
    (lo3, hi3), (lo2, hi2), nu_min = SIDES[sides]
    d3 = rng.random() < 0.5
    if storage == "auto":
        storage = F16S if d3 else F32
    side = int(round(math.exp(rng.uniform(math.log(lo3), math.log(hi3))) if d3 else
                     math.exp(rng.uniform(math.log(lo2), math.log(hi2)))))
    re = float(rng.uniform(50, 1000))
    lid = rng.random() < 0.5
    steps = max(1, int(round(int(rng.integers(1000, 20001)) * step_scale)))
    L = side - 2
    if not d3 and not lid:
        nu_min = max(nu_min, NU_MIN_BGK_CHANNEL)
    nu = max(U_LAT * L / re, nu_min)
    re_run = U_LAT * L / nu
    om = cases.omega_from_nu(nu)
    if d3:
        cfg = dict(lattice=D3Q27, nx=side, ny=side, nz=side, collision=MRT_GS, storage=storage,
                   bc_mode=BC_PHYSICAL, omega=om, s=cases.shear_mrt_rates(om))
        if lid:
            cfg["u_wall"] = (U_LAT, 0.0, 0.0)
            types = geo.lid_cavity_3d_physical(side, side, side)
        else:
            r = max(1, int(round(rng.uniform(0.05, 0.15) * side)))
            cfg.update(u_in=U_LAT, inlet_profile=INLET_UNIFORM, inlet_closure=INLET_EXTRAPOLATED)
            types = geo.channel_3d_physical(side, side, side, side // 4, side // 2, side // 2, r)
    else:
        cfg = dict(lattice=D2Q9, nx=side, ny=side, nz=1, collision=BGK, storage=storage,
                   bc_mode=BC_PHYSICAL, omega=om)
        if lid:
            cfg["u_wall"] = (U_LAT, 0.0, 0.0)
            types = geo.lid_cavity_2d_physical(side, side)
        else:
            r = max(1, int(round(rng.uniform(0.05, 0.15) * side)))
            cfg.update(u_in=U_LAT, inlet_profile=INLET_UNIFORM)
            types = geo.channel_2d_physical(side, side, side // 4, side // 2, r)
    u0 = (0.0 if lid else U_LAT, 0.0, 0.0)
    return dict(cfg=cfg, types=types.reshape(-1), steps=steps, re=re, re_run=re_run, u0=u0,
                seed=int(rng.integers(1 << 62)), nodes=side ** (3 if d3 else 2))


def make_mix(njobs, seed=0x5EED0005, rate=50.0, step_scale=1.0, sides="mixed", storage=F16S):
    rng = np.random.default_rng(seed)
    jobs = [make_job(rng, step_scale, sides, storage) for _ in range(njobs)]
    t = np.cumsum(rng.exponential(1.0 / rate, njobs))
    for j, a in zip(jobs, t):
        j["arrival_s"] = float(a)
    return jobs


def serve(jobs, budget_bytes, steps_per_tick=200, realtime=True, keep=(), progress=0):
    """Runs the mix through the server.  realtime: submit each job at its arrival time (wall clock,
    between ticks); otherwise submit everything at t = 0.  keep: indices of jobs whose final
    rho/u fields are kept (returned under "fields", by job index).  progress: print a line to
    stderr every that many finished jobs.  Returns per-job results and totals."""
    res = []
    keep = set(keep)
    fields = {}
    with Server(budget_bytes, steps_per_tick=steps_per_tick, check_every=32) as srv:
        t0 = time.perf_counter()
        k = 0
        ids = {}
        while len(res) < len(jobs):
            now = time.perf_counter() - t0
            while k < len(jobs) and (not realtime or jobs[k]["arrival_s"] <= now):
                j = jobs[k]
                ids[srv.submit(j["cfg"], j["types"], j["steps"], 1.0, j["u0"], j["seed"], 1e-4,
                               want_fields=k in keep)] = k
                k += 1
            active, queued = srv.tick()
            before = len(res)
            res += srv.poll()
            if progress and len(res) // progress > before // progress:
                print(f"serve: {len(res)}/{len(jobs)} jobs done at {time.perf_counter() - t0:.1f} s, "
                      f"{active} resident, {queued} queued", file=sys.stderr, flush=True)
            if active == 0 and queued == 0 and k < len(jobs):
                time.sleep(max(0.0, min(0.002, jobs[k]["arrival_s"] - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        stats = srv.stats()
        for r in res:
            if ids[r["job_id"]] in keep and r["status"] == 0:
                fields[ids[r["job_id"]]] = srv.take(r["job_id"])
    # latency from the job's arrival (realtime) or its submission (everything at t = 0)
    lat = np.array([r["finish_s"] - (jobs[ids[r["job_id"]]]["arrival_s"] if realtime else r["submit_s"])
                    for r in res])
    lups = sum(r["nodes"] * r["steps"] for r in res)
    bad = sum(1 for r in res if r["status"] != 0)
    failures = []
    for r in res:
        if r["status"] != 0:
            j = jobs[ids[r["job_id"]]]
            c = j["cfg"]
            failures.append(dict(job=ids[r["job_id"]], status=int(r["status"]), steps_done=int(r["steps"]),
                                 steps=j["steps"], kind=("D2Q9 BGK" if c["lattice"] == D2Q9 else "D3Q27 MRT-GS") +
                                 (" lid" if "u_wall" in c else " channel"), side=c["nx"], re=j["re"],
                                 re_run=j["re_run"], tau=1.0 / c["omega"]))
    out = dict(wall_s=wall, jobs=len(res), failed=bad, failures=failures, p50_s=float(np.percentile(lat, 50)),
               p99_s=float(np.percentile(lat, 99)), jobs_per_s=len(res) / wall, mlups=lups / wall / 1e6,
               server=stats)
    if keep:
        out["fields"] = fields
    return out


def fp64_errors(job, fields, steps=None):
    """rel-L2 of rho and of the velocity vector field of a served job's fields against the same job
    run alone with FP64 storage (same initial state, same step count)."""
    cfg = dict(job["cfg"], storage=0)
    with Lattice(cfg, job["types"]) as lat:
        lat.init_equilibrium(1.0, job["u0"], job["seed"], 1e-4)
        lat.step(job["steps"] if steps is None else steps)
        ref = lat.fields()
    e_rho = float(np.linalg.norm(fields[0] - ref[0]) / np.linalg.norm(ref[0]))
    ua, ub = np.stack(fields[1:4]), np.stack(ref[1:4])
    return e_rho, float(np.linalg.norm(ua - ub) / max(np.linalg.norm(ub), 1e-300))


def sequential(jobs, progress=0, plain=False):
    """One job at a time in arrival order (no idle time between jobs), each on a lattice of its own.
    plain=False: the same solver the server runs (the job's storage, multi-step launches where they
    apply); plain=True: the plain solver, FP32 storage and one fused launch per step (the
    reference's one-lattice-per-program shape on this library's single-step kernel)."""
    import os
    t0 = time.perf_counter()
    lups = 0
    t_setup = t_run = t_end = 0.0
    saved = os.environ.get("LBM_MULTISTEP")
    if plain:
        os.environ["LBM_MULTISTEP"] = "off"  # read by lbm_create
    try:
        for k, j in enumerate(jobs):
            if progress and k and k % progress == 0:
                print(f"sequential{' plain' if plain else ''}: {k}/{len(jobs)} jobs done at "
                      f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            a = time.perf_counter()
            cfg = dict(j["cfg"], storage=F32) if plain else j["cfg"]
            b, c = _run_alone(cfg, j)
            d = time.perf_counter()
            t_setup += b - a
            t_run += c - b
            t_end += d - c
            lups += j["nodes"] * j["steps"]
    finally:
        if plain:
            if saved is None:
                os.environ.pop("LBM_MULTISTEP", None)
            else:
                os.environ["LBM_MULTISTEP"] = saved
    wall = time.perf_counter() - t0
    return dict(wall_s=wall, jobs=len(jobs), jobs_per_s=len(jobs) / wall, mlups=lups / wall / 1e6,
                setup_s=t_setup, run_s=t_run, release_s=t_end, storage="F32" if plain else "job's",
                launches="one per step" if plain else "multi-step where it applies")


def _run_alone(cfg, j):
    """Creates, initialises, steps and releases one job's lattice; returns the times at which the
    steps started and ended."""
    with Lattice(cfg, j["types"]) as lat:
        lat.init_equilibrium(1.0, j["u0"], j["seed"], 1e-4)
        b = time.perf_counter()
        lat.step(j["steps"])
        lat.monitors()
        c = time.perf_counter()
    return b, c
