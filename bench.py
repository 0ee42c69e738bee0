"""Benchmark of the MI355X-native SFINCS explicit step.

Workload (one GPU): configuration 1 -- regular 2048 x 2048 grid, dx = 25 m, 10-level
subgrid tables from a seeded 5 m fBm DEM, 8 point discharges, 1.5 m tidal boundary.
A step is one model time step (forcing + momentum with CFL reduction + continuity) of
the whole grid with all state resident in HBM; value = cell-updates/s.

Prints one JSON line. `roofline.achieved` is the momentum kernel's algorithmic bytes per
launch (per-branch byte counts below times the branch counters of the device) divided by
its mean launch time, measured with HIP events on the launch stream.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# Bytes each sweep touches per element class (dense layout; neighbour values are counted
# once, with the element that owns them). Momentum per cell: zs 8 (+ V 8 + zsderv 4 for
# subgrid, zb 4 for regular). Per uv-point by branch:
#   copied (kcuv not 1/6): kcuv 1 + old q,uv 8 + new q,uv 8                      = 17
#   dry:                   kcuv 1 + zmin/zbuvmx 4 + new q,uv 8 + kfuv 1           = 14
#     (per-point kernel:   + kfuv read 1                                          = 15)
#   dry, zeros held:       kcuv 1 + kfuv 1 + zmin/zbuvmx 4 (per-point kernel)     =  6
#   quiescent tile skipped: its flags, 5 B per tile of 512 points (P = 2)
#   wet (regular):         kcuv 1 + zbuvmx 4 + gn2uv 4 + kfuv 2 + mask 1 + q,uv 16 = 28
#   full (subgrid):        kcuv 1 + zmin 4 + s4 16 + fnfit 4 + kfuv 2 + mask 1 + q,uv 16 = 44
#   partial (subgrid):     kcuv 1 + zmin 4 + s4 16 + 2 levels 32 + kfuv 2 + mask 1 + q,uv 16 = 72
MOM_CELL = {True: 20, False: 12}
MOM_UV = {"copied": 17, "dry": 14, "dry_point": 15, "held": 6, "tile_skipped": 5 / 512,
          "wet": {True: 44, False: 28}, "partial": 72}
# Continuity per cell, reads: subgrid kcs 1 + faces 8 + V 8 + zs 8 + zs0 4 + zsderv 4 + zsmax 4
# + zmin/zmax/volmax 12 + 2 levels 8 = 57; regular kcs 1 + faces 8 + zs 8 + zsmax 4 = 21.
# Writes: a cell stores a field only when its value changes, counted on the device per
# field (zs 8, V 8, zs0 4, zsderv 4, zsmax 4). Steady cells (counted) read their kcs byte
# and faces only: 9.
CONT_READ = {True: 57, False: 21}
CONT_WRITE = (8, 8, 4, 4, 4)
CONT_STEADY = 9


def traffic_of(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE passes scaled by a copy of known size), or None."""
    tf = os.path.join(ROOT, "profiles", "momentum_traffic.json")
    if not os.path.exists(tf):
        return None
    with open(tf) as f:
        return json.load(f).get("bytes_per_launch", {}).get(kernel)


def cont_bytes(reads_per_cell, n_cells, stores_per_step):
    steady = stores_per_step[5]
    return (n_cells - steady) * reads_per_cell + steady * CONT_STEADY + \
        sum(n * b for n, b in zip(stores_per_step[:5], CONT_WRITE))


def byte_model(sizes, counts_per_step, stores_per_step, point_kernel):
    sg = bool(sizes["subgrid"])
    np_ = sizes["np"]
    c0, c1, c2, c3, c4, c5 = counts_per_step
    invalid = 2 * np_ - sizes["npuv"]
    copied = max(0.0, c0 - invalid)
    dry = MOM_UV["dry_point"] if point_kernel else MOM_UV["dry"]
    # cells of skipped tiles (two points per cell) are not read either
    mom = (np_ - c5 / 2) * MOM_CELL[sg] + copied * MOM_UV["copied"] + c1 * dry + c2 * MOM_UV["wet"][sg] + \
        c3 * MOM_UV["partial"] + c4 * MOM_UV["held"] + c5 * MOM_UV["tile_skipped"]
    cont = cont_bytes(CONT_READ[sg], np_, stores_per_step)
    return mom, cont, dict(copied=copied, dry=c1, dry_zeros_held=c4, quiescent_tile=c5, wet=c2, partial=c3)


def host_info():
    """lscpu's model, sockets, cores, SMT and NUMA figures, the CPUs this process may run
    on, and the OpenMP placement used by the CPU baseline."""
    import subprocess
    info = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "NUMA node(s)": "numa_nodes", "CPU(s)": "cpus"}
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except Exception as e:     # reported, never hidden
        info["lscpu_error"] = repr(e)
    info["affinity_cpus"] = len(os.sched_getaffinity(0))
    for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES"):
        info[k] = os.environ.get(k)
    return info


def cpu_baseline(sc, model, inst, nsteps, host):
    """The reference's own compute_fluxes + compute_water_levels (oracle/_ref, OpenMP)
    timed with the reference's per-phase timers on a bounded sample of the same run,
    started from the GPU state after warm-up. Falls back to the C restatement."""
    # one OpenMP thread per core of this process's CPU share, kept on its core (set before
    # the reference library starts its OpenMP runtime)
    os.environ.setdefault("OMP_NUM_THREADS", str(len(os.sched_getaffinity(0))))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.OracleRun(model).configure(sc.params, model)
    for f in ("zs", "q", "uv", "kfuv", "zs0", "zsderv", "zsmax") + (("z_volume",) if model.sizes()["subgrid"] else ()):
        orc.state(f)[...] = inst.get(f)
    t, dt, nt = inst.time()
    _, st = inst.update_to(t)            # no step: the CFL minimum of the last momentum sweep
    orc.state("uvmean")[...] = inst.get("uvmean")
    orc.set("t", t)
    orc.set("nt", nt)
    orc.set("min_dt", st["min_dt"])
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    if O.ref_available():
        r = O.RefRun(orc)
        r.R.ref_reset_timers()
        for _ in range(nsteps):
            r.step()
        tf, tc = r.timers()
        secs = tf + tc
        kind = "reference"
    else:
        t0 = time.perf_counter()
        orc.update(1e30, nsteps)
        secs = time.perf_counter() - t0
        kind = "port"
        cores = 1
    np_ = model.sizes()["np"]
    return dict(value=np_ * nsteps / secs, unit="cell-updates/s", cores=cores, kind=kind, host=host,
                sample=f"{nsteps} consecutive steps of the same 2048x2048 subgrid run, started from the GPU "
                       f"state after the timed window (t = {t:.0f} s); timed by the reference's own "
                       f"tloopflux+tloopcont accumulators",
                seconds=secs)


def whole_period(model, tuning, t_end, twin=True):
    """The whole simulated period from t = 0 on a fresh device instance (the tuning of the
    timed window pinned), timed end to end: sfincs_update to t_end, graph-captured. With
    `twin`, the FP64 twin runs the same period on its own adaptive steps and the FP32
    result's water-level error against it is stated over the active cells."""
    import torch
    from sfincs_amd import _lib
    a = _lib.Instance(model)
    a.pin_tuning(tuning)
    b0 = a.balance()
    torch.cuda.synchronize()
    t_a = time.perf_counter()
    rc, st = a.update(float(t_end))
    torch.cuda.synchronize()
    secs = time.perf_counter() - t_a
    zs = a.get("zs")
    assert np.all(np.isfinite(zs)), "non-finite water levels"
    b1 = a.balance()
    np_ = model.sizes()["np"]
    kcs = model.table("kcs")
    water = float(np.sum(a.get("z_volume")[kcs == 1])) if model.sizes()["subgrid"] else None
    dv = b1["volume"] - b0["volume"] + (b0["storage"] - b1["storage"])
    inflow = {k: b1[k] - b0[k] for k in ("boundary", "sources", "rain", "qext")}
    out = {"model_time_s": st["t"], "steps": st["nt"], "seconds": secs,
           "cell_updates_per_s": np_ * st["nt"] / secs, "ms_per_step": secs / max(1, st["nt"]) * 1e3, "rc": rc,
           "volume_balance": {"interior_volume_start_m3": b0["volume"], "interior_volume_end_m3": b1["volume"],
                              "inflow_m3": inflow, "abs_error_m3": abs(dv - sum(inflow.values())),
                              "rel_error": abs(dv - sum(inflow.values())) / water if water else None,
                              "note": "FP64 accumulators on the device: |dV - (boundary + sources + rain + qext)| "
                                      "over the interior cells, relative to their water volume at the end"}}
    a.close()
    if twin:
        b = _lib.Instance(model, fp64=True)
        t_b = time.perf_counter()
        rc64, st64 = b.update(float(t_end))
        torch.cuda.synchronize()
        d = np.abs(zs - b.get("zs"))
        out["fp64_twin"] = {
            "steps": st64["nt"], "seconds": time.perf_counter() - t_b, "max_abs_dzs_m": float(d.max()),
            "mean_abs_dzs_m": float(d.mean()), "p99_abs_dzs_m": float(np.quantile(d, 0.99)),
            "cells_over_1e-4_m": int(np.count_nonzero(d > 1e-4)),
            "note": "FP32 arithmetic (zs, V f64) against the same kernels in FP64, each on its own adaptive dt"}
        b.close()
    return out


# Indexed (quadtree / masked) regular meshes, per uv-point by branch: as the dense regular
# figures plus the two cell indices (8 B), the flags byte and the kfuv byte read up front
# (dry with zeros held: kcuv 1, indices 8, threshold 4, kfuv 1 = 14); wet points also read
# the eight neighbour indices (32 B). Per cell: zs 8 + zb 4. Continuity per cell, reads: kcs 1,
# four face indices 16, iref 1, zs 8, zsmax 4, four faces 16; writes zs 8 and zsmax 4 when
# they change (counted).
IDX_MOM_CELL = 12
IDX_MOM_UV = {"copied": 17, "dry": 24, "held": 14, "wet": 69, "partial": 69}
IDX_CONT_READ = 46


def quadtree(args):
    """Configuration 2: quadtree mesh of about 7.7M active cells (4 levels, 100 m .. 12.5 m,
    refined where the seeded fBm bed lies within -5..+5 m), Gaussian surge, 20 m/s SW wind.
    value = cell-updates/s of the whole step with state resident in HBM."""
    import torch
    from sfincs_amd import _lib
    from sfincs_amd import workloads as W
    t0 = time.perf_counter()
    sc = W.config2()
    sc.params["t1"] = 1.0e7
    model = sc.build_model()
    sizes = model.sizes()
    print(f"# generated quadtree: np {sizes['np']} npuv {sizes['npuv']} ncuv {sizes['ncuv']} in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    inst = _lib.Instance(model)
    tuning = inst.autotune(11)
    print(f"# tuning {tuning}", file=sys.stderr, flush=True)
    batch = _lib.Batch()
    batch.add(inst, 1.0e9)
    batch.advance(args.warmup)
    torch.cuda.synchronize()
    inst.set_counting(True)
    inst.profile(args.profile_steps)
    counts = [c / args.profile_steps for c in inst.counters()]
    stores = [c / args.profile_steps for c in inst.store_counters()]
    inst.set_counting(False)
    ms = inst.profile(args.profile_steps)
    batch.close()
    batch = _lib.Batch()
    batch.add(inst, 1.0e9)
    batch.advance(8)
    copied, dry, wet, part, held, _ = counts
    mom_b = sizes["np"] * IDX_MOM_CELL + copied * IDX_MOM_UV["copied"] + dry * IDX_MOM_UV["dry"] + \
        held * IDX_MOM_UV["held"] + (wet + part) * IDX_MOM_UV["wet"] + sizes["ncuv"] * 36
    cont_b = cont_bytes(IDX_CONT_READ, sizes["np"], stores)
    torch.cuda.synchronize()
    t_a = time.perf_counter()
    batch.advance(args.steps)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t_a
    t_model, dt_model, _ = inst.time()
    assert np.all(np.isfinite(inst.get("zs"))), "non-finite water levels"
    mom_gbs = mom_b / (ms[1] * 1e-3) / 1e9
    out = {
        "metric": "cell-updates/s (mom+cont), quadtree configuration", "value": sizes["np"] * args.steps / secs,
        "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": secs / args.steps * 1e3, "higher_is_better": True, "dtype": "f32 (zs f64)",
        "data": "synthetic (seeded fBm DEM, PCG64(1002))",
        "config": {"workload": "config 2: quadtree 4 levels 100..12.5 m, surge 3 m, wind 20 m/s SW",
                   "np": sizes["np"], "npuv": sizes["npuv"], "ncuv": sizes["ncuv"], "model_time_s": t_model,
                   "dt_s": dt_model, "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": mom_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": mom_gbs / PEAK_HBM_GBS, "traffic": traffic_of("k_mom_idx<float, false>"),
                     "kernel": "k_mom_idx<float, regular> + k_cuv",
                     "bytes_per_launch": mom_b, "launch_ms": ms[1]},
        "continuity": {"achieved_GBs": cont_b / (ms[2] * 1e-3) / 1e9, "bytes_per_launch": cont_b, "launch_ms": ms[2]},
        "forcing_launch_ms": ms[0],
        "branches_per_step": dict(copied=copied, dry=dry, dry_zeros_held=held, wet=wet, partial=part),
        "continuity_stores_per_step": dict(zs=stores[0], zsmax=stores[4], steady_skipped=stores[5]),
    }
    print(json.dumps(out))


def ensemble(args):
    """Configuration 3: 64 members of log-uniform size (256^2..2048^2, regular or subgrid),
    each with its own surge amplitude, rain and adaptive dt, 3 h simulated. The same
    members run one by one (sfx_update per member) and together in one batch; the JSON
    line reports both aggregate rates, the speed-up, time-to-result percentiles and the
    largest zs difference of any member between the two runs."""
    import torch
    from sfincs_amd import ensemble as E
    from sfincs_amd import workloads as W
    t0 = time.perf_counter()
    scs = W.ensemble_members(args.members, t_end=args.t_end, scale=args.scale)
    models = []
    for j, sc in enumerate(scs):
        models.append(sc.build_model())
        print(f"# member {j}: {sc.name}", file=sys.stderr, flush=True)
    t_ends = [sc.t_end for sc in scs]
    cells = [m.sizes()["np"] for m in models]
    print(f"# generated {len(models)} members, {sum(cells)} cells in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    print("# running the members one by one", file=sys.stderr, flush=True)
    seq = E.run_sequential(models, t_ends, keep_state=True)
    print(f"# sequential {seq['wall_s']:.2f}s {seq['cell_updates_per_s']:.3e}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    print("# running the members in one batch", file=sys.stderr, flush=True)
    bat = E.run_batched(models, t_ends, round_steps=args.round_steps, keep_state=True)
    print(f"# batched {bat['wall_s']:.2f}s {bat['cell_updates_per_s']:.3e}", file=sys.stderr, flush=True)
    dzs = max(float(np.max(np.abs(a - b))) for a, b in zip(seq["zs"], bat["zs"]))
    same_steps = seq["steps"] == bat["steps"]
    # each member's rate alone (its cells x steps over its own wall time in the sequential
    # run): the batch can at best run every member's work at the rate of the fastest one
    done = [0.0] + list(seq["time_to_result_s"])
    work = [c * n for c, n in zip(cells, seq["steps"])]
    rates = [w / max(done[j + 1] - done[j], 1e-9) for j, w in enumerate(work)]
    best = max(rates)
    alone = {"min": min(rates), "median": float(np.median(rates)), "max": best,
             "largest_member_rate": rates[int(np.argmax(cells))]}
    out = {
        "metric": "ensemble aggregate cell-updates/s (batched) and p50/p99 time-to-result",
        "value": bat["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32 (zs, z_volume f64)", "data": "synthetic (seeded, PCG64((1000+3)*100000+j))",
        "config": {"workload": f"config 3: {len(models)} members, log-uniform 256..2048 per side x {args.scale}, "
                               f"regular/subgrid mix, {args.t_end / 3600:.2f} h simulated each",
                   "cells_total": sum(cells), "round_steps": args.round_steps},
        "batched": {k: bat[k] for k in ("wall_s", "cell_updates", "cell_updates_per_s", "p50_s", "p99_s")},
        "sequential": {k: seq[k] for k in ("wall_s", "cell_updates", "cell_updates_per_s", "p50_s", "p99_s")},
        "speedup": bat["cell_updates_per_s"] / seq["cell_updates_per_s"],
        "sequential_member_rate": alone,
        "speedup_if_every_member_ran_at_the_best_alone_rate": seq["wall_s"] * best / sum(work),
        "max_abs_dzs_batched_vs_alone": dzs, "same_step_counts": same_steps,
    }
    print(json.dumps(out))


def serve(args):
    """Configuration 4 (scaled by the flags): Poisson arrivals of requests drawn from the
    configuration-3 distribution plus 10% quadtree meshes, served by continuous batching
    with admission by HBM footprint. Reports p50/p99 time-to-result and aggregate rate."""
    import torch
    from sfincs_amd import ensemble as E
    from sfincs_amd import workloads as W
    class Request:
        """A request as it arrives: a prepared model (the counterpart of an input deck
        written by the user's preprocessing) and its end time."""

        def __init__(self, model, t_end):
            self.model, self.t_end = model, t_end

        def build_model(self):
            return self.model

    r = W.rng(1004)
    t = 0.0
    reqs = []
    j = 0
    while True:
        t += float(r.exponential(1.0 / args.rate))
        if t > args.window:
            break
        dur = float(r.uniform(3600.0, 43200.0)) * args.duration_scale
        seed = 1004 * 100000 + j
        if r.uniform() < 0.1:
            nb = int(r.integers(40, 120) * args.scale) + 4
            make = (lambda nb=nb, seed=seed, dur=dur:
                    W.quadtree_case(nb, nb, nlev=4, dx0=100.0, seed=seed, t_end=dur, wind=False))
        else:
            make = (lambda seed=seed, dur=dur, j=j: W.member(seed, t_end=dur, scale=args.scale, name=f"request{j}"))
        reqs.append((t, make, dur))
        j += 1
    # the requests' inputs are prepared before the stream starts (preprocessing is not
    # part of serving); models are built in a thread pool
    import concurrent.futures as cf
    built = [None] * len(reqs)
    with cf.ThreadPoolExecutor(max_workers=args.workers) as pool:
        futs = {pool.submit(lambda mk=mk: mk().build_model()): i for i, (_, mk, _) in enumerate(reqs)}
        for n_done, f in enumerate(cf.as_completed(futs)):
            built[futs[f]] = f.result()
            if n_done % 10 == 0:
                print(f"# prepared {n_done + 1}/{len(reqs)} request models", file=sys.stderr, flush=True)
    reqs = [(a, Request(m, dur)) for (a, _, dur), m in zip(reqs, built)]
    total, free = torch.cuda.mem_get_info()
    budget = int(min(args.hbm_fraction * free, free - (8 << 30)))
    srv = E.Server(budget, round_steps=args.round_steps, workers=args.workers)
    for a, sc in reqs:
        srv.submit(a, sc)
    print(f"# serving {len(reqs)} requests", file=sys.stderr, flush=True)

    def progress(event, info):
        print(f"# {event} {info}", file=sys.stderr, flush=True)

    res = srv.run(progress=progress)
    out = {
        "metric": "served stream: p50/p99 time-to-result and aggregate cell-updates/s",
        "value": res["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32 (zs, z_volume f64)", "data": "synthetic (seeded, PCG64(1004))",
        "config": {"workload": f"config 4 scaled: Poisson {args.rate} req/s for {args.window} s, sizes x {args.scale}, "
                               f"durations U(1,12) h x {args.duration_scale}, 10% quadtree; request models prepared "
                               f"before the stream, time-to-result = queueing + HBM upload + stepping",
                   "requests": len(reqs), "hbm_budget_bytes": budget, "round_steps": args.round_steps},
        "p50_s": res["p50_s"], "p99_s": res["p99_s"], "wall_s": res["wall_s"],
        "cell_updates": res["cell_updates"], "max_concurrent": res["max_concurrent"],
        "max_footprint_bytes": res["max_footprint_bytes"],
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=1200,
                    help="steps of the reference's CPU path timed for cpu_baseline (about 10 s of work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--start-time", type=float, default=10800.0,
                    help="model time (s) the run is advanced to, untimed, before tuning, warm-up and the timed "
                         "window (default 3 h: the tide's crest, halfway through the 6 h period)")
    ap.add_argument("--no-whole-period", action="store_true",
                    help="skip the whole 6 h period (and its FP64 twin) run after the timed window")
    ap.add_argument("--no-twin", action="store_true", help="skip the FP64 twin of the whole period")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--mode", choices=("step", "quadtree", "ensemble", "serve"), default="step",
                    help="step: configuration 1 (the default line); quadtree: configuration 2; ensemble: "
                         "configuration 3; serve: configuration 4 (scaled by --window/--scale/--duration-scale)")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--t-end", type=float, default=10800.0)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--round-steps", type=int, default=32)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--window", type=float, default=600.0)
    ap.add_argument("--duration-scale", type=float, default=1.0)
    ap.add_argument("--hbm-fraction", type=float, default=0.9)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--tuning", default="",
                    help="pin this tuning (the JSON of a bench line's config.tuning) instead of autotuning, "
                         "so that a profiled run times the shapes of a reported one")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed region, run a 1 GiB elementwise copy (torch.mul by 1) three "
                         "times: a kernel of known HBM traffic for scaling PMC byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(ensembles shard by member across GPUs; each GPU runs its own batch)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(2)
    if args.mode == "quadtree":
        return quadtree(args)
    if args.mode == "ensemble":
        return ensemble(args)
    if args.mode == "serve":
        return serve(args)
    from sfincs_amd import _lib
    from sfincs_amd import workloads as W

    host = host_info()     # before any OpenMP runtime binds this thread to a core
    t0 = time.perf_counter()
    sc = W.config1(args.size)
    t_period = sc.t_end
    sc.params["t1"] = 1.0e7          # the timed window must not reach the end time
    model = sc.build_model()
    t_gen = time.perf_counter() - t0
    sizes = model.sizes()
    inst = _lib.Instance(model)
    # the window starts mid-run (untimed advance): tuning, warm-up and timing see the
    # state of the tide's crest, not the quiet first minutes
    if args.start_time > 0:
        inst.update(float(args.start_time))
    t_start = inst.time()[0]
    if args.tuning:
        inst.pin_tuning(json.loads(args.tuning))
        tuning = inst.tuning()
    else:
        tuning = inst.autotune(11)
    batch = _lib.Batch()
    batch.add(inst, 1.0e9)
    print(f"# generated {sizes['np']} cells in {t_gen:.1f}s; window starts at t={t_start:.0f}s; tuning {tuning}",
          file=sys.stderr, flush=True)

    batch.advance(args.warmup)
    torch.cuda.synchronize()
    # branch counters over a short run, then live kernel timing with HIP events on the
    # launch stream (counting off)
    inst.set_counting(True)
    inst.profile(args.profile_steps)
    counts = [c / args.profile_steps for c in inst.counters()]
    stores = [c / args.profile_steps for c in inst.store_counters()]
    inst.set_counting(False)
    ms = inst.profile(args.profile_steps)
    batch.close()
    batch = _lib.Batch()
    batch.add(inst, 1.0e9)
    batch.advance(8)
    mom_b, cont_b, branches = byte_model(sizes, counts, stores, tuning["momentum_kernel"] < 0)
    print(f"# per launch ms forcing/momentum/continuity = {ms}; branches {branches}", file=sys.stderr, flush=True)

    # timed region: exactly K steps (graph-captured), barrier + synchronize on both sides
    torch.cuda.synchronize()
    t_a = time.perf_counter()
    batch.advance(args.steps)
    torch.cuda.synchronize()
    t_b = time.perf_counter()
    secs = t_b - t_a
    t_model, dt_model, nt_model = inst.time()
    zs = inst.get("zs")
    assert np.all(np.isfinite(zs)), "non-finite water levels"

    if args.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(f"# calibration copy: 3 x {src.numel() * 4} bytes read and written", file=sys.stderr, flush=True)
        del src, dst

    value = sizes["np"] * args.steps / secs
    mom_gbs = mom_b / (ms[1] * 1e-3) / 1e9
    cont_gbs = cont_b / (ms[2] * 1e-3) / 1e9
    step_gbs = (mom_b + cont_b) / (secs / args.steps) / 1e9
    mk = tuning["momentum_kernel"]
    kernel = (f"k_mom_point<float, true, {-mk}>" if mk < 0 else
              "k_mom_cell<float, true>" if mk == 0 else "k_mom_strip<float, true>")
    traffic = traffic_of(kernel)
    out = {
        "metric": "cell-updates/s (mom+cont) & % HBM peak; ensemble p50/p99 time-to-result",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (zs, z_volume f64)",
        "data": "synthetic (seeded fBm DEM, PCG64(1001))",
        "config": {"workload": f"config 1: regular {args.size}x{args.size} subgrid (10 levels), dx=25 m, "
                               f"8 discharges, 1.5 m tide", "np": sizes["np"], "npuv": sizes["npuv"],
                   "tuning": tuning, "window_start_model_time_s": t_start,
                   "model_time_s": t_model, "dt_s": dt_model},
        "roofline": {"bound": "hbm", "achieved": mom_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": mom_gbs / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": kernel,
                     "bytes_per_launch": mom_b, "launch_ms": ms[1]},
        "continuity": {"achieved_GBs": cont_gbs, "frac": cont_gbs / PEAK_HBM_GBS, "bytes_per_launch": cont_b,
                       "launch_ms": ms[2]},
        "forcing_launch_ms": ms[0],
        "step_GBs": step_gbs,
        "branches_per_step": branches,
        "continuity_stores_per_step": dict(zip(("zs", "z_volume", "zs0", "zsderv", "zsmax", "steady_skipped"), stores)),
    }
    batch.close()
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(sc, model, inst, args.cpu_steps, host)
        except Exception as e:    # reported, never hidden
            out["cpu_baseline"] = {"error": repr(e)}
    if not args.no_whole_period:
        out["whole_period"] = whole_period(model, tuning, t_period, twin=not args.no_twin)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
