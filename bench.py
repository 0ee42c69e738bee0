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
#   wet (regular):         kcuv 1 + zbuvmx 4 + gn2uv 4 + kfuv 2 + mask 1 + q,uv 16 = 28
#   full (subgrid):        kcuv 1 + zmin 4 + s4 16 + fnfit 4 + kfuv 2 + mask 1 + q,uv 16 = 44
#   partial (subgrid):     kcuv 1 + zmin 4 + s4 16 + 2 levels 32 + kfuv 2 + mask 1 + q,uv 16 = 72
MOM_CELL = {True: 20, False: 12}
MOM_UV = {"copied": 17, "dry": 14, "wet": {True: 44, False: 28}, "partial": 72}
# Continuity per cell: subgrid kcs 1 + faces 8 + V 16 + zs 16 + zs0 8 + zsderv 4 + zsmax 8
# + zmin/zmax/volmax 12 + 2 levels 8 = 81; regular kcs 1 + faces 8 + zs 16 + zsmax 8 = 33.
CONT_CELL = {True: 81, False: 33}


def byte_model(sizes, counts_per_step):
    sg = bool(sizes["subgrid"])
    np_ = sizes["np"]
    c0, c1, c2, c3 = counts_per_step
    invalid = 2 * np_ - sizes["npuv"]
    copied = max(0.0, c0 - invalid)
    mom = np_ * MOM_CELL[sg] + copied * MOM_UV["copied"] + c1 * MOM_UV["dry"] + c2 * MOM_UV["wet"][sg] + \
        c3 * MOM_UV["partial"]
    cont = np_ * CONT_CELL[sg]
    return mom, cont, dict(copied=copied, dry=c1, wet=c2, partial=c3)


def cpu_baseline(sc, model, inst, nsteps):
    """The reference's own compute_fluxes + compute_water_levels (oracle/_ref, OpenMP)
    timed with the reference's per-phase timers on a bounded sample of the same run,
    started from the GPU state after warm-up. Falls back to the C restatement."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.OracleRun(model).configure(sc.params, model)
    for f in ("zs", "q", "uv", "kfuv", "zs0", "zsderv", "zsmax") + (("z_volume",) if model.sizes()["subgrid"] else ()):
        orc.state(f)[...] = inst.get(f)
    t, dt, nt = inst.time()
    orc.state("uvmean")[...] = inst.get("uvmean")
    orc.set("t", t)
    orc.set("nt", nt)
    orc.set("min_dt", dt / float(np.float32(sc.params.get("alpha", 0.5))))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
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
    return dict(value=np_ * nsteps / secs, unit="cell-updates/s", cores=cores, kind=kind,
                sample=f"{nsteps} consecutive steps of the same 2048x2048 subgrid run, started from the GPU "
                       f"state after warm-up; timed by the reference's own tloopflux+tloopcont accumulators",
                seconds=secs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(ensembles shard by member across GPUs; each GPU runs its own batch)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(2)
    from sfincs_amd import _lib
    from sfincs_amd import workloads as W

    t0 = time.perf_counter()
    sc = W.config1(args.size)
    sc.params["t1"] = 1.0e7          # the timed window must not reach the end time
    model = sc.build_model()
    t_gen = time.perf_counter() - t0
    sizes = model.sizes()
    inst = _lib.Instance(model)
    tuning = inst.autotune(3)
    batch = _lib.Batch()
    batch.add(inst, 1.0e9)
    print(f"# generated {sizes['np']} cells in {t_gen:.1f}s; tuning {tuning}", file=sys.stderr, flush=True)

    batch.advance(args.warmup)
    torch.cuda.synchronize()
    # live kernel timing with HIP events on the launch stream + branch counters
    inst.set_counting(True)
    ms = inst.profile(args.profile_steps)
    counts = [c / args.profile_steps for c in inst.counters()]
    inst.set_counting(False)
    batch.close()
    batch = _lib.Batch()
    batch.add(inst, 1.0e9)
    batch.advance(8)
    mom_b, cont_b, branches = byte_model(sizes, counts)
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

    value = sizes["np"] * args.steps / secs
    mom_gbs = mom_b / (ms[1] * 1e-3) / 1e9
    cont_gbs = cont_b / (ms[2] * 1e-3) / 1e9
    step_gbs = (mom_b + cont_b) / (secs / args.steps) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "momentum_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
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
                   "tile": tuning[:2], "model_time_s": t_model, "dt_s": dt_model},
        "roofline": {"bound": "hbm", "achieved": mom_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": mom_gbs / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": "k_momentum<float> (dense subgrid)",
                     "bytes_per_launch": mom_b, "launch_ms": ms[1]},
        "continuity": {"achieved_GBs": cont_gbs, "frac": cont_gbs / PEAK_HBM_GBS, "bytes_per_launch": cont_b,
                       "launch_ms": ms[2]},
        "forcing_launch_ms": ms[0],
        "step_GBs": step_gbs,
        "branches_per_step": branches,
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(sc, model, inst, args.cpu_steps)
        except Exception as e:    # reported, never hidden
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
