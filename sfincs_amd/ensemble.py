"""Ensembles and serving on one GPU.

run_batched:   members of uneven size advance together in one batch (one launch per phase
               per step over an (instance, tile) work list, graph-captured); each member
               keeps its own adaptive dt and stops at its own end time. Members join the
               batch shortest expected work first, while the cells on the GPU stay below a
               fill target, so small members are not held to the rounds of the largest.
run_sequential: the same members one after the other, each on its own (the reference's
               way of running ensembles, docs/output.rst:19-20), for comparison.
Server:        continuous batching -- arrived requests are admitted shortest expected work
               first while their HBM footprint fits the budget and the cells on the GPU
               stay below the fill target, advance in the shared batch while on the GPU,
               and retire (freeing their memory) when they reach their end time.

Model building (table generation, mesh construction) happens on the host before a member
is timed; both runners time only the GPU time stepping of members already in HBM.
"""
import concurrent.futures as cf
import heapq
import time

import numpy as np

from . import _lib


# Cells on the GPU at which one batch round keeps every CU busy: a 2048 x 2048 grid alone
# runs the step at the device's full rate; smaller batches leave CUs idle, larger ones
# only lengthen every member's rounds.
FILL_CELLS = 4 << 20


def _np(model):
    return model.sizes()["np"]


def expected_work(model, t_end):
    """Relative work to the end time: cells x simulated time / cell size (the CFL step
    scales with the grid spacing)."""
    table = getattr(model, "table", None)
    dx = float(table("dxrm").min()) if table is not None else 1.0
    return _np(model) * float(t_end) / max(dx, 1e-3)


def _instances(models, device, tune_min_cells):
    insts = []
    for m in models:
        inst = _lib.Instance(m, device)
        if tune_min_cells is not None and _np(m) >= tune_min_cells:
            inst.autotune(5)
        insts.append(inst)
    return insts


def _summary(models, insts, wall, done_at, keep_state):
    cells = sum(_np(m) * inst.time()[2] for m, inst in zip(models, insts))
    out = dict(members=len(insts), wall_s=wall, cell_updates=cells, cell_updates_per_s=cells / wall,
               time_to_result_s=done_at,
               p50_s=float(np.percentile(done_at, 50)), p99_s=float(np.percentile(done_at, 99)),
               steps=[inst.time()[2] for inst in insts])
    if keep_state:
        out["zs"] = [inst.get("zs") for inst in insts]
    return out


def run_batched(models, t_ends, round_steps=32, device=0, tune_min_cells=1 << 20, keep_state=False,
                fill_cells=FILL_CELLS):
    """Every member through one batch until each reaches its end time. Members join
    shortest expected work first while the cells in the batch stay below `fill_cells`
    (at least one member is always in); a member leaves the batch when it is done."""
    insts = _instances(models, device, tune_min_cells)
    order = sorted(range(len(insts)), key=lambda j: expected_work(models[j], t_ends[j]))
    batch = _lib.Batch(device)
    done_at = [None] * len(insts)
    live = []
    cells = 0
    t0 = time.perf_counter()
    while order or live:
        while order and (not live or cells + _np(models[order[0]]) <= fill_cells):
            j = order.pop(0)
            batch.add(insts[j], t_ends[j])
            live.append(j)
            cells += _np(models[j])
        batch.advance(round_steps)
        now = time.perf_counter() - t0
        for j in list(live):
            if batch.done(insts[j]):
                done_at[j] = now
                batch.remove(insts[j])
                live.remove(j)
                cells -= _np(models[j])
    wall = time.perf_counter() - t0
    out = _summary(models, insts, wall, done_at, keep_state)
    batch.close()
    for i in insts:
        i.close()
    return out


def run_sequential(models, t_ends, device=0, tune_min_cells=1 << 20, keep_state=False):
    """The same members one after the other, each to its end time with sfx_update."""
    insts = _instances(models, device, tune_min_cells)
    done_at = []
    t0 = time.perf_counter()
    for te, inst in zip(t_ends, insts):
        inst.update(te - inst.time()[0])
        done_at.append(time.perf_counter() - t0)
    wall = time.perf_counter() - t0
    out = _summary(models, insts, wall, done_at, keep_state)
    for i in insts:
        i.close()
    return out


class Server:
    """Continuous-batching scenario server on one GPU.

    submit() enqueues (arrival_time, scenario); run() admits requests as they arrive while
    their footprint fits `hbm_budget` bytes, advances the shared batch `round_steps` steps
    at a time, and retires finished members. Host-side model building overlaps the GPU
    work in a thread pool."""

    def __init__(self, hbm_budget, round_steps=32, device=0, workers=4, keep_state=False, fill_cells=FILL_CELLS):
        self.budget = hbm_budget
        self.fill_cells = fill_cells
        self.round_steps = round_steps
        self.device = device
        self.pool = cf.ThreadPoolExecutor(max_workers=workers)
        self.queue = []          # (arrival, seq, scenario)
        self.seq = 0
        self.keep_state = keep_state

    def submit(self, arrival, scenario):
        heapq.heappush(self.queue, (arrival, self.seq, scenario))
        self.seq += 1

    def run(self, clock=None, progress=None):
        """Serves every submitted request. `progress(event, info)` is called on each
        admission and completion when given."""
        clock = clock or time.perf_counter
        t0 = clock()
        batch = _lib.Batch(self.device)
        pending = []             # (arrival, seq, future-model, scenario), arrived, not yet on the GPU
        live = {}                # seq -> (inst, model, arrival, footprint, scenario, admitted)
        used = 0
        on_gpu = 0               # cells of the requests on the GPU
        results = []
        max_live = 0
        max_used = 0
        cells = 0
        while self.queue or pending or live:
            now = clock() - t0
            while self.queue and self.queue[0][0] <= now:
                arrival, seq, sc = heapq.heappop(self.queue)
                pending.append((arrival, seq, self.pool.submit(sc.build_model), sc))
            # admission: among the arrived requests whose model is built, shortest expected
            # work first, while the memory fits and the GPU is not already full
            while pending:
                ready = [k for k, p in enumerate(pending) if p[2].done()]
                if not ready:
                    break
                k = min(ready, key=lambda r: expected_work(pending[r][2].result(), pending[r][3].t_end))
                arrival, seq, fut, sc = pending[k]
                model = fut.result()
                need = model.footprint()
                if live and (used + need > self.budget or on_gpu + _np(model) > self.fill_cells):
                    break
                pending.pop(k)
                inst = _lib.Instance(model, self.device)
                batch.add(inst, sc.t_end)
                live[seq] = (inst, model, arrival, need, sc, clock() - t0)
                used += need
                on_gpu += _np(model)
                if progress:
                    progress("admit", dict(seq=seq, live=len(live), used=used, pending=len(pending),
                                           queued=len(self.queue)))
            max_live = max(max_live, len(live))
            max_used = max(max_used, used)
            if not live:
                # nothing on the GPU: wait for the next model or arrival
                if pending:
                    cf.wait([p[2] for p in pending], timeout=0.05, return_when=cf.FIRST_COMPLETED)
                elif self.queue:
                    time.sleep(max(0.0, min(0.05, self.queue[0][0] - (clock() - t0))))
                continue
            batch.advance(self.round_steps)
            now = clock() - t0
            for seq in list(live):
                inst, model, arrival, need, sc, admitted = live[seq]
                if batch.done(inst):
                    nt = inst.time()[2]
                    cells += model.sizes()["np"] * nt
                    r = dict(seq=seq, arrival=arrival, admitted=admitted, done=now, ttr=now - arrival,
                             steps=nt, np=model.sizes()["np"])
                    if self.keep_state:
                        r["zs"] = inst.get("zs")
                    results.append(r)
                    if progress:
                        progress("done", dict(seq=seq, ttr=r["ttr"], steps=nt, done=len(results)))
                    batch.remove(inst)
                    inst.close()
                    on_gpu -= model.sizes()["np"]
                    model.close()
                    used -= need
                    del live[seq]
        wall = clock() - t0
        batch.close()
        ttr = [r["ttr"] for r in results]
        return dict(requests=len(results), wall_s=wall, cell_updates=cells, cell_updates_per_s=cells / wall,
                    p50_s=float(np.percentile(ttr, 50)) if ttr else None,
                    p99_s=float(np.percentile(ttr, 99)) if ttr else None,
                    max_concurrent=max_live, max_footprint_bytes=max_used, results=results)
