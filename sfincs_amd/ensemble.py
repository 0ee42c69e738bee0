"""Ensembles and serving on one GPU.

run_batched:   members of uneven size advance together in one batch (one launch per phase
               per step over an (instance, tile) work list, graph-captured); each member
               keeps its own adaptive dt and stops at its own end time.
run_sequential: the same members one after the other, each on its own (the reference's
               way of running ensembles, docs/output.rst:19-20), for comparison.
Server:        continuous batching -- requests are admitted as they arrive while their HBM
               footprint fits the budget, advance in the shared batch while on the GPU,
               and retire (freeing their memory) when they reach their end time.
"""
import concurrent.futures as cf
import heapq
import time

import numpy as np

from . import _lib


def _np(model):
    return model.sizes()["np"]


def run_batched(scenarios, round_steps=32, device=0):
    models = [sc.build_model() for sc in scenarios]
    insts = [_lib.Instance(m, device) for m in models]
    batch = _lib.Batch(device)
    for sc, inst in zip(scenarios, insts):
        batch.add(inst, sc.t_end)
    t0 = time.perf_counter()
    done_at = [None] * len(insts)
    active = len(insts)
    while active:
        active = batch.advance(round_steps)
        now = time.perf_counter() - t0
        for j, inst in enumerate(insts):
            if done_at[j] is None and batch.done(inst):
                done_at[j] = now
    wall = time.perf_counter() - t0
    cells = sum(_np(m) * inst.time()[2] for m, inst in zip(models, insts))
    out = dict(members=len(insts), wall_s=wall, cell_updates=cells, cell_updates_per_s=cells / wall,
               time_to_result_s=done_at,
               p50_s=float(np.percentile(done_at, 50)), p99_s=float(np.percentile(done_at, 99)))
    batch.close()
    for i in insts:
        i.close()
    return out


def run_sequential(scenarios, device=0):
    models = [sc.build_model() for sc in scenarios]
    t0 = time.perf_counter()
    cells = 0
    done_at = []
    for sc, m in zip(scenarios, models):
        inst = _lib.Instance(m, device)
        m_t1 = sc.t_end
        # run to the member's end time (its t1)
        rc, st = inst.update(m_t1 - inst.time()[0])
        cells += _np(m) * st["nt"]
        done_at.append(time.perf_counter() - t0)
        inst.close()
    wall = time.perf_counter() - t0
    return dict(members=len(models), wall_s=wall, cell_updates=cells, cell_updates_per_s=cells / wall,
                time_to_result_s=done_at,
                p50_s=float(np.percentile(done_at, 50)), p99_s=float(np.percentile(done_at, 99)))


class Server:
    """Continuous-batching scenario server on one GPU.

    submit() enqueues (arrival_time, scenario); run() admits requests as they arrive while
    their footprint fits `hbm_budget` bytes, advances the shared batch `round_steps` steps
    at a time, and retires finished members. Host-side model building overlaps the GPU
    work in a thread pool."""

    def __init__(self, hbm_budget, round_steps=32, device=0, workers=4):
        self.budget = hbm_budget
        self.round_steps = round_steps
        self.device = device
        self.pool = cf.ThreadPoolExecutor(max_workers=workers)
        self.queue = []          # (arrival, seq, scenario)
        self.seq = 0

    def submit(self, arrival, scenario):
        heapq.heappush(self.queue, (arrival, self.seq, scenario))
        self.seq += 1

    def run(self, clock=None):
        clock = clock or time.perf_counter
        t0 = clock()
        batch = _lib.Batch(self.device)
        pending = []             # (arrival, seq, future-model, scenario) in arrival order
        live = {}                # seq -> (inst, model, arrival, footprint, scenario)
        used = 0
        results = []
        max_live = 0
        cells = 0
        while self.queue or pending or live:
            now = clock() - t0
            while self.queue and self.queue[0][0] <= now:
                arrival, seq, sc = heapq.heappop(self.queue)
                pending.append((arrival, seq, self.pool.submit(sc.build_model), sc))
            admitted = False
            while pending:
                arrival, seq, fut, sc = pending[0]
                if not fut.done():
                    break
                model = fut.result()
                need = model.footprint()
                if live and used + need > self.budget:
                    break
                pending.pop(0)
                inst = _lib.Instance(model, self.device)
                batch.add(inst, sc.t_end)
                live[seq] = (inst, model, arrival, need, sc)
                used += need
                admitted = True
            max_live = max(max_live, len(live))
            if not live:
                # nothing on the GPU: wait for the next arrival or model
                if pending:
                    cf.wait([pending[0][2]], timeout=0.05)
                elif self.queue:
                    time.sleep(max(0.0, min(0.05, self.queue[0][0] - (clock() - t0))))
                continue
            batch.advance(self.round_steps)
            now = clock() - t0
            for seq in list(live):
                inst, model, arrival, need, sc = live[seq]
                if batch.done(inst):
                    nt = inst.time()[2]
                    zsmax = inst.get("zsmax")
                    cells += model.sizes()["np"] * nt
                    results.append(dict(seq=seq, arrival=arrival, done=now, ttr=now - arrival, steps=nt,
                                        np=model.sizes()["np"], zsmax_max=float(np.max(zsmax))))
                    batch.remove(inst)
                    inst.close()
                    model.close()
                    used -= need
                    del live[seq]
        wall = clock() - t0
        batch.close()
        ttr = [r["ttr"] for r in results]
        return dict(requests=len(results), wall_s=wall, cell_updates=cells, cell_updates_per_s=cells / wall,
                    p50_s=float(np.percentile(ttr, 50)) if ttr else None,
                    p99_s=float(np.percentile(ttr, 99)) if ttr else None,
                    max_concurrent=max_live, results=results)
