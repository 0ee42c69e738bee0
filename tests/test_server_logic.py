"""Host logic of the continuous-batching server, on the CPU with stand-in device objects:
admission in arrival order while the HBM footprint fits the budget, members that join a
running batch, retirement at the end time and release of their memory."""
import numpy as np

from sfincs_amd import ensemble as E


class FakeModel:
    def __init__(self, np_, steps, footprint):
        self._np, self.steps, self._fp = np_, steps, footprint
        self.closed = False

    def sizes(self):
        return {"np": self._np}

    def footprint(self):
        return self._fp

    def close(self):
        self.closed = True


class FakeInstance:
    live = 0
    peak = 0

    def __init__(self, model, device=0):
        self.model = model
        self.nt = 0
        FakeInstance.live += 1
        FakeInstance.peak = max(FakeInstance.peak, FakeInstance.live)

    def time(self):
        return float(self.nt), 1.0, self.nt

    def get(self, name):
        return np.full(self.model.sizes()["np"], float(self.nt))

    def close(self):
        FakeInstance.live -= 1


class FakeBatch:
    def __init__(self, device=0):
        self.members = []
        self.joined = []

    def add(self, inst, t_end):
        self.joined.append((inst, len(self.members)))
        self.members.append(inst)

    def remove(self, inst):
        self.members.remove(inst)

    def advance(self, n):
        for m in self.members:
            m.nt = min(m.model.steps, m.nt + n)
        return sum(1 for m in self.members if m.nt < m.model.steps)

    def done(self, inst):
        return inst.nt >= inst.model.steps

    def close(self):
        pass


class Req:
    def __init__(self, model, t_end=1.0):
        self.model, self.t_end = model, t_end

    def build_model(self):
        return self.model


def test_admission_by_footprint_and_retirement(monkeypatch):
    monkeypatch.setattr(E._lib, "Instance", FakeInstance)
    monkeypatch.setattr(E._lib, "Batch", FakeBatch)
    FakeInstance.live = FakeInstance.peak = 0
    srv = E.Server(hbm_budget=250, round_steps=10, workers=2, keep_state=True)
    models = [FakeModel(100 + k, steps=30 + 10 * k, footprint=100) for k in range(6)]
    for k, m in enumerate(models):
        srv.submit(0.0, Req(m))
    out = srv.run()
    assert out["requests"] == 6
    assert out["max_concurrent"] == 2          # 2 x 100 fits 250, a third does not
    assert out["max_footprint_bytes"] <= 250
    assert FakeInstance.peak == 2 and FakeInstance.live == 0
    assert all(m.closed for m in models)
    by_seq = {r["seq"]: r for r in out["results"]}
    for k, m in enumerate(models):
        assert by_seq[k]["steps"] == m.steps
        assert by_seq[k]["np"] == m.sizes()["np"]
        assert np.all(by_seq[k]["zs"] == m.steps)
    # admission follows arrival order
    adm = [by_seq[k]["admitted"] for k in range(6)]
    assert adm == sorted(adm)
    assert out["cell_updates"] == sum(m.sizes()["np"] * m.steps for m in models)


def test_oversized_request_runs_alone(monkeypatch):
    """A request larger than the budget is still served, alone, once the GPU is empty."""
    monkeypatch.setattr(E._lib, "Instance", FakeInstance)
    monkeypatch.setattr(E._lib, "Batch", FakeBatch)
    FakeInstance.live = FakeInstance.peak = 0
    srv = E.Server(hbm_budget=100, round_steps=5, workers=1)
    srv.submit(0.0, Req(FakeModel(10, steps=20, footprint=60)))
    srv.submit(0.0, Req(FakeModel(10, steps=20, footprint=500)))
    out = srv.run()
    assert out["requests"] == 2 and out["max_concurrent"] == 1


def test_shortest_work_first_within_the_fill_target(monkeypatch):
    """Requests that have arrived join shortest expected work first, and no more join
    while the cells on the GPU would pass the fill target (unless the GPU is empty)."""
    joined = []

    class LoggingBatch(FakeBatch):
        def add(self, inst, t_end):
            joined.append(inst.model.sizes()["np"])
            super().add(inst, t_end)

    monkeypatch.setattr(E._lib, "Instance", FakeInstance)
    monkeypatch.setattr(E._lib, "Batch", LoggingBatch)
    FakeInstance.live = FakeInstance.peak = 0
    srv = E.Server(hbm_budget=10 ** 9, round_steps=10, workers=1, keep_state=True, fill_cells=250)
    srv.submit(0.0, Req(FakeModel(10, steps=200, footprint=10)))     # on the GPU first, long
    sizes = [200, 50, 120, 60, 300]
    for n in sizes:
        srv.submit(0.0, Req(FakeModel(n, steps=40, footprint=10)))
    out = srv.run()
    assert out["requests"] == 6
    assert joined[0] == 10
    assert joined[1:4] == [50, 60, 120]          # 10 + 50 + 60 + 120 <= 250; adding 200 would pass it
    assert joined[4:] == [200, 300]
    assert FakeInstance.peak <= 4                # never more cells than the target (besides a lone request)


def test_run_batched_joins_small_members_first(monkeypatch):
    monkeypatch.setattr(E._lib, "Instance", FakeInstance)
    monkeypatch.setattr(E._lib, "Batch", FakeBatch)
    FakeInstance.live = FakeInstance.peak = 0
    sizes = [400, 100, 300, 50]
    models = [FakeModel(n, steps=30, footprint=1) for n in sizes]
    out = E.run_batched(models, [1.0] * 4, round_steps=10, tune_min_cells=None, keep_state=True, fill_cells=450)
    assert out["steps"] == [30] * 4
    done = out["time_to_result_s"]
    assert done[3] <= done[1] <= done[2] <= done[0]           # the small members finish first
    assert FakeInstance.peak == 4                              # instances exist for every member
