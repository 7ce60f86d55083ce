"""Request-loop logic of loma_phd_amd.serve on the CPU, with a stand-in
backend whose steps finish after a fixed number of polls: admission by
footprint, arrival order, eviction of a non-finite run, rejection of a run
larger than the budget, the admission window, and the report."""
import pytest

from loma_phd_amd.grid import MixingLayerConfig
from loma_phd_amd.serve import Request, Server, poisson_requests


class FakeRun:
    def __init__(self, req):
        self.req = req
        self.steps = 0
        self.left = 0


class FakeBackend:
    def __init__(self, polls_per_step=3, nan_at=None):
        self.polls = polls_per_step
        self.nan_at = nan_at or {}  # rid -> step number that turns non-finite
        self.resident = 0
        self.peak = 0
        self.started = []
        self.freed = []

    def footprint(self, req):
        return req.cfg.points * 10

    def prepare(self, req):
        return ("ic", req.rid)

    def start(self, req, ic):
        assert ic == ("ic", req.rid)
        self.resident += self.footprint(req)
        self.peak = max(self.peak, self.resident)
        self.started.append(req.rid)
        return FakeRun(req)

    def begin(self, run):
        assert run.left == 0, "one step in flight per run"
        run.left = self.polls

    def poll(self, run):
        assert run.left > 0, "poll without a step in flight"
        run.left -= 1
        if run.left:
            return False, 0.0, False
        run.steps += 1
        return True, 0.5, self.nan_at.get(run.req.rid) == run.steps

    def finish(self, run):
        self.free(run)
        return {"steps": run.steps}

    def free(self, run):
        self.resident -= self.footprint(run.req)
        self.freed.append(run.req.rid)


class Clock:
    def __init__(self, dt=1e-3):
        self.t, self.dt = 0.0, dt

    def __call__(self):
        self.t += self.dt
        return self.t


def req(rid, t, g=(64, 129, 64), steps=5):
    return Request(rid, t, MixingLayerConfig(mgalx=g[0], my=g[1], mgalz=g[2]), steps)


def small_budget(n):
    return n * 64 * 129 * 64 * 10


def test_all_served_within_budget_in_order():
    be = FakeBackend()
    reqs = [req(i, 0.001 * i, steps=4 + i) for i in range(6)]
    rep = Server(be, mem_budget=small_budget(2), max_slots=8, workers=0,
                 clock=Clock(), idle_sleep=0).run(reqs)
    assert rep["counts"]["done"] == 6
    assert be.peak <= small_budget(2)
    assert rep["peak_bytes"] <= small_budget(2) and rep["peak_active"] == 2
    assert be.started == sorted(be.started)  # arrival order
    assert be.resident == 0 and sorted(be.freed) == list(range(6))
    for i in range(6):
        o = rep["outcomes"][i]
        assert o.steps_done == 4 + i and o.state == {"steps": 4 + i}
    assert rep["steps"] == sum(4 + i for i in range(6))
    assert rep["grid_point_steps"] == rep["steps"] * 64 * 129 * 64
    assert rep["p50_ms"] > 0 and rep["p99_ms"] >= rep["p50_ms"]


def test_slots_limit_concurrency():
    be = FakeBackend()
    reqs = [req(i, 0.0) for i in range(5)]
    rep = Server(be, mem_budget=small_budget(100), max_slots=3, workers=0,
                 clock=Clock(), idle_sleep=0).run(reqs)
    assert rep["peak_active"] == 3 and rep["counts"]["done"] == 5


def test_nonfinite_run_is_evicted_others_finish():
    be = FakeBackend(nan_at={1: 2})
    reqs = [req(i, 0.0, steps=6) for i in range(3)]
    rep = Server(be, mem_budget=small_budget(3), workers=0, clock=Clock(),
                 idle_sleep=0).run(reqs)
    o = rep["outcomes"]
    assert o[1].status == "nonfinite" and o[1].steps_done == 2 and o[1].state is None
    assert o[0].status == o[2].status == "done"
    assert o[0].steps_done == o[2].steps_done == 6
    assert be.resident == 0


def test_too_large_is_rejected_and_does_not_block():
    be = FakeBackend()
    reqs = [req(0, 0.0, g=(384, 385, 192)), req(1, 0.0)]
    rep = Server(be, mem_budget=small_budget(1), workers=0, clock=Clock(),
                 idle_sleep=0).run(reqs)
    assert rep["outcomes"][0].status == "rejected"
    assert rep["outcomes"][1].status == "done"
    assert be.started == [1]


def test_window_cuts_admission_and_resident_runs():
    be = FakeBackend(polls_per_step=2)
    reqs = [req(0, 0.0, steps=10**6), req(1, 10.0)]  # the second arrives after the window
    rep = Server(be, mem_budget=small_budget(4), workers=0, clock=Clock(),
                 idle_sleep=0).run(reqs, window=0.5)
    o = rep["outcomes"]
    assert o[0].status == "cut" and 0 < o[0].steps_done < 10**6
    assert o[1].status == "cut" and o[1].steps_done == 0
    assert be.resident == 0


def test_threaded_workers_same_outcome():
    be = FakeBackend()
    reqs = [req(i, 0.0005 * i, steps=3) for i in range(5)]
    rep = Server(be, mem_budget=small_budget(2), workers=2, idle_sleep=0).run(reqs)
    assert rep["counts"]["done"] == 5 and be.resident == 0


def test_poisson_requests_distribution():
    r = poisson_requests(seed=3, rate=2.0, horizon=200.0)
    assert 300 < len(r) < 500  # mean 400
    ts = [q.arrival for q in r]
    assert ts == sorted(ts) and ts[-1] < 200.0
    assert all(200 <= q.steps <= 2000 for q in r)
    assert all(1.0 <= q.cfg.s <= 8.0 and 160.0 <= q.cfg.Re <= 800.0 for q in r)
    grids = {(q.cfg.mgalx, q.cfg.my, q.cfg.mgalz) for q in r}
    assert (64, 129, 64) in grids and (384, 385, 192) in grids
    # same seed, same requests
    r2 = poisson_requests(seed=3, rate=2.0, horizon=200.0)
    assert [(q.arrival, q.steps, q.cfg.seed) for q in r2] == \
        [(q.arrival, q.steps, q.cfg.seed) for q in r]


def test_footprint_matches_grid_scaling():
    capi = pytest.importorskip("loma_phd_amd.capi")
    a = capi.footprint(64, 129, 64, capi.LOMA_FP64)
    b = capi.footprint(128, 129, 64, capi.LOMA_FP64)
    m = capi.footprint(64, 129, 64, capi.LOMA_MIXED)
    z = capi.footprint(64, 129, 64, capi.LOMA_FP64, capi.LOMA_FLAG_SCALAR_Z)
    assert 1.8 * a < b < 2.1 * a   # dominated by per-plane buffers
    assert m < a < z
    with pytest.raises(capi.LomaError):
        capi.footprint(64, 2000, 64, capi.LOMA_FP64)


def test_run_pool_reuse_and_eviction():
    from loma_phd_amd.serve import RunPool
    closed = []
    pool = RunPool(closed.append)
    pool.put("g1", "run-a", 100)
    pool.put("g2", "run-b", 200)
    assert pool.idle_bytes() == 300
    # a request of grid g1 takes the kept run whatever the room
    assert pool.make_room(7, "g1", 100, room=0)
    assert pool.take(7) == "run-a" and pool.idle_bytes() == 200
    # grid g3 needs 250 of 300: the kept g2 run is freed
    assert pool.make_room(8, "g3", 250, room=300)
    assert closed == ["run-b"] and pool.idle_bytes() == 0 and pool.take(8) is None
    # too large even with an empty pool
    pool.put("g2", "run-c", 200)
    assert not pool.make_room(9, "g4", 500, room=400)
    assert closed == ["run-b", "run-c"]
    pool.put("g5", "run-d", 10)
    pool.close_all()
    assert closed[-1] == "run-d" and pool.idle_bytes() == 0


def test_slow_preparation_does_not_block_later_requests():
    """A request whose initial condition is still being prepared is passed
    over; the prepared requests behind it are admitted in arrival order."""
    import threading
    gate = threading.Event()

    class SlowFirst(FakeBackend):
        def prepare(self, req):
            if req.rid == 0:
                gate.wait(timeout=30)
            return ("ic", req.rid)

        def finish(self, run):
            if run.req.rid == 2:
                gate.set()  # the first request's preparation ends late
            return super().finish(run)

    be = SlowFirst()
    reqs = [req(0, 0.0, steps=2), req(1, 0.0, steps=2), req(2, 0.0, steps=2)]
    rep = Server(be, mem_budget=small_budget(8), workers=4, idle_sleep=0).run(reqs)
    assert rep["counts"]["done"] == 3
    assert be.started[:2] == [1, 2] and be.started[2] == 0
