"""Request loop for many device-resident mixing-layer runs (continuous batching).

Run requests arrive over time.  Each names a grid, a density ratio, a
Reynolds number, a seed and a step budget.  The loop

  * prepares each request's initial condition as it arrives, on a worker
    thread so the loop keeps polling meanwhile (HipBackend draws the seeded
    coefficients there and builds the state on the device, loma_seed);
  * admits prepared requests in arrival order when a slot is free and the
    run's device footprint (loma_footprint: the exact bytes loma_init will
    allocate) fits in the memory budget beside the runs already resident; a
    request still being prepared does not hold back those behind it, a
    prepared one that does not fit does;
  * keeps every admitted run's spectral state on the GPU and keeps one RK3
    step in flight per run, each run on its own non-blocking stream
    (loma_step_begin / loma_step_poll), so the steps of all resident runs
    overlap on the device and a newly admitted run joins the mix at once;
  * retires a run when its budget is spent (final state downloaded, device
    memory freed), or evicts it when its state turns non-finite.

It reports per-step latency (host wall time from enqueue to observed
completion, p50/p99) beside aggregate grid-point-steps/s.

The reference has no request loop: its driver advances one run per MPI job
(trunk/lomacte3/src/main.f:215-236).  One served run is that driver's
cross1 time loop (cross.f:225-1226), one step per loma_step_begin.
"""
from __future__ import annotations

import concurrent.futures as cf
import dataclasses
import math
import threading
import time
from collections import deque
from typing import Any, Callable

import numpy as np

from .grid import MixingLayerConfig


@dataclasses.dataclass
class Request:
    """One run request: when it arrives (s after the loop starts), what to
    run and for how many RK3 steps."""
    rid: int
    arrival: float
    cfg: MixingLayerConfig
    steps: int
    precision: int = 1  # LOMA_MIXED

    @property
    def points(self) -> int:
        return self.cfg.points


@dataclasses.dataclass
class Outcome:
    rid: int
    status: str = "waiting"  # done | nonfinite | rejected | cut | failed
    steps_done: int = 0
    footprint: int = 0
    admit_s: float = math.nan
    end_s: float = math.nan
    error: str = ""
    state: Any = None


class RunPool:
    """Finished runs kept for the next request of the same grid.

    A kept run holds its device memory (grid tables, Helmholtz factors,
    buffers), so a later request of its grid skips set-up, allocation and
    release.  `make_room` is the admission test: it takes a kept run of the
    request's grid if there is one (its bytes move from the pool to the
    caller), or else frees kept runs of other grids, least recently kept
    first, until the request's footprint fits beside what stays kept.
    """

    def __init__(self, close: Callable[[Any], None]):
        self._close = close
        self._idle: list = []      # (key, run, bytes), oldest first
        self._taken: dict = {}     # rid -> run reserved for that request
        self._lock = threading.Lock()

    def idle_bytes(self) -> int:
        with self._lock:
            return sum(b for _, _, b in self._idle)

    def make_room(self, rid, key, nbytes: int, room: int) -> bool:
        with self._lock:
            for q, (k, run, b) in enumerate(self._idle):
                if k == key:
                    del self._idle[q]
                    self._taken[rid] = run
                    return True
            evict = []
            while self._idle and nbytes + sum(b for _, _, b in self._idle) > room:
                evict.append(self._idle.pop(0)[1])
            fits = nbytes + sum(b for _, _, b in self._idle) <= room
        for run in evict:
            self._close(run)
        return fits

    def take(self, rid):
        with self._lock:
            return self._taken.pop(rid, None)

    def put(self, key, run, nbytes: int):
        with self._lock:
            self._idle.append((key, run, nbytes))

    def close_all(self):
        with self._lock:
            runs = [r for _, r, _ in self._idle] + list(self._taken.values())
            self._idle, self._taken = [], {}
        for run in runs:
            self._close(run)


class HipBackend:
    """Runs requests on the GPU through the C ABI (capi.Run).  With `reuse`,
    finished runs are kept per grid (RunPool) and re-targeted at the next
    request of that grid (loma_set_physics + upload)."""

    def __init__(self, keep_state: bool = False, use_graph: bool = True, reuse: bool = True,
                 device_ic: bool = True, autotune: bool = True):
        from . import capi  # loads libloma_hip.so; fails loudly if it is missing
        self.capi = capi
        self.keep_state = keep_state
        self.use_graph = use_graph
        # autotune: every run is made with LOMA_FLAG_AUTOTUNE, so the first run
        # of each grid times its transform tiles and later runs of that grid
        # take the cached choice (the library keeps it per grid, precision
        # and Z for the life of the process)
        self.autotune = autotune
        # device_ic: only the seeded draws are made on the host and the state
        # is built on the device (loma_seed).  The host initial condition is
        # numpy work that holds the GIL for seconds per large grid and would
        # stall the loop thread.
        self.device_ic = device_ic
        self.pool = RunPool(lambda run: self._span("close", run.close)) if reuse else None
        self.spans: list = []  # (name, t_start, t_end) of every set-up / retirement call

    def _span(self, name, fn, *args):
        t = time.perf_counter()
        try:
            return fn(*args)
        finally:
            self.spans.append((name, t, time.perf_counter()))

    def key(self, req: Request):
        c = req.cfg
        return (c.mgalx, c.my, c.mgalz, req.precision, self.flags(req), c.Lx, c.Lz, c.Ly,
                c.mesh_ratio, c.ncfl)

    def idle_bytes(self) -> int:
        return self.pool.idle_bytes() if self.pool else 0

    def make_room(self, req: Request, nbytes: int, room: int) -> bool:
        if self.pool is None:
            return nbytes <= room
        return self.pool.make_room(req.rid, self.key(req), nbytes, room)

    def close(self):
        """Free the kept runs."""
        if self.pool:
            self.pool.close_all()

    def flags(self, req: Request) -> int:
        f = self.capi.LOMA_FLAG_SCALAR_Z if req.cfg.scalar_z else 0
        return f | (self.capi.LOMA_FLAG_AUTOTUNE if self.autotune else 0)

    def footprint(self, req: Request) -> int:
        c = req.cfg
        return self.capi.footprint(c.mgalx, c.my, c.mgalz, req.precision, self.flags(req))

    def prepare(self, req: Request):
        """What set-up needs of `req`'s initial condition: the seeded draws
        (device_ic) or the whole host-built state."""
        from . import ic
        if self.device_ic:
            return self._span("prepare", lambda: {"coef": ic.ic_coefficients(req.cfg),
                                                  "y": req.cfg.y()})
        return self._span("prepare", ic.initial_condition, req.cfg)

    def start(self, req: Request, ic):
        c = req.cfg
        run = self.pool.take(req.rid) if self.pool else None
        if run is not None:
            self._span("set_physics", run.set_physics, c.Re, c.pr, c.cfl, c.dt0)
        else:
            run = self._span("init", lambda: self.capi.Run(
                c.mgalx, c.my, c.mgalz, ic["y"], c.Re, c.alp, c.bet, pr=c.pr, cfl=c.cfl,
                ncfl=c.ncfl, dt0=c.dt0, precision=req.precision, flags=self.flags(req)))
        run.req_key = self.key(req)
        try:
            if "coef" in ic:
                from .ic import peak_wavenumber
                self._span("seed", run.seed, ic["coef"], c.s, c.amp, c.env_width,
                           peak_wavenumber(c))
            else:
                self._span("upload", lambda: run.upload(
                    ic["vor"], ic["phi"], ic["psi"], ic["scal"], ic["u00"], ic["v00"],
                    ic["w00"], z=ic["z"] if c.scalar_z else None))
            self._span("use_graph", run.use_graph, self.use_graph)
        except Exception:
            self._span("close", run.close)
            raise
        return run

    def begin(self, run):
        run.step_begin()

    def poll(self, run):
        return run.step_poll()

    def finish(self, run):
        """Final state (with keep_state); the run is closed here, or kept by
        `retired` once the loop has returned its bytes to the budget."""
        try:
            out = self._span("download", run.download) if self.keep_state else None
        except Exception:
            self._span("close", run.close)
            raise
        if self.pool is None:
            self._span("close", run.close)
        return out

    def retired(self, run):
        if self.pool is not None:
            self.pool.put(run.req_key, run, run.bytes)

    def free(self, run):
        self._span("close", run.close)


class _Inline:
    """Executor stand-in that runs jobs at submission (workers=0)."""

    def submit(self, fn, *args):
        f: cf.Future = cf.Future()
        try:
            f.set_result(fn(*args))
        except BaseException as e:  # noqa: BLE001 - handed to the caller
            f.set_exception(e)
        return f

    def shutdown(self, wait=True):
        pass


@dataclasses.dataclass
class _Live:
    req: Request
    out: Outcome
    run: Any = None
    t_begin: float = 0.0


class Server:
    """Continuous-batching request loop over one GPU.

    mem_budget: device bytes the resident runs may hold together.
    max_slots: resident runs at most (each holds one stream and one step in
    flight).
    workers: host threads for initial conditions, run set-up and retirement
    (0 runs them inline in the loop).
    """

    def __init__(self, backend, mem_budget: int, max_slots: int = 32, workers: int = 2,
                 clock: Callable[[], float] = time.perf_counter, idle_sleep: float = 1e-4):
        self.b = backend
        self.budget = int(mem_budget)
        self.max_slots = int(max_slots)
        self.workers = workers
        self.clock = clock
        self.idle_sleep = idle_sleep

    def _timed(self, name, fn, *args):
        t = time.perf_counter()
        out = fn(*args)
        t1 = time.perf_counter()
        dt = t1 - t
        if dt > self._slow.get(name, 0.0):
            self._slow[name] = dt
        if dt > 0.02:
            self._stalls.append((name, t, t1))
        return out

    def _stall_overlaps(self, stalls) -> dict:
        """For loop-thread intervals over 20 ms: how many of them overlapped
        each kind of backend call (initial condition, set-up, retirement)."""
        spans = getattr(self.b, "spans", None)
        if not spans:
            return {}
        out: dict = {}
        for _, a0, a1 in stalls:
            kinds = {n for n, b0, b1 in spans if b0 < a1 and b1 > a0}
            for k in kinds or {"none"}:
                out[k] = out.get(k, 0) + 1
        return out

    def _idle(self) -> int:
        f = getattr(self.b, "idle_bytes", None)
        return f() if f else 0

    def _room(self, req, nbytes: int, used: int) -> bool:
        """Admission test: the backend's own (it may hold finished runs for
        reuse and free them to make room), else plain budget arithmetic."""
        mk = getattr(self.b, "make_room", None)
        if mk is None:
            return used + nbytes <= self.budget
        return mk(req, nbytes, self.budget - used)

    def run(self, requests, window: float | None = None) -> dict:
        """Serve `requests` (arrival times relative to the start).  With
        `window`, admission stops at that time and runs still resident are cut
        after their step in flight; otherwise every request is served to its
        end.  Returns the report dict."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        pool = (cf.ThreadPoolExecutor(max_workers=self.workers) if self.workers > 0
                else _Inline())
        outs = {r.rid: Outcome(r.rid) for r in reqs}
        lat, dev_ms, lat_grid = [], [], []
        pending = deque(reqs)          # not yet arrived
        waiting: deque = deque()       # arrived, in arrival order: (req, prep future)
        starting: list = []            # admitted, set-up in flight: (_Live, future)
        active: list[_Live] = []       # resident, one step in flight
        retiring: list = []            # (_Live, future) download / free in flight
        used = 0
        peak_used = 0
        peak_active = 0
        gp_steps = 0
        t0 = self.clock()
        closed = False
        gaps = []          # loop passes that took longer than 20 ms (s)
        self._slow = {}    # slowest main-thread call per kind (s)
        self._stalls = []  # main-thread calls over 20 ms: (kind, start, end)
        self._gap_spans = []  # loop passes over 20 ms: ("pass", start, end)
        t_prev = t0
        try:
            while pending or waiting or starting or active or retiring:
                t_now = self.clock()
                if t_now - t_prev > 0.02:
                    gaps.append(t_now - t_prev)
                    self._gap_spans.append(("pass", t_prev, t_now))
                t_prev = t_now
                now = t_now - t0
                busy = False
                if window is not None and now >= window and not closed:
                    closed = True
                    for r in pending:
                        outs[r.rid].status = "cut"
                    pending.clear()
                    for r, f in waiting:
                        f.cancel()
                        outs[r.rid].status = "cut"
                    waiting.clear()
                # arrivals: start preparing the initial condition
                while pending and pending[0].arrival <= now:
                    r = pending.popleft()
                    outs[r.rid].footprint = self.b.footprint(r)
                    waiting.append((r, pool.submit(self.b.prepare, r)))
                    busy = True
                # admission in arrival order among the prepared requests: one
                # still being prepared is passed over, a prepared one that does
                # not fit stops admission (so large runs are not starved)
                qi = 0
                while qi < len(waiting):
                    r, f = waiting[qi]
                    if not f.done():
                        qi += 1
                        continue
                    o = outs[r.rid]
                    if o.footprint > self.budget or f.exception() is not None:
                        del waiting[qi]
                        o.status = "rejected" if o.footprint > self.budget else "failed"
                        o.error = "" if f.exception() is None else repr(f.exception())
                        o.end_s = now
                        busy = True
                        continue
                    nres = len(active) + len(starting)
                    if nres >= self.max_slots or not self._timed("admit", self._room, r,
                                                                  o.footprint, used):
                        break
                    del waiting[qi]
                    used += o.footprint
                    peak_used = max(peak_used, used + self._idle())
                    o.admit_s = now
                    lv = _Live(r, o)
                    starting.append((lv, pool.submit(self.b.start, r, f.result())))
                    busy = True
                # set-up finished: first step
                for item in list(starting):
                    lv, f = item
                    if not f.done():
                        continue
                    starting.remove(item)
                    busy = True
                    if f.exception() is not None:
                        lv.out.status, lv.out.error = "failed", repr(f.exception())
                        lv.out.end_s = self.clock() - t0
                        used -= lv.out.footprint
                        continue
                    lv.run = f.result()
                    if closed or lv.req.steps <= 0:
                        lv.out.status = "cut" if closed else "done"
                        retiring.append((lv, pool.submit(self.b.finish, lv.run)))
                        continue
                    self._timed("begin", self.b.begin, lv.run)
                    lv.t_begin = self.clock()
                    active.append(lv)
                peak_active = max(peak_active, len(active))
                # poll the steps in flight
                for lv in list(active):
                    done, ms, nonfinite = self._timed("poll", self.b.poll, lv.run)
                    if not done:
                        continue
                    t = self.clock()
                    busy = True
                    lat.append(t - lv.t_begin)
                    dev_ms.append(ms)
                    c = lv.req.cfg
                    lat_grid.append(f"{c.mgalx}x{c.my}x{c.mgalz}")
                    lv.out.steps_done += 1
                    gp_steps += lv.req.points
                    if nonfinite:
                        lv.out.status = "nonfinite"
                        active.remove(lv)
                        retiring.append((lv, pool.submit(self.b.free, lv.run)))
                    elif lv.out.steps_done >= lv.req.steps or closed:
                        lv.out.status = "done" if lv.out.steps_done >= lv.req.steps else "cut"
                        active.remove(lv)
                        retiring.append((lv, pool.submit(self.b.finish, lv.run)))
                    else:
                        self._timed("begin", self.b.begin, lv.run)
                        lv.t_begin = self.clock()
                # retirement finished: memory returns to the budget
                for item in list(retiring):
                    lv, f = item
                    if not f.done():
                        continue
                    retiring.remove(item)
                    busy = True
                    used -= lv.out.footprint
                    lv.out.end_s = self.clock() - t0
                    if f.exception() is not None:
                        lv.out.status, lv.out.error = "failed", repr(f.exception())
                    elif lv.out.status in ("done", "cut"):
                        lv.out.state = f.result()
                        keep = getattr(self.b, "retired", None)
                        if keep:  # the backend may keep the run for reuse
                            keep(lv.run)
                if not busy:
                    time.sleep(self.idle_sleep)
        finally:
            # an exception in the loop: leave no run resident
            for lv in active:
                try:
                    self.b.free(lv.run)
                except Exception:  # noqa: BLE001
                    pass
            pool.shutdown(wait=True)
        wall = self.clock() - t0
        lat_ms = np.asarray(lat) * 1e3
        by_grid = {}
        for g in sorted(set(lat_grid), key=lambda x: [int(v) for v in x.split("x")]):
            sel = np.asarray([q == g for q in lat_grid])
            by_grid[g] = dict(steps=int(sel.sum()),
                              p50_ms=float(np.percentile(lat_ms[sel], 50)),
                              p99_ms=float(np.percentile(lat_ms[sel], 99)))
        return dict(
            outcomes=outs,
            wall_s=wall,
            steps=int(lat_ms.size),
            grid_point_steps=int(gp_steps),
            gp_steps_per_s=gp_steps / wall if wall > 0 else 0.0,
            p50_ms=float(np.percentile(lat_ms, 50)) if lat_ms.size else math.nan,
            p99_ms=float(np.percentile(lat_ms, 99)) if lat_ms.size else math.nan,
            device_p50_ms=float(np.percentile(dev_ms, 50)) if dev_ms else math.nan,
            by_grid=by_grid,
            slow_loop_passes=len(gaps),
            slowest_loop_pass_ms=float(max(gaps) * 1e3) if gaps else 0.0,
            slowest_call_ms={k: float(v * 1e3) for k, v in self._slow.items()},
            slow_calls=len(self._stalls),
            slow_call_overlaps=self._stall_overlaps(self._stalls),
            slow_pass_overlaps=self._stall_overlaps(self._gap_spans),
            peak_bytes=int(peak_used),
            peak_active=int(peak_active),
            counts={s: sum(1 for o in outs.values() if o.status == s)
                    for s in ("done", "cut", "nonfinite", "rejected", "failed")},
        )


# ---- the served-ensemble workload ----------------------------------------
SERVED_GRIDS = [(64, 129, 64), (128, 129, 64), (192, 193, 96), (256, 257, 128),
                (384, 385, 192)]


def poisson_requests(seed: int, rate: float, horizon: float, grids=SERVED_GRIDS,
                     ratio=(1.0, 8.0), re=(160.0, 800.0), steps=(200, 2000),
                     precision: int = 1, scalar_z: bool = True):
    """Requests arriving as a Poisson process of `rate` per second over
    `horizon` seconds; grid, density ratio U[ratio], Re U[re] and step budget
    U[steps] drawn per request from one seeded generator."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= horizon:
            break
        g = grids[int(rng.integers(len(grids)))]
        cfg = MixingLayerConfig(mgalx=g[0], my=g[1], mgalz=g[2],
                                s=float(rng.uniform(*ratio)), Re=float(rng.uniform(*re)),
                                seed=int(rng.integers(2**31)), scalar_z=scalar_z)
        out.append(Request(rid, t, cfg, int(rng.integers(steps[0], steps[1] + 1)),
                           precision))
        rid += 1
    return out
