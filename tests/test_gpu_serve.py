"""Served ensemble on the GPU: the footprint used for admission is exactly
what a run allocates; stepping with loma_step_begin/poll is the same as
loma_advance; and each run's result is the same alone and in a mix of
other runs that arrive, overlap and leave while it advances."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("vor", "phi", "psi", "scal", "u00", "v00", "w00")


@pytest.mark.parametrize("grid,precision,flags", [
    ((64, 129, 64), 0, 0), ((64, 129, 64), 1, 2), ((192, 193, 96), 1, 2),
    ((96, 97, 48), 0, 2)])
def test_footprint_is_allocation(gpu, grid, precision, flags):
    from loma_phd_amd import capi
    print("footprint", grid, precision, flags, flush=True)
    y = np.linspace(-20.0, 20.0, grid[1])
    r = capi.Run(*grid, y, 160.0, 0.25, 0.25, precision=precision, flags=flags)
    assert r.bytes == capi.footprint(*grid, precision, flags)
    r.close()


def _run(cfg, precision):
    from loma_phd_amd import capi, ic
    st = ic.initial_condition(cfg)
    flags = capi.LOMA_FLAG_SCALAR_Z if cfg.scalar_z else 0
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, precision, flags)
    run.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"],
               z=st["z"] if cfg.scalar_z else None)
    return run


@pytest.mark.parametrize("graph", [False, True])
def test_step_poll_equals_advance(gpu, graph):
    from loma_phd_amd.grid import MixingLayerConfig
    print("step/poll graph", graph, flush=True)
    cfg = MixingLayerConfig(mgalx=64, my=129, mgalz=64, s=4.0, Re=400.0, seed=5)
    a, b = _run(cfg, 1), _run(cfg, 1)
    a.use_graph(graph)
    b.use_graph(graph)
    a.advance(4)
    for _ in range(4):
        b.step_begin()
        done = False
        while not done:
            done, ms, nonfinite = b.step_poll()
        assert ms > 0.0 and not nonfinite
    assert b.scalars()["istep"] == 4
    da, db = a.download(), b.download()
    for f in FIELDS + ("z",):
        assert np.array_equal(da[f], db[f]), f
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [1, 0])
def test_alone_equals_mix(gpu, precision):
    """Run 0 served alone and served among three other runs (different grids,
    density ratios, Reynolds numbers, arrival times and budgets) ends in the
    same state.  Tolerance: 1e-12 relative L2 per field (the runs share no
    data, so the states are expected to agree bit for bit)."""
    from loma_phd_amd.grid import MixingLayerConfig
    from loma_phd_amd.serve import HipBackend, Request, Server
    print("alone vs mix precision", precision, flush=True)

    def reqs():
        c = [MixingLayerConfig(mgalx=64, my=129, mgalz=64, s=4.0, Re=400.0, seed=11),
             MixingLayerConfig(mgalx=128, my=129, mgalz=64, s=7.0, Re=800.0, seed=12),
             MixingLayerConfig(mgalx=64, my=129, mgalz=64, s=1.5, Re=160.0, seed=13),
             MixingLayerConfig(mgalx=192, my=193, mgalz=96, s=3.0, Re=600.0, seed=14)]
        steps = [12, 5, 20, 8]
        arrival = [0.0, 0.0, 0.05, 0.1]
        return [Request(q, arrival[q], c[q], steps[q], precision) for q in range(4)]

    be = HipBackend(keep_state=True, reuse=False)
    alone = Server(be, mem_budget=64 << 30).run(reqs()[:1])
    assert alone["outcomes"][0].status == "done"
    # the mix reuses finished runs of a grid for later requests of that grid
    be = HipBackend(keep_state=True, reuse=True)
    mix = Server(be, mem_budget=64 << 30, max_slots=8).run(reqs())
    be.close()
    print("mix report", {k: v for k, v in mix.items() if k != "outcomes"}, flush=True)
    for q in range(4):
        assert mix["outcomes"][q].status == "done", (q, mix["outcomes"][q])
    assert mix["peak_active"] >= 2  # the runs did overlap
    a, m = alone["outcomes"][0].state, mix["outcomes"][0].state
    for f in FIELDS + ("z",):
        e = rel_l2(m[f], a[f])
        assert e <= 1e-12, (f, e)


def test_budget_holds_back_admission(gpu):
    """With room for one run only, the second waits until the first has left."""
    from loma_phd_amd import capi
    from loma_phd_amd.grid import MixingLayerConfig
    from loma_phd_amd.serve import HipBackend, Request, Server
    print("budget admission", flush=True)
    c = MixingLayerConfig(mgalx=64, my=129, mgalz=64, s=2.0, Re=160.0, seed=21)
    fp = capi.footprint(64, 129, 64, 1, capi.LOMA_FLAG_SCALAR_Z)
    reqs = [Request(0, 0.0, c, 6, 1), Request(1, 0.0, c, 6, 1)]
    be = HipBackend()
    rep = Server(be, mem_budget=fp + fp // 2).run(reqs)
    be.close()
    o = rep["outcomes"]
    assert o[0].status == o[1].status == "done"
    assert rep["peak_active"] == 1 and rep["peak_bytes"] == fp
    # both arrive at t = 0 and are prepared on worker threads; whichever is
    # prepared first is admitted first, and the other waits until it has left
    first, second = sorted((o[0], o[1]), key=lambda x: x.admit_s)
    assert second.admit_s >= first.end_s


def test_reused_run_equals_fresh(gpu):
    """A finished run re-targeted at a new request of its grid (other Re,
    density ratio and seed; loma_set_physics + upload) ends in the same state
    as a fresh run of that request; a budget for one run of either grid makes
    the pool free the kept run of the other grid."""
    from loma_phd_amd import capi
    from loma_phd_amd.grid import MixingLayerConfig
    from loma_phd_amd.serve import HipBackend, Request, Server
    print("reuse", flush=True)
    a = MixingLayerConfig(mgalx=64, my=129, mgalz=64, s=2.0, Re=160.0, seed=31)
    b = MixingLayerConfig(mgalx=64, my=129, mgalz=64, s=6.0, Re=700.0, seed=32)
    fresh = HipBackend(keep_state=True, reuse=False)
    ref = Server(fresh, mem_budget=64 << 30).run([Request(0, 0.0, b, 7, 1)])
    be = HipBackend(keep_state=True, reuse=True)
    # request 1 arrives 200 ms after request 0 (whose draws take about
    # 10 ms), so request 0 runs first and request 1 gets its re-targeted run
    rep = Server(be, mem_budget=64 << 30, max_slots=1).run(
        [Request(0, 0.0, a, 5, 1), Request(1, 0.2, b, 7, 1)])
    assert rep["outcomes"][1].admit_s >= rep["outcomes"][0].end_s
    assert be.idle_bytes() > 0
    be.close()
    x, y = rep["outcomes"][1].state, ref["outcomes"][0].state
    for f in ("vor", "phi", "psi", "scal", "u00", "v00", "w00", "z"):
        assert np.array_equal(x[f], y[f]), f
    # room for one run: the second grid evicts the first grid's kept run
    c = MixingLayerConfig(mgalx=96, my=129, mgalz=64, s=3.0, Re=300.0, seed=33)
    fa = capi.footprint(64, 129, 64, 1, capi.LOMA_FLAG_SCALAR_Z)
    fc = capi.footprint(96, 129, 64, 1, capi.LOMA_FLAG_SCALAR_Z)
    be = HipBackend()
    rep = Server(be, mem_budget=max(fa, fc) + min(fa, fc) // 2).run(
        [Request(0, 0.0, a, 3, 1), Request(1, 0.0, c, 3, 1)])
    assert all(o.status == "done" for o in rep["outcomes"].values())
    assert rep["peak_bytes"] <= max(fa, fc) + min(fa, fc) // 2
    be.close()
