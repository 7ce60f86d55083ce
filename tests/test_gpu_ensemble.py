"""Ensembles: several independent runs of different grids advanced together
(one graph launch per step, batched launches: one per kernel and group of
runs that share an instantiation) give the same state as each run alone,
bit for bit; the tile autotuner changes the schedule but not the result.
The grids put three runs of different y sizes into one phase-kernel group
(M = 3) and two runs of one grid into one transform group."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("vor", "phi", "psi", "scal", "u00", "v00", "w00")
GRIDS = [(48, 97, 36, 2.0, 160.0, 1), (64, 129, 64, 4.0, 400.0, 2), (40, 65, 48, 8.0, 800.0, 3),
         (64, 129, 64, 2.0, 160.0, 4)]


def _make(grid, precision, flags=0):
    from loma_phd_amd import capi, ic
    from loma_phd_amd.grid import MixingLayerConfig
    mgalx, my, mgalz, s, Re, seed = grid
    cfg = MixingLayerConfig(mgalx=mgalx, my=my, mgalz=mgalz, s=s, Re=Re, seed=seed)
    st = ic.initial_condition(cfg)
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, precision, flags)
    run.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    return run


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_matches_alone(gpu, precision):
    from loma_phd_amd import capi
    print(f"ensemble precision={precision}", flush=True)
    alone = []
    for g in GRIDS:
        r = _make(g, precision)
        r.advance(3)
        alone.append(r.download())
        r.close()
    runs = [_make(g, precision) for g in GRIDS]
    for r in runs:
        r.use_graph(True)
    capi.advance_many(runs, 2)
    capi.advance_many(runs, 1)
    for q, r in enumerate(runs):
        got = r.download()
        for f in FIELDS:
            assert np.array_equal(got[f], alone[q][f]), (q, f, rel_l2(got[f], alone[q][f]))
        r.close()


def test_autotune_same_result(gpu):
    from loma_phd_amd import capi
    print("autotune", flush=True)
    g = GRIDS[1]
    a = _make(g, 0)
    b = _make(g, 0, capi.LOMA_FLAG_AUTOTUNE)
    print("tiles", a.tiles(), b.tiles(), flush=True)
    a.advance(2)
    b.advance(2)
    da, db = a.download(), b.download()
    for f in FIELDS:
        assert rel_l2(db[f], da[f]) < 1e-14, f


def test_ensemble_batch_shape_trials(gpu):
    """Runs made with LOMA_FLAG_AUTOTUNE: the ensemble's first four steps
    try two batch shapes (one launch per group, at most two runs per launch)
    and keep the faster; every step, trial or not, leaves each run where it
    would be alone, bit for bit."""
    from loma_phd_amd import capi
    print("ensemble batch-shape trials", flush=True)
    flags = capi.LOMA_FLAG_AUTOTUNE
    alone = []
    for g in GRIDS:
        r = _make(g, 1, flags)
        r.advance(6)
        alone.append(r.download())
        r.close()
    runs = [_make(g, 1, flags) for g in GRIDS]
    for r in runs:
        r.use_graph(True)
    for _ in range(6):
        capi.advance_many(runs, 1)
    for q, r in enumerate(runs):
        got = r.download()
        for f in FIELDS:
            assert np.array_equal(got[f], alone[q][f]), (q, f)
        r.close()
