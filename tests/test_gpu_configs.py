"""The production workloads at their real sizes.

Single production run, 512x513x256 (density ratio 4, Re 800, seeded
perturbation 1e-3 with a Gaussian envelope of width 2):
  * one FP64 RK3 step on the GPU against the FP64 oracle on the host from
    the same state (seeded on the device): 1e-10 per field;
  * ten steps of the mixed form against ten FP64 steps on the GPU from the
    same state: every physical field (rho u, rho v, rho w, T) within 1e-5.

Ensemble batch on its own grids (128x129x64, 192x193x96, 256x257x128, mixed
precision, density ratios and Reynolds numbers drawn as the workload draws
them): each run's state after three steps in one ensemble graph equals the
same run advanced alone, bit for bit.  Then one run is seeded again with
another density ratio and the ensemble advanced again: the run matches the
same run advanced alone from that new state, so no cached ensemble graph
keeps the old boundary densities.
"""
import dataclasses

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("vor", "phi", "psi", "scal", "u00", "v00", "w00")


def _config_b():
    from loma_phd_amd.grid import CONFIG_B
    return dataclasses.replace(CONFIG_B, scalar_z=False)


def _seeded(cfg, precision, flags=0):
    from loma_phd_amd import capi, ic
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, cfg.y(), cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, precision, flags)
    ic.seed_run(run, cfg)
    return run


def test_config_b_fp64_step_vs_oracle(gpu, oracle_mod):
    from loma_phd_amd import capi
    O = oracle_mod
    cfg = _config_b()
    print("config B: one FP64 step vs the oracle", flush=True)
    run = _seeded(cfg, capi.LOMA_FP64)
    st = run.download()
    g = O.Grid(cfg.mgalx, cfg.my, cfg.mgalz, cfg.y(), cfg.Re, cfg.alp, cfg.bet, cfg.pr)
    S = O.State(g, dt0=cfg.dt0, cfl=cfg.cfl, ncfl=cfg.ncfl, flags=0)
    S.set(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    run.advance(1)
    a = run.download()
    sa = run.scalars()
    run.close()
    print("oracle step (host, about 20 s)", flush=True)
    assert S.step(1) == 0
    b = S.get()
    errs = {f: rel_l2(a[f], b[f]) for f in FIELDS}
    print("config B fp64 step vs oracle", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for f, e in errs.items():
        assert e < 1e-10, (f, e)
    sb = S.scalars()
    assert abs(sa["dt"] - sb["dt"]) <= 1e-12 * sb["dt"]


def test_config_b_mixed_vs_fp64_10_steps(gpu):
    from loma_phd_amd import capi, diagnostics, ic
    cfg = _config_b()
    print("config B: ten mixed steps vs ten FP64 steps", flush=True)
    r64 = _seeded(cfg, capi.LOMA_FP64)
    st = r64.download()
    rmx = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, cfg.y(), cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, capi.LOMA_MIXED, 0)
    rmx.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    for r in (r64, rmx):
        r.use_graph(True)
        r.advance(10)
    a, b = rmx.download(), r64.download()
    ta, tb = rmx.scalars()["time"], r64.scalars()["time"]
    rmx.close()
    r64.close()
    errs = diagnostics.field_errors(a, b, ic.HostOps(cfg.y()), cfg.alp, cfg.bet)
    print("config B mixed vs fp64 after 10 steps, physical fields",
          {k: f"{v:.2e}" for k, v in errs.items()}, "time", ta, tb, flush=True)
    for f, e in errs.items():
        assert e < 1e-5, (f, e)
    assert abs(ta - tb) <= 1e-5 * tb


def _d_configs():
    from loma_phd_amd import grid
    # the workload's draw, one run of each of its grids
    cfgs = grid.ensemble_configs(32)
    out = {}
    for c in cfgs:
        out.setdefault((c.mgalx, c.my, c.mgalz), dataclasses.replace(c, scalar_z=False))
    return [out[g] for g in sorted(out)]


def test_config_d_ensemble_matches_alone(gpu):
    from loma_phd_amd import capi
    cfgs = _d_configs()
    assert len(cfgs) == 3
    print("config D grids", [(c.mgalx, c.my, c.mgalz, round(c.s, 3), round(c.Re, 1))
                             for c in cfgs], flush=True)
    alone = []
    for c in cfgs:
        r = _seeded(c, capi.LOMA_MIXED)
        r.advance(3)
        alone.append(r.download())
        r.close()
    runs = [_seeded(c, capi.LOMA_MIXED) for c in cfgs]
    for r in runs:
        r.use_graph(True)
    capi.advance_many(runs, 2)
    capi.advance_many(runs, 1)
    for q, r in enumerate(runs):
        got = r.download()
        for f in FIELDS:
            assert np.array_equal(got[f], alone[q][f]), (q, f, rel_l2(got[f], alone[q][f]))
    # seed run 0 again with another density ratio: the cached ensemble graph
    # must not replay the old boundary densities
    c2 = dataclasses.replace(cfgs[0], s=cfgs[0].s * 1.7 + 0.5, seed=cfgs[0].seed + 1)
    print("re-seeded run 0 with s =", c2.s, flush=True)
    ref = _seeded(c2, capi.LOMA_MIXED)
    ref.advance(2)
    want = ref.download()
    ref.close()
    from loma_phd_amd import ic
    ic.seed_run(runs[0], c2)
    runs[0].use_graph(True)
    capi.advance_many(runs, 2)
    got = runs[0].download()
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), (f, rel_l2(got[f], want[f]))
    for r in runs:
        r.close()
