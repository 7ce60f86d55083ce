"""The statistics window on the GPU (loma_stats_*) against the oracle.

The oracle's statistics are pinned to the reference's own .sta/.spe/.cf
output (tests/test_oracle_stats.py); here the device path runs cross1's
schedule (6 steps, nhist 3, ntimes 2: windows over steps 2-3 and 5-6,
loma_phd_amd.stats.advance_with_stats) from a seeded state next to the
oracle running the same schedule.

* FP64: every .sta column, the spectra and masst / enerdis / energy(9)
  within 1e-9 relative L2 of the oracle (the two sum in different orders;
  the columns whose values sit at round-off of the others, means of fields
  that vanish in this flow, are held in absolute terms instead).
* Mixed (FP32 transforms): the same within 1e-5 of the oracle, relative to
  each column's norm or to 1e-3 of the largest column's, whichever is
  larger: the FP32 rounding is relative to the mean flow, so a column that
  is itself small (the means of w, rho w and the x vorticity, the
  velocity-divergence terms) carries it in absolute terms.
* The window only reads the state: a run with the window open ends bit for
  bit where the same run without it ends.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEP, NHIST, NTIMES = 6, 3, 2


def _cfg(mgalx, my, mgalz):
    from loma_phd_amd.grid import MixingLayerConfig
    return MixingLayerConfig(mgalx=mgalx, my=my, mgalz=mgalz, Re=160.0, s=2.0, seed=5,
                             scalar_z=False)


def _rows(my):
    # spectra rows across the layer (1-based, as the reference's jspecy)
    return np.unique(np.linspace(my // 8, my - my // 8, 9).astype(int))


def _gpu_windows(cfg, st, precision, stats=True):
    from loma_phd_amd import capi, stats as S
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, precision, 0)
    run.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    run.use_graph(True)
    wins = []
    if stats:
        run.stats_config(_rows(cfg.my))
        wins = S.advance_with_stats(run, NSTEP, nhist=NHIST, ntimes=NTIMES)
    else:
        run.advance(NSTEP)
    final = run.download()
    run.close()
    return wins, final


def _oracle_windows(O, cfg, st):
    g = O.Grid(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr)
    S = O.State(g, dt0=cfg.dt0, cfl=cfg.cfl, ncfl=cfg.ncfl, flags=0)
    S.set(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    rows = list(_rows(cfg.my) - 1) + [(cfg.my + 1) // 2 - 1]
    S.stats_config(rows)
    out = []
    for istep in range(1, NSTEP + 1):
        if (istep - 1) % NHIST == NHIST - NTIMES:
            S.stats_accumulate(True)
        S.step(1)
        if istep % NHIST == 0:
            out.append(S.stats())
            S.stats_reset()
            S.stats_accumulate(False)
    return out


def _check(gw, ow, tol, label, floor=1e-9):
    """Each .sta column: ||gpu - oracle|| / max(||oracle||, floor * the
    largest column norm) <= tol; the floor keeps columns at round-off (w2m,
    means the flow does not carry) to an absolute bar."""
    from loma_phd_amd.stats import STA_COLUMNS
    assert [s for s, _ in gw] == [3, 6]
    for (step, g), o in zip(gw, ow):
        assert g["nacum"] == o["nacum"] == NTIMES
        big = np.linalg.norm(o["sta"], axis=0).max()
        errs = {}
        for c, name in enumerate(STA_COLUMNS):
            a, b = g["sta"][:, c], o["sta"][:, c]
            errs[name] = np.linalg.norm(a - b) / max(np.linalg.norm(b), floor * big)
        print(f"{label} window after step {step}: " +
              " ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
        worst = max(errs.values())
        for name, err in errs.items():
            assert err <= tol, (label, step, name, err)
        gs = g["spe"].astype(np.float64)
        e_spe = np.linalg.norm(gs - o["spe"]) / np.linalg.norm(o["spe"])
        # spectra come back as float32, as the .spe file stores them
        assert e_spe <= max(tol, 2e-7), (label, step, e_spe)
        for k in ("masst", "enerdis"):
            assert abs(g[k] - o[k]) <= tol * abs(o[k]), (label, step, k, g[k], o[k])
        e_en = np.abs(g["energy"] - o["energy"]) / np.maximum(
            np.abs(o["energy"]), floor * np.abs(o["energy"]).max())
        assert np.all(e_en <= tol), (label, step, e_en)
        print(f"{label} window after step {step}: worst .sta column {worst:.2e}, spectra "
              f"{e_spe:.2e}, energy {e_en.max():.2e}", flush=True)


@pytest.mark.parametrize("shape", [(64, 129, 64), (96, 257, 72)])
def test_stats_fp64_match_oracle(gpu, oracle_mod, shape):
    from loma_phd_amd import capi, ic
    cfg = _cfg(*shape)
    print(f"statistics window fp64 {shape}", flush=True)
    st = ic.initial_condition(cfg)
    gw, _ = _gpu_windows(cfg, st, capi.LOMA_FP64)
    ow = _oracle_windows(oracle_mod, cfg, st)
    _check(gw, ow, 1e-9, "fp64")


def test_stats_mixed_match_oracle(gpu, oracle_mod):
    from loma_phd_amd import capi, ic
    cfg = _cfg(64, 129, 64)
    print("statistics window mixed (64, 129, 64)", flush=True)
    st = ic.initial_condition(cfg)
    gw, _ = _gpu_windows(cfg, st, capi.LOMA_MIXED)
    ow = _oracle_windows(oracle_mod, cfg, st)
    _check(gw, ow, 1e-5, "mixed", floor=1e-3)


def test_stats_window_leaves_the_advance_unchanged(gpu):
    from loma_phd_amd import capi, ic
    cfg = _cfg(64, 129, 64)
    print("statistics window: same final state with and without", flush=True)
    st = ic.initial_condition(cfg)
    _, a = _gpu_windows(cfg, st, capi.LOMA_FP64, stats=True)
    _, b = _gpu_windows(cfg, st, capi.LOMA_FP64, stats=False)
    for f in ("vor", "phi", "psi", "scal", "u00", "v00", "w00"):
        assert np.array_equal(a[f], b[f]), f
