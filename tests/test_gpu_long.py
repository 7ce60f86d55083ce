"""Long run of the production case (runs only with LOMA_LONG_TESTS=1; it
takes several minutes): 2000 RK3 steps of 512x513x256 (density ratio 4,
Re 800, seeded perturbation 1e-3 with a Gaussian envelope) in FP64 and in
mixed precision from the same initial condition.  The momentum thickness
delta_m (cross.f:1416-1429) is computed on the host from the downloaded
state every 500 steps.  Bar: the mixed form's delta_m growth over the run is
within 1% of the FP64 form's; the per-field error of the mixed state against
the FP64 state is printed at every checkpoint.  With LOMA_LONG_OUT set, the
histories are written there as JSON."""
import json
import os
import time

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("LOMA_LONG_TESTS") != "1",
                                 reason="long run: set LOMA_LONG_TESTS=1")]


def _delta_m(state, ops, cfg):
    from loma_phd_amd import diagnostics as D
    ru, _, _ = D.momentum(state, ops, cfg.alp, cfg.bet)
    u00 = D.mean_velocity(state, ru)
    return float(D.momentum_thickness(u00, ops.y))


def test_momentum_thickness_growth_2000_steps(gpu):
    from loma_phd_amd import capi, diagnostics as D, ic
    from loma_phd_amd.grid import CONFIG_B
    import dataclasses
    cfg = dataclasses.replace(CONFIG_B, scalar_z=False)
    nsteps = int(os.environ.get("LOMA_LONG_STEPS", "2000"))
    every = 500
    t0 = time.time()
    st = ic.initial_condition(cfg)
    ops = ic.HostOps(st["y"])
    print(f"initial condition {time.time() - t0:.1f} s", flush=True)
    runs = {}
    for name, prec in (("fp64", capi.LOMA_FP64), ("mixed", capi.LOMA_MIXED)):
        r = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                     cfg.cfl, cfg.ncfl, cfg.dt0, prec, 0)
        r.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
        r.use_graph(True)
        runs[name] = r
    dm0 = _delta_m(st, ops, cfg)
    hist = {"step": [0], "fp64": [dm0], "mixed": [dm0], "time_fp64": [0.0],
            "time_mixed": [0.0], "err": [{}]}
    print(f"delta_m(0) = {dm0:.8f}", flush=True)
    done = 0
    while done < nsteps:
        n = min(every, nsteps - done)
        states = {}
        for name, r in runs.items():
            t1 = time.time()
            r.advance(n)
            states[name] = r.download()
            hist[f"time_{name}"].append(r.scalars()["time"])
            print(f"{name}: steps {done}..{done + n} in {time.time() - t1:.1f} s", flush=True)
        done += n
        hist["step"].append(done)
        for name in runs:
            hist[name].append(_delta_m(states[name], ops, cfg))
        err = D.field_errors(states["mixed"], states["fp64"], ops, cfg.alp, cfg.bet)
        hist["err"].append(err)
        print(f"step {done}: delta_m fp64 {hist['fp64'][-1]:.8f} mixed "
              f"{hist['mixed'][-1]:.8f} t {hist['time_fp64'][-1]:.5f}/"
              f"{hist['time_mixed'][-1]:.5f} errors {err}", flush=True)
    for r in runs.values():
        r.close()
    g64 = hist["fp64"][-1] - hist["fp64"][0]
    gmx = hist["mixed"][-1] - hist["mixed"][0]
    rel = abs(gmx - g64) / abs(g64)
    hist["growth_fp64"], hist["growth_mixed"], hist["growth_rel_diff"] = g64, gmx, rel
    print(f"delta_m growth fp64 {g64:.6e} mixed {gmx:.6e} relative difference {rel:.3e}",
          flush=True)
    out = os.environ.get("LOMA_LONG_OUT")
    if out:
        with open(out, "w") as f:
            json.dump(hist, f, indent=1)
    assert g64 != 0.0
    assert rel <= 0.01
