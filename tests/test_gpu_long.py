"""Long run of the production case: 2000 RK3 steps of 512x513x256 (density
ratio 4, Re 800, seeded perturbation 1e-3 with a Gaussian envelope) in FP64
and in mixed precision from the same initial condition (seeded on the
device).  The momentum thickness delta_m (cross.f:1416-1429) is computed on
the host from the downloaded states at the start and the end.  Bar: the
mixed form's delta_m growth over the run is within 1% of the FP64 form's.
The per-field error of the mixed state against the FP64 state at the end is
printed.  LOMA_LONG_STEPS changes the step count; with LOMA_LONG_OUT set,
the results are written there as JSON.  About 2 minutes on one MI355X."""
import dataclasses
import json
import os
import time

import pytest

pytestmark = pytest.mark.gpu


def _delta_m(state, ops, cfg):
    from loma_phd_amd import diagnostics as D
    ru, _, _ = D.momentum(state, ops, cfg.alp, cfg.bet)
    u00 = D.mean_velocity(state, ru)
    return float(D.momentum_thickness(u00, ops.y))


def test_momentum_thickness_growth_2000_steps(gpu):
    from loma_phd_amd import capi, diagnostics as D, ic
    from loma_phd_amd.grid import CONFIG_B
    cfg = dataclasses.replace(CONFIG_B, scalar_z=False)
    nsteps = int(os.environ.get("LOMA_LONG_STEPS", "2000"))
    chunk = 250
    seed = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, cfg.y(), cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                    cfg.cfl, cfg.ncfl, cfg.dt0, capi.LOMA_FP64, 0)
    ic.seed_run(seed, cfg)
    st = seed.download()
    seed.close()
    ops = ic.HostOps(cfg.y())
    dm0 = _delta_m(st, ops, cfg)
    print(f"delta_m(0) = {dm0:.8f}", flush=True)
    out = {"steps": nsteps, "delta_m0": dm0}
    finals = {}
    for name, prec in (("fp64", capi.LOMA_FP64), ("mixed", capi.LOMA_MIXED)):
        r = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, cfg.y(), cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                     cfg.cfl, cfg.ncfl, cfg.dt0, prec, 0)
        r.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
        r.use_graph(True)
        t1 = time.time()
        done = 0
        while done < nsteps:
            n = min(chunk, nsteps - done)
            r.advance(n)
            done += n
            print(f"{name}: {done} steps, {time.time() - t1:.1f} s, "
                  f"t = {r.scalars()['time']:.5f}", flush=True)
        finals[name] = r.download()
        out[f"time_{name}"] = r.scalars()["time"]
        r.close()
    dm = {k: _delta_m(v, ops, cfg) for k, v in finals.items()}
    err = D.field_errors(finals["mixed"], finals["fp64"], ops, cfg.alp, cfg.bet)
    g64, gmx = dm["fp64"] - dm0, dm["mixed"] - dm0
    rel = abs(gmx - g64) / abs(g64)
    out.update(delta_m=dm, growth_fp64=g64, growth_mixed=gmx, growth_rel_diff=rel, errors=err)
    print(f"delta_m after {nsteps} steps: fp64 {dm['fp64']:.8f} mixed {dm['mixed']:.8f}; "
          f"growth fp64 {g64:.6e} mixed {gmx:.6e} relative difference {rel:.3e}; "
          f"field errors {err}", flush=True)
    dest = os.environ.get("LOMA_LONG_OUT")
    if dest:
        with open(dest, "w") as f:
            json.dump(out, f, indent=1)
    assert g64 != 0.0
    assert rel <= 0.01
