"""Momentum and vorticity thickness computed on the device every step
(LOMA_FLAG_HISTORY) against the host diagnostics of the same state:
delta_m of the state at the start of a step (the reference's rkstep-1
value, cross.f:1414-1429) and delta_w at the start of its third substep
(cross.f:692-699).  Bar: 1e-10 relative in FP64; 1e-6 in the mixed form,
whose mean velocity carries an FP32-transformed perturbation part."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-6)])
def test_thickness_matches_host(gpu, precision, tol):
    from loma_phd_amd import capi, diagnostics as D, ic
    from loma_phd_amd.grid import MixingLayerConfig
    print("thickness precision", precision, flush=True)
    cfg = MixingLayerConfig(mgalx=64, my=129, mgalz=64, s=3.0, Re=300.0, seed=4,
                            scalar_z=False)
    st = ic.initial_condition(cfg)
    ops = ic.HostOps(st["y"])
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, precision, capi.LOMA_FLAG_HISTORY)
    run.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    run.advance(2)
    start = run.download()          # the state at the start of the next step
    run.substep(1)
    run.substep(2)
    third = run.download()          # the state at the start of its third substep
    run.substep(3)
    dm, dw = run.thickness()

    def mean_u(s):
        return D.mean_velocity(s, D.momentum(s, ops, cfg.alp, cfg.bet)[0])

    dm_host = D.momentum_thickness(mean_u(start), st["y"])
    dw_host = D.vorticity_thickness(mean_u(third), ops)
    print("delta_m", dm, dm_host, "delta_w", dw, dw_host, flush=True)
    assert abs(dm - dm_host) <= tol * abs(dm_host)
    assert abs(dw - dw_host) <= tol * abs(dw_host)
    run.close()


def test_thickness_needs_flag(gpu):
    from loma_phd_amd import capi
    from loma_phd_amd.grid import MixingLayerConfig
    cfg = MixingLayerConfig(mgalx=48, my=97, mgalz=36, scalar_z=False)
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, cfg.y(), cfg.Re, cfg.alp, cfg.bet)
    with pytest.raises(capi.LomaError):
        run.thickness()
    run.close()


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-6)])
def test_cf_history_matches_oracle(gpu, oracle_mod, precision, tol):
    """trapz and the momentum integrals of the 00 modes (the .cf columns of
    cross.f:1013-1036) after two steps, GPU against the FP64 oracle."""
    from loma_phd_amd import capi, ic
    from loma_phd_amd.grid import MixingLayerConfig
    O = oracle_mod
    print("cf history precision", precision, flush=True)
    cfg = MixingLayerConfig(mgalx=48, my=97, mgalz=36, s=3.0, Re=300.0, seed=6,
                            scalar_z=False)
    st = ic.initial_condition(cfg)
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, precision, 0)
    run.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    run.advance(2)
    h = run.history()
    run.close()
    g = O.Grid(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr)
    S = O.State(g, dt0=cfg.dt0, cfl=cfg.cfl, ncfl=cfg.ncfl, flags=0)
    S.set(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    assert S.step(2) == 0
    o = S.history_cf()
    print("gpu", h, "oracle", o, flush=True)
    assert h["dm"] != h["dm"]  # NaN without LOMA_FLAG_HISTORY
    for name in ("trapz", "massu", "massv", "massw"):
        scale = max(abs(o["massu"]), 1.0) if name != "trapz" else max(abs(o["trapz"]), 1e-12)
        assert abs(h[name] - o[name]) <= tol * scale, (name, h[name], o[name])

