"""Device checkpoints and restart-on-NaN (loma_snapshot, loma_restore,
loma_set_cfl; loma_phd_amd.driver.advance_with_restart).

* A snapshot holds everything a step carries to the next: three steps from
  a snapshot, restored, and the same three steps again give the same state
  bit for bit, and the same step count and time.
* Restart on a non-finite state (initnan, trunk/mixing/src/main.f:953): a
  run started with a CFL number of 1e4 overflows within its first chunk of
  steps (the condition is made on the host through the CFL number; every
  kernel runs its normal bounded path on the non-finite values); the driver
  restores the checkpoint, lowers the CFL number to 0.5 and finishes.  The
  result equals a run advanced at CFL 0.5 from the start, bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("vor", "phi", "psi", "scal", "u00", "v00", "w00")


def _run(cfg, st, cfl, precision=0):
    from loma_phd_amd import capi
    r = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                 cfl, cfg.ncfl, cfg.dt0, precision, 0)
    r.upload(st["vor"], st["phi"], st["psi"], st["scal"], st["u00"], st["v00"], st["w00"])
    return r


@pytest.mark.parametrize("precision", [0, 1])
def test_snapshot_restore_repeats_steps(gpu, precision):
    from loma_phd_amd import ic
    from loma_phd_amd.grid import MixingLayerConfig
    print(f"snapshot/restore precision {precision}", flush=True)
    cfg = MixingLayerConfig(mgalx=48, my=97, mgalz=36, s=3.0, Re=300.0, seed=8,
                            scalar_z=False)
    st = ic.initial_condition(cfg)
    r = _run(cfg, st, cfg.cfl, precision)
    r.use_graph(True)
    r.advance(2)
    r.snapshot()
    r.advance(3)
    a, sa = r.download(), r.scalars()
    r.restore()
    assert r.scalars()["istep"] == 2
    r.advance(3)
    b, sb = r.download(), r.scalars()
    r.close()
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
    assert sa["time"] == sb["time"] and sa["dt"] == sb["dt"] and sa["istep"] == sb["istep"]


def test_restart_on_nan(gpu):
    from loma_phd_amd import driver, ic
    from loma_phd_amd.grid import MixingLayerConfig
    print("restart on NaN", flush=True)
    cfg = MixingLayerConfig(mgalx=48, my=97, mgalz=36, s=2.0, Re=160.0, seed=3,
                            scalar_z=False)
    st = ic.initial_condition(cfg)
    ref = _run(cfg, st, 0.5)
    ref.advance(10)
    want = ref.download()
    ref.close()
    r = _run(cfg, st, 1e4)
    log = driver.advance_with_restart(r, 10, every=5, cfl=1e4, cfl_factor=0.5e-4)
    got = r.download()
    sc = r.scalars()
    r.close()
    print("restart log", log, flush=True)
    assert log.restarts == 1 and log.events[0][1] == 0 and log.steps_done == 10
    assert abs(log.cfl - 0.5) < 1e-12
    assert sc["nan"] == 0 and np.isfinite(sc["dt"])
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f


def test_restart_file_continues_the_run(gpu, tmp_path):
    """A restart file written from a device run (the escru layout with its
    FP64 sidecar, loma_phd_amd.restart) holds the state exactly: read back,
    it equals the downloaded state, and a run started from the file steps
    bit for bit like a run started from the downloaded state itself.  (Both
    restarted runs take the boundary densities from the 00 mode of T as
    cross1 does on entry, cross.f:208-213; T's end rows carry the last
    substep's increment, so a restart is not the uninterrupted run, in the
    reference either.)  Without the sidecar the file's FP32 fields are the
    state to FP32 rounding."""
    from loma_phd_amd import capi, ic, restart
    from loma_phd_amd.grid import MixingLayerConfig
    print("restart file round trip 48x97x36", flush=True)
    cfg = MixingLayerConfig(mgalx=48, my=97, mgalz=36, s=3.0, Re=300.0, seed=8,
                            scalar_z=False)
    st = ic.initial_condition(cfg)
    a = _run(cfg, st, cfg.cfl)
    a.advance(3)
    sa = a.download()
    sc = a.scalars()
    a.close()
    path = str(tmp_path / "run.004")
    restart.write(path, dict(sa, y=st["y"]), time=sc["time"], Re=cfg.Re, alp=cfg.alp,
                  bet=cfg.bet, with_z=False)
    back = restart.read(path)
    assert back["time"] == sc["time"]
    for f in FIELDS:
        assert np.array_equal(back[f], sa[f]), f
    raw = restart.read(path, prefer_sidecar=False)
    for f in ("vor", "phi", "psi", "scal"):
        e = np.abs(raw[f] - sa[f]).max() / np.abs(sa[f]).max()
        assert e <= 2 ** -23, (f, e)

    def start(s):
        r = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                     cfg.cfl, cfg.ncfl, sc["dt"], 0, 0)
        r.upload(s["vor"], s["phi"], s["psi"], s["scal"], s["u00"], s["v00"], s["w00"])
        r.advance(2)
        out = r.download()
        r.close()
        return out

    da, db = start(sa), start(back)
    for f in FIELDS:
        assert np.array_equal(da[f], db[f]), f
