"""Pins the oracle's whole time advance against the reference's own cross1.

tests/golden/ref_cross1_192x453x96.npz holds what the reference solver
(trunk/lomacte3, cross1 at cross.f:24-1259, built in place by
oracle/Makefile.ref and run with mpirun over its shipped 8 compute + 4 save
ranks by tests/golden/make_ref_cross1.py) wrote after 1, 3 and 10 RK3 steps
from a seeded FP32 restart.  Here the FP64 oracle starts from the same FP32
state and advances with the reference's FP32 time step (ORC_DT_FP32) and its
8-rank CFL reduction (cross.f:1778-1799).

The reference keeps its state, transforms and spectral work arrays in FP32
(cross.f:77-91, cftsingle.f:16), so agreement is bounded by FP32 rounding of
the mean flow.  Measured: in the energetic modes the state agrees to 1.6e-7
(vor), 1.8e-6 (phi), 1.1e-6 (psi), 8e-8 (T) after 10 steps, the change over
the run to 2e-5 (phi) or better, the means u00, v00 to 1e-7 and the time
after 10 steps to 1e-7.  At perturbation amplitude 1e-3 the vor and phi differences are 10x
larger, at 1e-2 10x smaller: they are absolute FP32 noise, not a term of the
advance.  A wrong RK coefficient, a swapped xi/gamma, a wrong constraint
source or acoustic split changes the increments by O(1) and fails every bar
below by orders of magnitude.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "ref_cross1_192x453x96.npz")

STATE_TOL = 2e-5      # energetic columns, relative to the column norm
INC_TOL = {1: 1e-2, 3: 2e-3, 10: 2e-3}   # change since the start, relative
NORM_TOL = 1e-3       # per-mode column norms, energy weighted (FP32 noise: phi 2e-4)
TIME_TOL = 1e-6       # time after n steps, relative
MEAN_TOL = 1e-6       # u00, v00, w00 (momentum of the mean), absolute
# .cf history columns, absolute: the reference sums them in FP32 from its
# FP32 state (measured: trapz 2.8e-8, massu 5.9e-6 on -6.5, massv 4.1e-7,
# massw 7e-14 over the 10 steps)
CF_TOL = {"trapz": 5e-8, "massu": 2e-5, "massv": 1e-6, "massw": 1e-12}


@pytest.fixture(scope="module")
def ref():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def start(ref):
    from loma_phd_amd import grid, ic
    g = [int(v) for v in ref["grid"]]
    Re, s, alp, bet, pr, cfl, amp, seed, dt0 = [float(v) for v in ref["physics"]]
    cfg = grid.MixingLayerConfig(mgalx=g[0], my=g[1], mgalz=g[2], Re=Re, s=s, amp=amp,
                                 seed=int(seed), cfl=cfl, dt0=dt0, pr=pr)
    assert abs(cfg.alp - alp) < 1e-12 and abs(cfg.bet - bet) < 1e-12
    st = ic.initial_condition(cfg)
    # the restart file stores FP32 (escru.f:125-140)
    s32 = dict(y=np.asarray(st["y"], np.float64))
    for f in ("vor", "phi", "psi", "scal"):
        s32[f] = np.asarray(st[f]).astype(np.complex64).astype(np.complex128)
    for f in ("u00", "v00", "w00"):
        s32[f] = np.asarray(st[f]).astype(np.float32).astype(np.float64)
    return cfg, s32


def _norms(f):
    return np.sqrt(np.sum(np.abs(f) ** 2, axis=2))


def _oracle(oracle_mod, cfg, s32, ranks):
    O = oracle_mod
    g = O.Grid(cfg.mgalx, cfg.my, cfg.mgalz, s32["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr)
    S = O.State(g, dt0=cfg.dt0, cfl=cfg.cfl, ncfl=cfg.ncfl, flags=1)  # ORC_DT_FP32
    S.set_cfl_ranks(ranks)
    S.set(s32["vor"], s32["phi"], s32["psi"], s32["scal"], s32["u00"], s32["v00"],
          s32["w00"])
    return g, S


def test_same_initial_condition(ref, start):
    """The rebuilt FP32 start is the one the reference was run from."""
    _, s32 = start
    for f in ("vor", "phi", "psi", "scal"):
        a, b = _norms(s32[f]), ref[f"ic_norms_{f}"]
        assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b), f


def test_steps_match_reference_cross1(ref, start, oracle_mod):
    cfg, s32 = start
    g, S = _oracle(oracle_mod, cfg, s32, int(ref["compute_ranks"]))
    done = 0
    for n in [int(v) for v in ref["steps"]]:
        assert S.step(n - done) == 0
        done = n
        a = S.get()
        t_ref = float(ref[f"time_{n}"])
        t = S.scalars()["time"]
        print(f"step {n}: time {t:.9g} vs {t_ref:.9g}")
        assert abs(t - t_ref) <= TIME_TOL * t_ref, (n, t, t_ref)
        for f in ("u00", "v00", "w00"):
            err = np.abs(a[f] - ref[f"{f}_{n}"]).max()
            print(f"  {f}: max abs err {err:.2e}")
            assert err <= MEAN_TOL, (n, f, err)
        # .cf history of step n (cross.f:1166-1177): trapz (column 5) and the
        # momentum of the 00 modes (columns 17-19)
        cf = ref["cf"][n - 1]
        h = S.history_cf()
        for name, col in (("trapz", 4), ("massu", 16), ("massv", 17), ("massw", 18)):
            tol = CF_TOL[name]
            err = abs(h[name] - cf[col])
            print(f"  .cf {name}: {h[name]:.9e} vs {cf[col]:.9e} (abs err {err:.1e})")
            assert err <= tol, (n, name, h[name], cf[col])
        for f in ("vor", "phi", "psi", "scal"):
            m = ref[f"modes_{f}"]
            cr = ref[f"cols_{f}_{n}"].astype(np.complex128)
            ca = a[f][m[:, 0], m[:, 1]]
            c0 = s32[f][m[:, 0], m[:, 1]]
            e_state = np.linalg.norm(ca - cr) / np.linalg.norm(cr)
            e_inc = np.linalg.norm(ca - cr) / np.linalg.norm(cr - c0)
            na, nr = _norms(a[f]), ref[f"norms_{f}_{n}"].astype(np.float64)
            e_norm = np.linalg.norm(na - nr) / np.linalg.norm(nr)
            print(f"  {f}: state {e_state:.2e} increment {e_inc:.2e} norms {e_norm:.2e}")
            assert e_state <= STATE_TOL, (n, f, e_state)
            if f != "scal":  # T changes by less than its FP32 spacing per step
                assert e_inc <= INC_TOL[n], (n, f, e_inc)
            assert e_norm <= NORM_TOL, (n, f, e_norm)


def test_single_rank_cfl_is_resolved(ref, start, oracle_mod):
    """Negative control: the fixture resolves the reference's 8-rank CFL
    reduction.  With the one-rank reduction (the single-GPU semantics) the
    time step after the first step departs from the reference's by 1.5e-4,
    far above TIME_TOL."""
    cfg, s32 = start
    _, S = _oracle(oracle_mod, cfg, s32, 1)
    assert S.step(2) == 0
    sc = S.scalars()
    t3 = sc["time"] + sc["dt"]  # time after step 3 = time after 2 + dt of step 3
    t_ref = float(ref["time_3"])
    assert abs(t3 - t_ref) > 20 * TIME_TOL * t_ref, (t3, t_ref)
