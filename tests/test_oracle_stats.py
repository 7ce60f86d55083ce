"""The oracle's statistics window against the reference's own.

tests/golden/ref_stats_192x453x96.npz holds what the reference's cross1
wrote with statistics on (make_ref_stats.py): two .sta/.spe files, each
accumulated at rkstep 1 of two steps (steps 2-3 and 5-6 of a 7-step run
from the seeded 192x453x96 restart), and the .cf lines whose masst, enerdis
and energy(9) columns come from them.  The oracle runs the same schedule
from the same FP32 initial condition (dt in FP32, CFL maxima reduced over 8
ranks as the reference's MPI_ALLREDUCE does) and must give:

  * every .sta column (means, spectral sums, physical sums, cross.f:1981-2291,
    1586-1607; escru.f:276-484) within the FP32-storage floor of the
    reference: 1e-4 relative L2 per column, looser only for the columns
    whose values are themselves near the floor (the means of w, rho w and
    the x vorticity, the velocity-divergence terms);
  * the spectra at every row inside the grid, summed over all kx planes and
    for the first 8 planes: 2e-4;
  * masst within 1e-6 relative (its first term reads wkstats(0,18), the
    element before the array's column in memory: rww of the top row; the
    oracle restates that and the bar would fail without it), enerdis 1e-5,
    energy(9) 1e-3 (the reference sums uner in FP32);
  * nacum = 2 for both files.
About a minute on 8 cores."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "ref_stats_192x453x96.npz")

NAMES = ("um vm wm rum rvm rwm Tm w1m w2m w3m them up vp wp uvr uwr vwr w1p w2p w3p Tp ep "
         "ruu ruv ruw rvv rvw rww rhom thep theup mum num").split()
# measured 1e-8 .. 3e-5 for the columns with the default bar; the loose ones
# (measured 2e-5 .. 2e-3) are the means of w, rho w and the x vorticity and
# the divergence terms, of size 1e-6 .. 1e-8 relative to u
COL_TOL = dict.fromkeys(NAMES, 1e-4)
COL_TOL.update(wm=2e-3, rwm=2e-3, w1m=1e-2, them=2e-3, thep=2e-3, theup=2e-3)
SPE_TOL = 2e-4


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


@pytest.fixture(scope="module")
def ref():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def oracle_stats(ref):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_ref_cross1 as base
    from oracle import oracle as O
    cfg = base.config()
    assert [cfg.mgalx, cfg.my, cfg.mgalz] == [int(v) for v in ref["grid"]]
    _, s32 = base.fp32_initial_condition(cfg)
    g = O.Grid(cfg.mgalx, cfg.my, cfg.mgalz, s32["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr)
    S = O.State(g, dt0=cfg.dt0, cfl=cfg.cfl, ncfl=cfg.ncfl, flags=1)
    S.set_cfl_ranks(8)
    S.set(s32["vor"], s32["phi"], s32["psi"], s32["scal"], s32["u00"], s32["v00"], s32["w00"])
    S.stats_config(ref["jsp"][ref["jsp_kept"]] - 1)
    nstep, nhist, ntimes = (int(v) for v in ref["window"])
    out = []
    for istep in range(1, nstep):  # the cross1 schedule, cross.f:236-239, 1145-1161
        if (istep - 1) % nhist == nhist - ntimes:
            S.stats_accumulate(True)
        S.step(1)
        if istep % nhist == 0:
            out.append(S.stats())
            S.stats_reset()
            S.stats_accumulate(False)
    return out


def test_window_schedule(ref, oracle_stats):
    assert len(oracle_stats) == 2
    for k, s in enumerate(oracle_stats):
        assert s["nacum"] == ref[f"sta_nacum_{k}"] == 2
        assert s["nacumsp"] == ref[f"spe_nacum_{k}"] == 2


def test_sta_columns_match_reference(ref, oracle_stats):
    for k, s in enumerate(oracle_stats):
        want = ref[f"sta_cols_{k}"]
        for c, name in enumerate(NAMES):
            if not np.any(want[:, c]):
                assert np.all(s["sta"][:, c] == 0.0), name  # w2m: the 00 mode of ikz u - ikx w
                continue
            e = _rel(s["sta"][:, c], want[:, c])
            print(f"file {k} {name}: {e:.2e}")
            assert e <= COL_TOL[name], (k, name, e)


def test_spectra_match_reference(ref, oracle_stats):
    for k, s in enumerate(oracle_stats):
        sp = s["spe"]
        for q in range(12):
            e_sum = _rel(sp[:, q].sum(axis=0), ref[f"spe_sum_{k}"][q])
            e_pl = _rel(sp[:8, q], ref[f"spe_planes_{k}"][:, q].astype(np.float64))
            print(f"file {k} spectrum {q}: sum {e_sum:.2e} planes {e_pl:.2e}")
            assert e_sum <= SPE_TOL and e_pl <= SPE_TOL, (k, q, e_sum, e_pl)


def test_cf_statistics_columns(ref, oracle_stats):
    # .cf lines of steps 1, 3, 6: masst (5), enerdis (6), energy(9) (7..15)
    for k, s in enumerate(oracle_stats):
        cf = ref["cf"][k + 1]
        assert abs(s["masst"] - cf[5]) <= 1e-6 * abs(cf[5]), (s["masst"], cf[5])
        assert abs(s["enerdis"] - cf[6]) <= 1e-5 * abs(cf[6]), (s["enerdis"], cf[6])
        e = np.abs(s["energy"] - cf[7:16]) / np.abs(cf[7:16])
        print(f"file {k} energy relative errors {e}")
        assert np.all(e <= 1e-3), e
