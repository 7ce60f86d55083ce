"""CPU-side checks of the C ABI library and the host logic (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import REPO, rel_l2


def test_library_exports_every_declared_symbol():
    from loma_phd_amd import capi
    L = capi.lib()
    header = open(os.path.join(REPO, "include", "loma_amd.h")).read()
    declared = set(re.findall(r"\b(loma_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(L, name), name
    assert set(capi.EXPORTS) == declared


def test_no_gpu_calls_fail_cleanly():
    from loma_phd_amd import capi
    n = capi.device_count()
    if n > 0:
        pytest.skip("a GPU is visible")
    y = np.linspace(-1, 1, 33)
    with pytest.raises(capi.LomaError) as e:
        capi.Run(32, 33, 32, y, 100.0, 1.0, 1.0)
    assert e.value.code == capi.LOMA_ERR_NODEV


def test_fd_matrices_match_oracle(oracle_mod):
    from loma_phd_amd import capi, grid
    for my in (65, 129, 513):
        y = grid.tanh_mesh(my, 20.0)
        g = oracle_mod.Grid(48, my, 48, y, 100.0, 0.25, 0.25)
        M = g.matrices()
        a1, b1, a2, b2, fm = capi.fd_matrices(y)
        assert np.array_equal(a1, M["dt11"]) and np.array_equal(b1, M["dt12"])
        assert np.array_equal(fm, M["fmap"])
        assert rel_l2(a2, M["dt21"]) < 1e-12 and rel_l2(b2, M["dt22"]) < 1e-12


def test_fd_matrices_reject_bad_sizes():
    from loma_phd_amd import capi
    with pytest.raises(capi.LomaError):
        capi.fd_matrices(np.linspace(-1, 1, 4))


def test_mesh_and_configs():
    from loma_phd_amd import grid
    y = grid.tanh_mesh(129, 20.0)
    assert y[0] == -20.0 and y[-1] == 20.0
    assert np.all(np.diff(y) > 0)
    assert np.allclose(y, -y[::-1])
    d = np.diff(y)
    assert d[0] > 2.5 * d[64]  # fine in the shear-layer centre
    a = grid.CONFIG_A
    assert (a.nplanes, a.mz, a.spectral_size()) == (21, 41, 111069)
    b = grid.CONFIG_B
    assert (b.nplanes, b.mz, b.spectral_size()) == (170, 169, 14738490)


def test_initial_condition_is_seeded_and_consistent(oracle_mod):
    from loma_phd_amd import grid, ic
    cfg = grid.MixingLayerConfig(mgalx=32, my=65, mgalz=32, Ly=10.0, seed=4)
    a = ic.initial_condition(cfg)
    b = ic.initial_condition(cfg)
    for k in ("vor", "phi", "psi"):
        assert np.array_equal(a[k], b[k])
    c = ic.initial_condition(grid.MixingLayerConfig(mgalx=32, my=65, mgalz=32, Ly=10.0,
                                                    seed=5))
    assert not np.array_equal(a["vor"], c["vor"])
    # psi solves Lapvdvhom of div(rho u) (checked with the oracle's operator)
    g = oracle_mod.Grid(32, 65, 32, a["y"], cfg.Re, cfg.alp, cfg.bet)
    from loma_phd_amd.diagnostics import momentum
    ops = ic.HostOps(a["y"])
    ru, rv, rw = momentum(a, ops, cfg.alp, cfg.bet)
    rK = g.rK()
    kx = cfg.alp * np.arange(g.nplanes)
    kz = np.sqrt(rK[0])
    kz = np.where(np.arange(g.mz) <= g.nz, 1, -1) * kz
    # the reconstructed momentum is divergence consistent: div = lap psi
    div = 1j * kx[:, None, None] * ru + g.deryr2(rv) + 1j * kz[None, :, None] * rw
    div[0, 0] = 0
    for i, k in ((1, 2), (3, 5), (2, g.mz - 3)):
        v, _ = g.lapvdvhom(div[i, k], rK[i, k])
        assert rel_l2(v, a["psi"][i, k]) < 1e-4
    # mean profiles
    assert a["scal"][0, 0, 0].real == pytest.approx((1 + cfg.s) / (2 * cfg.s))
    assert np.abs(a["scal"][1:]).max() == 0.0


def test_restart_round_trip(tmp_path):
    from loma_phd_amd import grid, ic, restart
    cfg = grid.MixingLayerConfig(mgalx=32, my=65, mgalz=32, Ly=10.0, seed=2)
    st = ic.initial_condition(cfg)
    p = str(tmp_path / "run.001")
    restart.write(p, st, time=1.5, Re=cfg.Re, alp=cfg.alp, bet=cfg.bet)
    r = restart.read(p, prefer_sidecar=False)
    for k in ("vor", "phi", "psi", "scal", "z"):
        assert rel_l2(r[k], st[k]) < 1e-7
    assert r["time"] == 1.5 and r["y"][0] == st["y"][0]
    r2 = restart.read(p)
    for k in ("vor", "phi", "psi", "scal", "z"):
        assert np.array_equal(r2[k], st[k])
    # byte layout: record 1 length and first plane record
    raw = open(p, "rb").read()
    n1 = np.frombuffer(raw[:4], "<i4")[0]
    assert n1 == 5 * 4 + 3 * 4 + 16 * 65 + 12 * 65
    n2 = np.frombuffer(raw[8 + n1:12 + n1], "<i4")[0]
    assert n2 == 5 * 4 * 2 * 65 * cfg.mz
    # first plane, k=0: (vor, phi, psi, scal, z) of Re part at j=0
    first = np.frombuffer(raw[12 + n1:12 + n1 + 20], "<f4")
    assert first[3] == np.float32(st["scal"][0, 0, 0].real)


def test_restart_assign_regrid():
    from loma_phd_amd import restart
    f = np.arange(2 * 5 * 4).reshape(2, 5, 4).astype(np.complex128)
    g = restart.assign(f, 4, 9)   # zero-pad kz: 5 -> 9 modes
    assert np.array_equal(g[:, :3], f[:, :3])
    assert np.array_equal(g[:, -2:], f[:, -2:])
    assert np.all(g[:, 3:7] == 0)
    h = restart.assign(g, 4, 5)
    assert np.array_equal(h, f)
