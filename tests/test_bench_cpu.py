"""CPU checks of the benchmark's contract and byte models, and of the seeded
draws shared by the host and device initial conditions (no GPU needed)."""
import os
import subprocess
import sys
import types

import numpy as np

from conftest import REPO

sys.path.insert(0, REPO)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_three_phase_byte_model_at_the_production_grid():
    """2640 S + 48 P (FP64) and 1728 S + 48 P (mixed) bytes per step at
    512x513x256: 42.14 and 28.70 GB."""
    import bench
    S, P = 170 * 169 * 513, 512 * 513 * 256
    assert abs(bench.model_bytes_per_step(S, P, False) / 1e9 - 42.137) < 1e-3
    assert abs(bench.model_bytes_per_step(S, P, True) / 1e9 - 28.695) < 1e-3


def test_algorithmic_bytes_per_kernel():
    """One compulsory read and write of each field per launch: the x lines
    at 512x513x256 move 19 W-sized transform fields and rhst twice."""
    import bench
    S, W, P = 170 * 169 * 513, 170 * 256 * 513, 512 * 513 * 256
    run = types.SimpleNamespace(S=S, W=W, P=P, precision=0)
    k, step = bench.algorithmic_bytes(run, False)
    assert k["xline"] == 19 * 16 * W + 2 * 8 * P == 7862870016
    assert k["zfft_inv"] == 8 * 16 * (S + W)
    assert k["zfft_fwd"] == 11 * 16 * (S + W)
    assert step == 3 * sum(k.values())
    mixed = types.SimpleNamespace(S=S, W=W, P=P, precision=1)
    km, _ = bench.algorithmic_bytes(mixed, False)
    assert km["xline"] == 19 * 8 * W + 2 * 8 * P
    kz, _ = bench.algorithmic_bytes(run, True)  # Z: 12 in, 12 out, rhsz too
    assert kz["xline"] == 24 * 16 * W + 4 * 8 * P


def test_seeded_draws_are_shared_and_reproducible():
    from loma_phd_amd import ic
    from loma_phd_amd.grid import MixingLayerConfig
    cfg = MixingLayerConfig(mgalx=48, my=97, mgalz=36, seed=12)
    c1, c2 = ic.ic_coefficients(cfg), ic.ic_coefficients(cfg)
    assert c1.shape == (3, cfg.nplanes, cfg.mz, ic.NYMAX)
    assert np.array_equal(c1, c2)
    other = ic.ic_coefficients(MixingLayerConfig(mgalx=48, my=97, mgalz=36, seed=13))
    assert not np.array_equal(c1, other)
    # |c_n| = |amp_n| <= 1 (amp ~ U(-1, 1), unit-modulus phase factor)
    assert np.all(np.abs(c1) <= 1.0 + 1e-15)
