"""The dense-operator trial: can the fixed y-derivatives run faster as one
dense matrix product on the matrix cores than as the banded solves?

The compact first derivative of the production grid (my = 513) is
D1 = diag(fmap) A1^-1 B1 (cfdiff.f:518-601: the 7-band B1 product, the
banded solve with A1, the mapping factor); applied to every mode column of
one field (S = 14,738,490 complex entries at 512x513x256, i.e. 57,460 real
columns) it is the FP64 product D1 (513 x 513) times U (513 x 57,460): 30.2
GFLOP.  This times that product with torch.matmul in FP64 (rocBLAS /
hipBLASLt, the FP64 matrix cores) against loma_deryr2_dev (the library's
banded wavefront solve) on the same columns, checks that both give the same
derivative, and prints one JSON line:

    python -m loma_phd_amd.dense_y_trial

The FP64 matrix-core peak of MI355X is 78.6 TFLOP/s, so the product cannot
take less than 0.38 ms per field; the banded solve reads and writes each
entry once (16 + 16 bytes per complex entry, 0.47 GB per field, 0.06 ms at
8 TB/s).
"""
from __future__ import annotations

import json
import time

import numpy as np


def dense_d1(y):
    from . import capi
    a1, b1, _, _, fmap = capi.fd_matrices(y)
    my = y.size
    A = np.zeros((my, my))
    B = np.zeros((my, my))
    for j in range(my):
        for d in range(7):
            c = j - 3 + d
            if 0 <= c < my:
                A[j, c] = a1[j, d]
                B[j, c] = b1[j, d]
    return fmap[:, None] * np.linalg.solve(A, B)


def main():
    import torch
    from . import capi
    from .grid import CONFIG_B
    cfg = CONFIG_B
    y = cfg.y()
    my = y.size
    S = cfg.spectral_size()
    ncols = S // my
    dev = torch.device("cuda:0")
    D = torch.from_numpy(dense_d1(y)).to(dev)
    g = torch.Generator(device="cpu").manual_seed(5)
    U = torch.randn((ncols, my, 2), dtype=torch.float64, generator=g).to(dev)  # column-major pairs
    run = capi.Run(cfg.mgalx, my, cfg.mgalz, y, cfg.Re, cfg.alp, cfg.bet)
    out_band = torch.empty_like(U)
    # dense: D @ [re | im] columns, as one (my x my) x (my x 2 ncols) product
    Ucols = U.permute(1, 0, 2).reshape(my, 2 * ncols).contiguous()

    def band():
        run.deryr2_dev(U.data_ptr(), out_band.data_ptr(), ncols)

    def dense():
        return D @ Ucols

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps * 1e3, r

    ms_band, _ = timed(band, 20)
    ms_dense, prod = timed(dense, 20)
    got = prod.reshape(my, ncols, 2).permute(1, 0, 2)
    err = (torch.linalg.norm(got - out_band) / torch.linalg.norm(out_band)).item()
    flop = 2.0 * my * my * 2 * ncols
    res = {
        "workload": f"d/dy of one field at {cfg.mgalx}x{my}x{cfg.mgalz}: {ncols} complex columns",
        "band_ms": round(ms_band, 4),
        "dense_fp64_matmul_ms": round(ms_dense, 4),
        "dense_tflops": round(flop / (ms_dense * 1e-3) / 1e12, 2),
        "dense_frac_of_fp64_mfma_peak": round(flop / (ms_dense * 1e-3) / 78.6e12, 3),
        "band_GBps": round(32.0 * S / (ms_band * 1e-3) / 1e9, 1),
        "rel_diff": err,
    }
    print(json.dumps(res), flush=True)
    run.close()


if __name__ == "__main__":
    main()
