"""One RK3 step of the 512x513x256 production case (density ratio 4, Re 800)
in one precision, then a plain 1 GiB device-to-device copy, for hardware
counter runs:

    rocprofv3 --pmc FETCH_SIZE -- python3 -m loma_phd_amd.profile_step fp64
    rocprofv3 --pmc WRITE_SIZE -- python3 -m loma_phd_amd.profile_step fp64

Optional arguments pick another grid and the passive scalar Z:

    python3 -m loma_phd_amd.profile_step mixed 256 257 128 z

The copy moves a known number of bytes (1 GiB read, 1 GiB written), so the
counters it reports give the factor between a counter's KiB and HBM bytes.
Graph replay is off so that every kernel is its own dispatch.
"""
from __future__ import annotations

import dataclasses
import sys


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    prec = argv[0] if argv else "fp64"
    import torch
    from . import capi, ic
    from .grid import CONFIG_B
    cfg = dataclasses.replace(CONFIG_B, scalar_z=False)
    if len(argv) >= 4:
        cfg = dataclasses.replace(cfg, mgalx=int(argv[1]), my=int(argv[2]), mgalz=int(argv[3]),
                                  scalar_z=len(argv) >= 5 and argv[4] == "z")
    p = capi.LOMA_FP64 if prec == "fp64" else capi.LOMA_MIXED
    flags = capi.LOMA_FLAG_SCALAR_Z if cfg.scalar_z else 0
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, cfg.y(), cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, p, flags)
    print("initial condition (built on the device)", flush=True)
    ic.seed_run(run, cfg)
    run.use_graph(False)
    run.advance(1)
    print(f"{prec} step done, {cfg.mgalx}x{cfg.my}x{cfg.mgalz}, tiles {run.tiles()}", flush=True)
    n = 1 << 27  # doubles: 1 GiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"copy of {n * 8} bytes done", flush=True)
    run.close()


if __name__ == "__main__":
    main()
