"""Writes tests/golden/ref_cross1_192x453x96.npz: the reference's own time
advance (trunk/lomacte3 cross1, cross.f:24-1259) run from a seeded restart.

The reference solver is built in place by `make -f oracle/Makefile.ref full`
(its shipped ctes3D: 192x453x96, 8 compute + 4 save ranks, the image's MPICH)
and run here with `mpirun -np 12` from:
  * a restart file in the reference layout (escru.f:125-140, read by getfil
    main.f:329-455) holding the seeded mixing-layer initial condition of
    loma_phd_amd.ic at 192x453x96 (s = 2, Re = 160, Lx = Lz = 8 pi,
    y in [-20, 20], perturbation amplitude 1e-2, seed 11), rounded to FP32
    as the file stores it;
  * an hre.dat (parsed by initcr, main.f:500-551): 11 steps, one image per
    step (nimag = 1), one .cf line per step (nhist = 1), CFL 0.5 every
    step (ncfl = 1), no statistics (nstart = 0).

The images written after 1, 3 and 10 steps are reduced to a small fixture:
  * time after each step (FP32, as the file stores it);
  * the 00 modes u00, v00, w00 (FP32);
  * for each of vor, phi, psi, scal: the y columns of its 8 most energetic
    (kx, kz) modes after 10 steps (complex64), and the L2 norm of every
    (kx, kz) column (float32), after 1, 3 and 10 steps;
  * the same column norms of the FP32 initial condition, so that a test can
    first check that it rebuilt the same input;
  * the .cf history lines of every step (cross.f:1166-1177) without the
    two timer columns;
  * the reference's per-step wall time (stdout column 15, cross.f:1183), the
    rank count and the host it ran on, and the per-step time of the FP64
    oracle on the same grid and host with all its cores (OpenMP), so that
    the restatement/reference speed ratio is measured on one machine.

Run in the container that holds the reference tree (about 2 minutes on 8
cores):
    python tests/golden/make_ref_cross1.py
"""
import os
import platform
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

BIN = os.path.join(REPO, "oracle", "_ref", "full", "lomacte3")
MPIRUN = "/opt/conda/bin/mpirun"
RANKS = 12          # numtot = numerop + numerosa of the shipped ctes3D
COMPUTE_RANKS = 8   # numerop
STEPS = (1, 3, 10)
NMODES = 8
OUT = os.path.join(HERE, "ref_cross1_192x453x96.npz")


def config():
    from loma_phd_amd import grid
    return grid.MixingLayerConfig(mgalx=192, my=453, mgalz=96, Re=160.0, s=2.0,
                                  amp=1e-2, seed=11)


def fp32_initial_condition(cfg):
    """The seeded initial condition as the restart file stores it (FP32)."""
    from loma_phd_amd import ic
    st = ic.initial_condition(cfg)
    out = dict(y=np.asarray(st["y"], np.float64))
    for f in ("vor", "phi", "psi", "scal"):
        out[f] = np.asarray(st[f]).astype(np.complex64).astype(np.complex128)
    for f in ("u00", "v00", "w00"):
        out[f] = np.asarray(st[f]).astype(np.float32).astype(np.float64)
    return st, out


def column_norms(f):
    return np.sqrt(np.sum(np.abs(f) ** 2, axis=2)).astype(np.float32)


def time_oracle(cfg, s32, steps=4):
    """Per-step seconds of the FP64 oracle (all cores) on the same state;
    the first step is left out like the reference's."""
    import time
    from oracle import oracle as O
    O.build()
    g = O.Grid(cfg.mgalx, cfg.my, cfg.mgalz, s32["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr)
    S = O.State(g, dt0=cfg.dt0, cfl=cfg.cfl, ncfl=cfg.ncfl, flags=1)
    S.set(s32["vor"], s32["phi"], s32["psi"], s32["scal"], s32["u00"], s32["v00"], s32["w00"])
    out = []
    for _ in range(steps):
        t = time.perf_counter()
        S.step(1)
        out.append(time.perf_counter() - t)
    return out[1:]


def hre_dat(cfg, d, nstep):
    return "\n".join([
        "CC Re alp bet -u", f"  {cfg.Re}  {cfg.alp:.10f}  {cfg.bet:.10f}  0.",
        "CC nstep nimag nhist", f"  {nstep}  1  1",
        "CC CFL ncfl", f"  {cfg.cfl}  {cfg.ncfl}",
        "CC first-file stats ntimes", "  1  0  1",
        "CC output", f"{d}/out", "CC input", f"{d}/ic.000", "CC stats", f"{d}/sta", ""])


def main():
    from loma_phd_amd import restart
    subprocess.check_call(["make", "-s", "-f", "oracle/Makefile.ref", "full"], cwd=REPO)
    cfg = config()
    st, ic32 = fp32_initial_condition(cfg)
    with tempfile.TemporaryDirectory(prefix="lomaref_") as d:
        restart.write(os.path.join(d, "ic.000"), st, time=0.0, Re=cfg.Re, alp=cfg.alp,
                      bet=cfg.bet, sidecar=False, with_z=False)
        nstep = max(STEPS) + 1  # images are written at the top of a step
        with open(os.path.join(d, "hre.dat"), "w") as f:
            f.write(hre_dat(cfg, d, nstep))
        r = subprocess.run([MPIRUN, "-np", str(RANKS), BIN], cwd=d, capture_output=True,
                           text=True, timeout=1800)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-3000:] + r.stderr[-3000:])
        walls = []
        for line in r.stdout.splitlines():
            m = re.match(r"^\s*(\d+)\s+(\S+D[+-]\d+\s+){14}", line)
            if m:
                cols = line.split()
                walls.append((int(cols[0]), float(cols[14].replace("D", "E"))))
        imgs = {n: restart.read(os.path.join(d, "out.%03d" % n), prefer_sidecar=False)
                for n in STEPS}
        # the .cf history, one line per step (format 324 of cross.f:1172-1176:
        # time, dm, dw, Redw, trapz, masst, enerdis, energy(9), massu,
        # massv, massw, wall, comm)
        with open(os.path.join(d, "sta.cf")) as f:
            cf = np.array([[float(v.replace("D", "E")) for v in line.split()]
                           for line in f if line.strip()])
    fx = dict(grid=np.array([cfg.mgalx, cfg.my, cfg.mgalz]),
              physics=np.array([cfg.Re, cfg.s, cfg.alp, cfg.bet, cfg.pr, cfg.cfl,
                                cfg.amp, cfg.seed, cfg.dt0]),
              steps=np.array(STEPS), compute_ranks=np.int64(COMPUTE_RANKS),
              ranks=np.int64(RANKS))
    for f in ("vor", "phi", "psi", "scal"):
        fx[f"ic_norms_{f}"] = column_norms(ic32[f])
        last = imgs[max(STEPS)][f]
        en = column_norms(last) ** 2
        en[0, 0] = en[0, 0] if f == "scal" else 0.0
        order = np.argsort(-en.ravel())[:NMODES]
        modes = np.array(np.unravel_index(order, en.shape)).T
        fx[f"modes_{f}"] = modes
        for n in STEPS:
            fx[f"cols_{f}_{n}"] = imgs[n][f][modes[:, 0], modes[:, 1]].astype(np.complex64)
            fx[f"norms_{f}_{n}"] = column_norms(imgs[n][f])
    fx["cf"] = cf[:, :19]  # the wall and communication timers are left out
    for n in STEPS:
        fx[f"time_{n}"] = np.float32(imgs[n]["time"])
        for f in ("u00", "v00", "w00"):
            fx[f"{f}_{n}"] = imgs[n][f].astype(np.float32)
    # per-step wall seconds, steps 2.. (the first carries the set-up)
    w = np.array([t for s, t in walls if s >= 2])
    fx["ref_step_wall_s"] = w
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    fx["host"] = np.array(f"{cpu}; {os.cpu_count()} cores; {platform.machine()}")
    fx["oracle_step_s"] = np.array(time_oracle(cfg, ic32))
    np.savez_compressed(OUT, **fx)
    print(f"wrote {OUT}: median reference step {np.median(w):.3f} s on {RANKS} ranks "
          f"({os.cpu_count()} cores); oracle step {np.median(fx['oracle_step_s']):.3f} s")


if __name__ == "__main__":
    main()
