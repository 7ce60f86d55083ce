"""Statistics files and the statistics schedule of a run.

The reference writes a statistics window to two Fortran sequential
unformatted files (writestats, trunk/lomacte3/src/escru.f:276-484):

  <name>_NNN.sta  record 1: my (int32), time, Re, alp, bet, a0 (float32),
                            nacum (int32)
                  record 2: y(j), fmap(j) pairs, j = 1..my (float64)
                  record 3: per y row, the 33 columns (float64) of
                            loma_stats_read (um ... num)
  <name>_NNN.spe  record 1: time, Re, alp, bet (float32), mx, my, mz,
                            nspec+1, nacumsp (int32)
                  record 2: jsptot (nspec+1 int32), then y(j), fmap(j)
                            pairs (float64)
                  then one record per kx plane: spwk(0:nz1, 1:nspec+1, 12)
                            (float32)

and cross1 accumulates over the last ntimes steps before every history
step (cross.f:234-239, 1145-1161).  `advance_with_stats` runs that schedule
on a device run (capi.Run): it opens the window before those steps, reads
it after the history step, hands the result to a callback (to write the
files and the .cf line) and resets it.
"""
from __future__ import annotations

import struct

import numpy as np

STA_COLUMNS = ("um vm wm rum rvm rwm Tm w1m w2m w3m them up vp wp uvr uwr vwr w1p w2p w3p "
               "Tp ep ruu ruv ruw rvv rvw rww rhom thep theup mum num").split()


def _record(f, payload: bytes):
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def read_records(path):
    """The records of a Fortran sequential unformatted file."""
    out = []
    with open(path, "rb") as f:
        while True:
            h = f.read(4)
            if not h:
                break
            n = struct.unpack("<i", h)[0]
            out.append(f.read(n))
            if struct.unpack("<i", f.read(4))[0] != n:
                raise ValueError(f"{path}: record length markers disagree")
    return out


def write_sta(path, stats, *, y, fmap, time, Re, alp, bet, a0=0.0):
    """The .sta file of one window (stats: the dict of Run.stats_read)."""
    sta = np.ascontiguousarray(stats["sta"], dtype="<f8")
    my = sta.shape[0]
    if sta.shape != (my, 33) or len(y) != my or len(fmap) != my:
        raise ValueError("sta must be (my, 33) with y and fmap of my rows")
    yf = np.stack([np.asarray(y, "<f8"), np.asarray(fmap, "<f8")], axis=1)
    with open(path, "wb") as f:
        _record(f, struct.pack("<i5fi", my, time, Re, alp, bet, a0, int(stats["nacum"])))
        _record(f, yf.tobytes())
        _record(f, sta.tobytes())


def write_spe(path, stats, *, jsptot, y, fmap, time, Re, alp, bet, mx, mz):
    """The .spe file of one window; jsptot are the 1-based spectra rows
    (the configured rows, then the centre row)."""
    spe = np.ascontiguousarray(stats["spe"], dtype="<f4")
    nplanes, nq, nsp, _ = spe.shape
    my = len(y)
    if nq != 12 or len(jsptot) != nsp:
        raise ValueError("spe must be (nplanes, 12, nspec+1, nz+1) with nspec+1 rows")
    yf = np.stack([np.asarray(y, "<f8"), np.asarray(fmap, "<f8")], axis=1)
    with open(path, "wb") as f:
        _record(f, struct.pack("<4f5i", time, Re, alp, bet, mx, my, mz, nsp,
                               int(stats["nacumsp"])))
        _record(f, np.asarray(jsptot, "<i4").tobytes() + yf.tobytes())
        for i in range(nplanes):
            _record(f, spe[i].tobytes())


def read_sta(path):
    r = read_records(path)
    my, time, Re, alp, bet, a0, nacum = struct.unpack("<i5fi", r[0])
    yf = np.frombuffer(r[1], "<f8").reshape(my, 2)
    return dict(my=my, time=time, Re=Re, alp=alp, bet=bet, a0=a0, nacum=nacum,
                y=yf[:, 0].copy(), fmap=yf[:, 1].copy(),
                sta=np.frombuffer(r[2], "<f8").reshape(my, 33).copy())


def read_spe(path):
    r = read_records(path)
    time, Re, alp, bet, mx, my, mz, nsp, nacumsp = struct.unpack("<4f5i", r[0])
    jsp = np.frombuffer(r[1][:4 * nsp], "<i4").copy()
    nz1 = (mz - 1) // 2
    spe = np.stack([np.frombuffer(b, "<f4").reshape(12, nsp, nz1 + 1) for b in r[2:]])
    return dict(time=time, Re=Re, alp=alp, bet=bet, mx=mx, my=my, mz=mz, nacumsp=nacumsp,
                jsptot=jsp, spe=spe)


def cf_line(history, stats=None):
    """The 19 columns of a .cf line (cross.f:1172-1176 without the two
    timers): time, dm, dw, Redw, trapz, masst, enerdis, energy(9), massu,
    massv, massw; history is Run.history(), stats the last window read
    (masst, enerdis and energy keep their last values between windows in the
    reference; zero before the first)."""
    st = stats or dict(masst=0.0, enerdis=0.0, energy=np.zeros(9))
    return np.array([history["time"], history["dm"], history["dw"], history["Redw"],
                     history["trapz"], st["masst"], st["enerdis"], *st["energy"],
                     history["massu"], history["massv"], history["massw"]])


def advance_with_stats(run, nsteps, *, nhist, ntimes, on_window=None, first_step=1):
    """Advance `run` by nsteps steps with cross1's statistics schedule
    (cross.f:234-239): the window is open for the steps istep with
    (istep-1) mod nhist >= nhist - ntimes, and read after every step with
    istep mod nhist == 0 that it covered; istep counts from first_step.
    `on_window(run, istep, stats)` receives each window before the reset.
    Steps outside the window run in batches (captured graphs)."""
    if not (1 <= ntimes <= nhist):
        raise ValueError("need 1 <= ntimes <= nhist")
    windows = []
    istep, last = first_step, first_step + nsteps - 1
    opened = False  # istati: set only at the window's first step (cross.f:236-237)
    while istep <= last:
        ph = (istep - 1) % nhist
        if ph < nhist - ntimes:
            n = min(nhist - ntimes - ph, last - istep + 1)
            run.advance(n)
            istep += n
            continue
        if ph == nhist - ntimes:
            run.stats_accumulate(True)
            opened = True
        n = min(nhist - ph, last - istep + 1)
        run.advance(n)
        istep += n
        if (istep - 1) % nhist == 0 and opened:  # writestats at ihist with istati
            st = run.stats_read()
            windows.append((istep - 1, st))
            if on_window is not None:
                on_window(run, istep - 1, st)
            run.stats_reset()
            run.stats_accumulate(False)
            opened = False
    return windows
