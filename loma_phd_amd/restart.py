"""Restart files in the reference's layout (Fortran unformatted sequential,
4-byte record markers), so runs can be resumed by either code.

Layout written by escru (trunk/lomacte3/src/escru.f:125-140) and read by
getfil (main.f:329-455):
  record 1: time f4, Re f4, alp f4, bet f4, a0 f4, mx i4, my i4, mz i4,
            (y f8, fmap f8) x my, (u00 f4, v00 f4, w00 f4) x my
  then mx/2 records, one per kx plane i:
            ((vor, phi, psi, scal)(j, k), j = 0..2my-1, k = 0..mz-1) as f4,
            i.e. an array (4, 2my, mz) in Fortran order, re/im interleaved in j.
The lomaHZB variant stores five fields per plane, the fifth being the mixture
fraction Z (branches/lomaHZB/src/escru.f:142-143).

Because the reference layout is FP32, an optional FP64 sidecar
(<name>.f64.npz) carries the full-precision state for exact resumption.
"""
from __future__ import annotations

import os

import numpy as np

_MARK = np.dtype("<i4")


def _record(f, payload: bytes):
    n = np.array([len(payload)], dtype=_MARK).tobytes()
    f.write(n)
    f.write(payload)
    f.write(n)


def _read_record(f) -> bytes:
    head = f.read(4)
    if len(head) < 4:
        raise EOFError("truncated restart file")
    n = int(np.frombuffer(head, dtype=_MARK)[0])
    data = f.read(n)
    tail = f.read(4)
    if len(data) != n or int(np.frombuffer(tail, dtype=_MARK)[0]) != n:
        raise ValueError("corrupt record markers in restart file")
    return data


def write(path, state, *, time, Re, alp, bet, a0=0.0, fmap=None, with_z=None,
          sidecar=True):
    """Write `state` (dict with vor, phi, psi, scal [nplanes][mz][my] complex,
    u00, v00, w00 [my], y [my], optional z) in the reference restart layout."""
    vor = np.asarray(state["vor"])
    nplanes, mz, my = vor.shape
    y = np.asarray(state["y"], dtype=np.float64)
    fm = np.ones(my) if fmap is None else np.asarray(fmap, dtype=np.float64)
    if with_z is None:
        with_z = "z" in state and state["z"] is not None
    nf = 5 if with_z else 4
    head = np.zeros(1, dtype=[("t", "<f4"), ("Re", "<f4"), ("alp", "<f4"), ("bet", "<f4"),
                              ("a0", "<f4"), ("mx", "<i4"), ("my", "<i4"), ("mz", "<i4")])
    head[0] = (time, Re, alp, bet, a0, 2 * nplanes, my, mz)
    yf = np.empty((my, 2), dtype="<f8")
    yf[:, 0], yf[:, 1] = y, fm
    uvw = np.empty((my, 3), dtype="<f4")
    uvw[:, 0], uvw[:, 1], uvw[:, 2] = state["u00"], state["v00"], state["w00"]
    names = ["vor", "phi", "psi", "scal"] + (["z"] if with_z else [])
    with open(path, "wb") as f:
        _record(f, head.tobytes() + yf.tobytes() + uvw.tobytes())
        for i in range(nplanes):
            plane = np.empty((mz, 2 * my, nf), dtype="<f4")  # Fortran (nf, 2my, mz)
            for q, n in enumerate(names):
                c = np.asarray(state[n][i])  # (mz, my) complex
                plane[:, 0::2, q] = c.real
                plane[:, 1::2, q] = c.imag
            _record(f, plane.tobytes())
    if sidecar:
        np.savez(path + ".f64.npz", **{n: np.asarray(state[n]) for n in names},
                 u00=state["u00"], v00=state["v00"], w00=state["w00"], y=y,
                 time=np.float64(time))


def read(path, *, my=None, mz=None, with_z=None, prefer_sidecar=True):
    """Read a restart file.  With my/mz given, modes are truncated or
    zero-padded like the reference's assign (main.f:841-902)."""
    side = path + ".f64.npz"
    with open(path, "rb") as f:
        r1 = _read_record(f)
        hdt = np.dtype([("t", "<f4"), ("Re", "<f4"), ("alp", "<f4"), ("bet", "<f4"),
                        ("a0", "<f4"), ("mx", "<i4"), ("my", "<i4"), ("mz", "<i4")])
        head = np.frombuffer(r1[:hdt.itemsize], dtype=hdt)[0]
        mxe, mye, mze = int(head["mx"]), int(head["my"]), int(head["mz"])
        off = hdt.itemsize
        yf = np.frombuffer(r1[off:off + 16 * mye], dtype="<f8").reshape(mye, 2)
        off += 16 * mye
        uvw = np.frombuffer(r1[off:off + 12 * mye], dtype="<f4").reshape(mye, 3)
        nplanes = mxe // 2
        planes = []
        for _ in range(nplanes):
            planes.append(_read_record(f))
    nfe = len(planes[0]) // (4 * 2 * mye * mze)
    names = ["vor", "phi", "psi", "scal", "z"][:nfe]
    raw = {n: np.zeros((nplanes, mze, mye), np.complex128) for n in names}
    for i, p in enumerate(planes):
        a = np.frombuffer(p, dtype="<f4").reshape(mze, 2 * mye, nfe)
        for q, n in enumerate(names):
            raw[n][i] = a[:, 0::2, q] + 1j * a[:, 1::2, q]
    out = dict(time=float(head["t"]), Re=float(head["Re"]), alp=float(head["alp"]),
               bet=float(head["bet"]), a0=float(head["a0"]), y=yf[:, 0].copy(),
               fmap=yf[:, 1].copy(), u00=uvw[:, 0].astype(np.float64),
               v00=uvw[:, 1].astype(np.float64), w00=uvw[:, 2].astype(np.float64))
    if prefer_sidecar and os.path.exists(side):
        with np.load(side, allow_pickle=False) as z:
            for n in names:
                raw[n] = z[n]
            for n in ("u00", "v00", "w00", "y"):
                out[n] = z[n]
            out["time"] = float(z["time"])
    my = my or mye
    mz = mz or mze
    for n in names:
        out[n] = assign(raw[n], my, mz)
    for n in ("u00", "v00", "w00"):
        v = np.zeros(my)
        m = min(my, mye)
        v[:m] = out[n][:m]
        out[n] = v
    return out


def assign(field, my, mz):
    """Re-grid modes from (mye, mze) to (my, mz) by truncation / zero fill of
    kz, keeping the positive and negative halves (assign, main.f:841-902)."""
    nplanes, mze, mye = field.shape
    mym, mzm = min(my, mye), min(mz, mze)
    klen = (mzm + 1) // 2
    out = np.zeros((nplanes, mz, my), np.complex128)
    out[:, :klen, :mym] = field[:, :klen, :mym]
    if klen > 1:
        out[:, mz - klen + 1:, :mym] = field[:, mze - klen + 1:, :mym]
    return out
