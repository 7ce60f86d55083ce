"""Diagnostics of a mixing-layer state (host side, from the spectral state).

  momentum(state)      rho*u_i spectral fields incl. the mean (the phase-A
                       algebra of cross.f:352-499)
  mean_velocity(...)   x-z mean of u = rho*u * T by Parseval (the 00 mode
                       that foup2 transforms back, cross.f:1416)
  momentum_thickness   delta_m = (y_max - y_min)/4 - int(u00^2)/Du^2 with
                       the trapezoid weights of main.f:689-696
                       (cross.f:1416-1429)
  vorticity_thickness  delta_w = 1/max|du00/dy| (cross.f:692-699)
  field_errors(a, b)   per-field relative L2 of (rho u, rho v, rho w, T, Z)
"""
from __future__ import annotations

import numpy as np

from .ic import HostOps


def _along_y(ops, fn, a):
    return np.moveaxis(fn(np.moveaxis(a, 2, 0)), 0, 2)


def momentum(state, ops: HostOps, alp, bet):
    vor, phi, psi = state["vor"], state["phi"], state["psi"]
    np_, mz, my = vor.shape
    nz = (mz - 1) // 2
    kx = alp * np.arange(np_)
    k = np.arange(mz)
    kz = np.where(k <= nz, bet * k, -bet * (mz - k))
    rK = kx[:, None] ** 2 + kz[None, :] ** 2
    v = np.zeros_like(phi)
    nzm = rK != 0.0
    if np.any(nzm):
        v[nzm] = ops.helm_batch(phi[nzm], rK[nzm])
    dv = _along_y(ops, ops.dy1, v)
    with np.errstate(divide="ignore", invalid="ignore"):
        ta = np.where(rK > 0, 1j * kx[:, None] / rK, 0.0)[:, :, None]
        tb = np.where(rK > 0, 1j * kz[None, :] / rK, 0.0)[:, :, None]
    mx = dv * ta - vor * tb
    mzf = dv * tb + vor * ta
    dpsi = _along_y(ops, ops.dy1, psi)
    ru = mx + 1j * kx[:, None, None] * psi
    rw = mzf + 1j * kz[None, :, None] * psi
    rv = v + dpsi
    ru[0, 0] = state["u00"]
    rv[0, 0] = state["v00"]
    rw[0, 0] = state["w00"]
    return ru, rv, rw


def plane_mean(f, g):
    """x-z mean of the product of two real fields from their half spectra
    (kx >= 0 planes; kx > 0 counted twice), exact for band-limited fields."""
    w = np.ones(f.shape[0])
    w[1:] = 2.0
    return np.real(np.einsum("i,iky,iky->y", w, f, np.conj(g)))


def mean_velocity(state, ru):
    return plane_mean(ru, state["scal"])


def momentum_thickness(u00, y):
    my = y.size
    trp = np.empty(my)
    trp[1:-1] = 0.25 * (y[2:] - y[:-2])
    trp[0] = 0.25 * (y[1] - y[0])
    trp[-1] = 0.25 * (y[-1] - y[-2])
    s = np.sum(trp[1:-1] * (u00[2:] ** 2 + u00[:-2] ** 2))
    s += trp[0] * (u00[1] ** 2 + u00[0] ** 2) + trp[-1] * (u00[-1] ** 2 + u00[-2] ** 2)
    return 0.25 * (y[-1] - y[0]) - s / (u00[-1] - u00[0]) ** 2


def vorticity_thickness(u00, ops: HostOps):
    return 1.0 / np.max(np.abs(ops.dy1(u00)))


def field_errors(a, b, ops: HostOps, alp, bet):
    """Per-field relative L2 error of state a against reference state b on
    the physical fields (rho u, rho v, rho w, T, Z), means included."""
    ra = momentum(a, ops, alp, bet)
    rb = momentum(b, ops, alp, bet)
    out = {}
    for n, x, y in (("u", ra[0], rb[0]), ("v", ra[1], rb[1]), ("w", ra[2], rb[2]),
                    ("T", a["scal"], b["scal"])) + ((("Z", a["z"], b["z"]),)
                                                    if "z" in a and "z" in b else ()):
        out[n] = float(np.linalg.norm((x - y).ravel()) / np.linalg.norm(y.ravel()))
    return out
