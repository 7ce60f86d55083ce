"""Seeded initial condition of the variable-density temporal mixing layer.

Recipe of the reference's generator (trunk/initloma3/src/main.f:95-609):
random-amplitude/phase velocity modes with a Gaussian y-window and a
(kh/k0)^2 exp(-(kh/k0)^2) spectrum shape, made solenoidal by a curl, 2-D
modes (kx = 0 or kz = 0) removed, normalised to a target rms at the centre
plane; mean profiles rho00 = 1 + (s-1)/(s+1) tanh(.), U = 0.5 tanh(2y);
rho u_i = rho00 * u_i; psi from lap(psi) = div(rho u) (Lapvdvhom, Dirichlet),
Omega_y = i kz rho u - i kx rho w, phi = lap(rho v - d psi/dy); T = 1/rho00
carried in the 00 mode only.  The reference seeds from the wall clock
(main.f:36-45); here a numpy PCG64 seed makes the state reproducible.

Host-side set-up code (runs once per run, not on the time-advance path).
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import solve_banded

from . import capi
from .grid import MixingLayerConfig


class HostOps:
    """Compact-FD y operators on the host, from the library's FD set-up."""

    def __init__(self, y):
        self.y = np.asarray(y, dtype=np.float64)
        self.my = self.y.size
        self.a1, self.b1, self.a2, self.b2, self.fmap = capi.fd_matrices(self.y)

    @staticmethod
    def _ab(band):
        my, nb = band.shape
        p = nb // 2
        ab = np.zeros((nb, my))
        for j in range(my):
            for d in range(nb):
                c = j - p + d
                if 0 <= c < my:
                    ab[2 * p - d, c] = band[j, d]
        return ab

    @staticmethod
    def _apply(band, u):
        my, nb = band.shape
        p = nb // 2
        uu = u.reshape(my, -1)
        out = np.zeros_like(uu)
        for d in range(nb):
            off = d - p
            lo, hi = max(0, -off), min(my, my - off)
            out[lo:hi] += band[lo:hi, d, None] * uu[lo + off:hi + off]
        return out.reshape(u.shape)

    def dy1(self, u):
        """d/dy along axis 0 of u (my, ...)"""
        shp = u.shape
        v = u.reshape(self.my, -1)
        r = self._apply(self.b1, v)
        x = solve_banded((3, 3), self._ab(self.a1), r)
        return (x * self.fmap[:, None]).reshape(shp)

    def dy2(self, u):
        shp = u.shape
        v = u.reshape(self.my, -1)
        r = self._apply(self.b2, v)
        return solve_banded((2, 2), self._ab(self.a2), r).reshape(shp)

    def helm(self, f, rK):
        """Lapvdvhom: (B2 - rK A2) v = A2 f with v = 0 at both ends."""
        H = self.b2 - rK * self.a2
        H[0] = 0.0
        H[0, 2] = 1.0
        H[-1] = 0.0
        H[-1, 2] = 1.0
        r = self._apply(self.a2, f)
        r[0] = 0.0
        r[-1] = 0.0
        return solve_banded((2, 2), self._ab(H), r)

    def helm_batch(self, F, rK):
        """helm() for many modes at once: F (nmodes, my) complex, rK (nmodes,).
        Banded elimination without pivoting, vectorised over modes."""
        nm, my = F.shape
        rK = np.asarray(rK, dtype=np.float64)
        R = self._apply(self.a2, F.T.copy()).T  # (nm, my)
        R[:, 0] = 0.0
        R[:, -1] = 0.0
        H = self.b2[None, :, :] - rK[:, None, None] * self.a2[None, :, :]  # (nm, my, 5)
        H[:, 0, :] = 0.0
        H[:, 0, 2] = 1.0
        H[:, -1, :] = 0.0
        H[:, -1, 2] = 1.0
        U1 = np.zeros((nm, my))
        U2 = np.zeros((nm, my))
        D = np.zeros((nm, my))
        w = H.copy()
        for k in range(my):
            D[:, k] = w[:, k, 2]
            if k + 1 < my:
                U1[:, k] = w[:, k, 3]
            if k + 2 < my:
                U2[:, k] = w[:, k, 4]
            for i in (1, 2):
                if k + i >= my:
                    continue
                l = w[:, k + i, 2 - i] / D[:, k]
                w[:, k + i, 2 - i] = 0.0
                w[:, k + i, 3 - i] -= l * U1[:, k]
                w[:, k + i, 4 - i] -= l * U2[:, k]
                R[:, k + i] -= l * R[:, k]
        X = np.zeros_like(R)
        for k in range(my - 1, -1, -1):
            s = R[:, k].copy()
            if k + 1 < my:
                s -= U1[:, k] * X[:, k + 1]
            if k + 2 < my:
                s -= U2[:, k] * X[:, k + 2]
            X[:, k] = s / D[:, k]
        return X


def mean_profiles(cfg: MixingLayerConfig, y):
    """rho00, U, Z mean profiles (initloma3 main.f:360-373 with dm0 = 1/4)."""
    s = cfg.s
    rho = 1.0 + (s - 1.0) / (s + 1.0) * np.tanh(-2.0 * y)
    U = 0.5 * np.tanh(2.0 * y)
    Z = 0.5 * (1.0 + np.tanh(2.0 * y))
    return rho, U, Z


NYMAX = 8  # y harmonics of each perturbation mode (main.f:123-291)


def peak_wavenumber(cfg: MixingLayerConfig) -> float:
    """k0 of the perturbation spectrum (kh/k0)^2 exp(-(kh/k0)^2)."""
    return max(cfg.alp, cfg.bet) * 4.0


def ic_coefficients(cfg: MixingLayerConfig) -> np.ndarray:
    """The seeded draws of the perturbation: for u, v, w and each mode,
    c_n = amp_n exp(2 pi i phs_n) with amp_n ~ U(-1, 1), phs_n ~ U(0, 1)
    (main.f:123-160), the coefficient of exp(i (n+1) (2 kh / NYMAX) y).
    Shape (3, nplanes, mz, NYMAX) complex128; the same draws feed the host
    initial condition and the device one (loma_seed)."""
    np_, mz = cfg.nplanes, cfg.mz
    rng = np.random.Generator(np.random.PCG64(cfg.seed))
    out = np.empty((3, np_, mz, NYMAX), np.complex128)
    for c in range(3):
        amp = 2.0 * (0.5 - rng.random((np_, mz, NYMAX)))
        phs = rng.random((np_, mz, NYMAX))
        out[c] = amp * np.exp(2j * np.pi * phs)
    return out


def seed_run(run, cfg: MixingLayerConfig):
    """The initial condition of `cfg` built on the device of `run`
    (capi.Run.seed): only the seeded draws are made here."""
    run.seed(ic_coefficients(cfg), cfg.s, cfg.amp, cfg.env_width, peak_wavenumber(cfg))


def initial_condition(cfg: MixingLayerConfig, ops: HostOps | None = None):
    """Spectral state dict (vor, phi, psi, scal, u00, v00, w00, z) in the
    reference layout [kx plane][kz][y], complex128."""
    y = cfg.y()
    ops = ops or HostOps(y)
    np_, mz, my = cfg.nplanes, cfg.mz, cfg.my
    nz = (mz - 1) // 2
    coefs = ic_coefficients(cfg)
    kx = cfg.alp * np.arange(np_)
    kk = np.arange(mz)
    kz = np.where(kk <= nz, cfg.bet * kk, -cfg.bet * (mz - kk))
    kh = np.sqrt(kx[:, None] ** 2 + kz[None, :] ** 2)
    nymax = NYMAX
    k0 = peak_wavenumber(cfg)
    ywin = np.exp(-(y / cfg.env_width) ** 2)
    vel = []
    shape = (kh / k0) ** 2 * np.exp(-(kh / k0) ** 2)
    for comp in range(3):
        # sum_n amp_n exp(i (ky_n y + 2 pi phs_n)) with ky_n = (n+1) ky_1:
        # exp(i ky_n y) = e1^(n+1), so the sum is a polynomial in e1, summed
        # by Horner's rule (one complex exponential per point, not nymax)
        c = coefs[comp]
        e1 = np.exp(1j * (2.0 * kh / nymax)[:, :, None] * y[None, None, :])
        f = c[:, :, nymax - 1, None] * e1
        for n in range(nymax - 2, -1, -1):
            f += c[:, :, n, None]
            f *= e1
        vel.append(f * (shape[:, :, None] * ywin[None, None, :]))
    u, v, w = vel
    ikx = 1j * kx[:, None, None]
    ikz = 1j * kz[None, :, None]
    # solenoidal by a curl (main.f:240-257)
    dw = ops.dy1(np.moveaxis(w, 2, 0))
    du = ops.dy1(np.moveaxis(u, 2, 0))
    dw = np.moveaxis(dw, 0, 2)
    du = np.moveaxis(du, 0, 2)
    cu = dw - ikz * v
    cv = ikz * u - ikx * w
    cw = ikx * v - du
    u, v, w = cu, cv, cw
    for a in (u, v, w):
        a[:, 0, :] = 0.0   # kz = 0 modes
        a[0, :, :] = 0.0   # kx = 0 plane
    # normalise rms at the centre plane (main.f:295-350)
    jc = my // 2
    wgt = np.where(np.arange(np_)[:, None] == 0, 1.0, 2.0)
    for a in (u, v, w):
        rms = np.sqrt(np.sum(wgt * np.abs(a[:, :, jc]) ** 2))
        if rms > 0:
            a *= cfg.amp / rms
    rho, U, Z = mean_profiles(cfg, y)
    ru, rv, rw = rho * u, rho * v, rho * w
    # psi from lap psi = div(rho u)   (main.f:452-463)
    dv = np.moveaxis(ops.dy1(np.moveaxis(rv, 2, 0)), 0, 2)
    div = ikx * ru + dv + ikz * rw
    rK = kx[:, None] ** 2 + kz[None, :] ** 2
    psi = ops.helm_batch(div.reshape(-1, my), rK.ravel()).reshape(div.shape)
    psi[rK == 0.0] = 0.0
    vor = ikz * ru - ikx * rw
    dpsi = np.moveaxis(ops.dy1(np.moveaxis(psi, 2, 0)), 0, 2)
    my_ = rv - dpsi
    d2 = np.moveaxis(ops.dy2(np.moveaxis(my_, 2, 0)), 0, 2)
    phi = d2 - rK[:, :, None] * my_
    scal = np.zeros_like(vor)
    scal[0, 0, :] = 1.0 / rho
    z = np.zeros_like(vor)
    z[0, 0, :] = Z
    return dict(vor=vor, phi=phi, psi=psi, scal=scal, u00=rho * U, v00=np.zeros(my),
                w00=np.zeros(my), z=z, y=y)
