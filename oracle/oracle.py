"""ctypes wrapper of the CPU FP64 oracle (oracle/loma_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libloma_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (-O2, OpenMP) into oracle/_build."""
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(HERE, "loma_oracle.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=gnu11",
               "-ffp-contract=off", src, "-o", LIB + ".tmp", "-lm"]
        subprocess.check_call(cmd)
        os.replace(LIB + ".tmp", LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.orc_grid_create.restype = vp
        L.orc_grid_create.argtypes = [ctypes.c_int] * 3 + [_dp] + [ctypes.c_double] * 4
        L.orc_grid_destroy.argtypes = [vp]
        L.orc_grid_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.orc_grid_matrices.argtypes = [vp] + [_dp] * 10
        L.orc_pred1m6o.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_prederiv2.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_gaussj.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_gaussj.restype = ctypes.c_int
        for n in ("orc_bandec", "orc_bandec7"):
            getattr(L, n).argtypes = [_dp, ctypes.c_int]
        for n in ("orc_banbks", "orc_banbks7"):
            getattr(L, n).argtypes = [_dp, ctypes.c_int, _dp]
        for n in ("orc_deryr2", "orc_deryyr2"):
            getattr(L, n).argtypes = [vp, _dp, _dp, ctypes.c_int]
        for n in ("orc_deryr", "orc_deryyr", "orc_inty8", "orc_inty8_ref"):
            getattr(L, n).argtypes = [vp, _dp, _dp]
        for n in ("orc_lapvdvhom", "orc_solvepsi"):
            getattr(L, n).argtypes = [vp, _dp, _dp, _dp, ctypes.c_double]
        L.orc_helmholtz_dirichlet.argtypes = [vp, _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.orc_visc.argtypes = [vp] + [_dp] * 6 + [ctypes.c_double] * 5
        for n in ("orc_cft", "orc_rft"):
            getattr(L, n).argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_state_create.restype = vp
        L.orc_state_create.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int]
        L.orc_state_destroy.argtypes = [vp]
        L.orc_state_set.argtypes = [vp] + [_dp] * 7
        L.orc_state_get.argtypes = [vp] + [_dp] * 7
        L.orc_state_history.argtypes = [vp] + [_dp] * 5
        L.orc_state_set_z.argtypes = [vp, _dp]
        L.orc_state_get_z.argtypes = [vp, _dp]
        L.orc_step.argtypes = [vp, ctypes.c_int]
        L.orc_step.restype = ctypes.c_int
        L.orc_substep.argtypes = [vp, ctypes.c_int]
        L.orc_substep.restype = ctypes.c_int
        L.orc_state_scalars.argtypes = [vp, _dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_state_set_cfl_ranks.argtypes = [vp, ctypes.c_int]
        L.orc_state_cf.argtypes = [vp, _dp]
        L.orc_foup2.argtypes = [vp, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_state_set_rhst.argtypes = [vp, _dp]
        L.orc_stats_config.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_stats_config.restype = ctypes.c_int
        L.orc_stats_accumulate.argtypes = [vp, ctypes.c_int]
        L.orc_stats_get.argtypes = [vp, _dp, _dp, _dp]
        L.orc_stats_reset.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Grid:
    """FP64 grid + compact-FD matrices (initcr/derivadas restatement)."""

    def __init__(self, mgalx, my, mgalz, y, Re, alp, bet, pr=0.7):
        self.y = _f64(y)
        assert self.y.size == my
        self.mgalx, self.my, self.mgalz = mgalx, my, mgalz
        self._h = lib().orc_grid_create(mgalx, my, mgalz, _p(self.y), Re, alp, bet, pr)
        d = (ctypes.c_int * 4)()
        lib().orc_grid_dims(self._h, d)
        self.mx, self.mz, self.nplanes, self.nz = list(d)
        self.Re, self.alp, self.bet, self.pr = Re, alp, bet, pr

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_grid_destroy(self._h)
            self._h = None

    def matrices(self):
        my = self.my
        out = {k: np.zeros(n * my) for k, n in
               (("prem1", 7), ("dt11", 7), ("dt12", 7), ("prem3", 5), ("dt21", 5),
                ("dt22", 5), ("fmap", 1), ("hy", 1), ("trp", 1), ("endder", 1))}
        lib().orc_grid_matrices(self._h, *[_p(out[k]) for k in
                                           ("prem1", "dt11", "dt12", "prem3", "dt21", "dt22",
                                            "fmap", "hy", "trp", "endder")])
        for k, n in (("prem1", 7), ("dt11", 7), ("dt12", 7), ("prem3", 5), ("dt21", 5),
                     ("dt22", 5)):
            out[k] = out[k].reshape(my, n)
        return out

    # complex columns: arrays of shape (m, my) complex128
    def _colop(self, fn, u):
        u = np.ascontiguousarray(u, dtype=np.complex128)
        shp = u.shape
        u2 = u.reshape(-1, self.my)
        out = np.zeros_like(u2)
        getattr(lib(), fn)(self._h, _p(u2.view(np.float64)), _p(out.view(np.float64)),
                           u2.shape[0])
        return out.reshape(shp)

    def deryr2(self, u):
        return self._colop("orc_deryr2", u)

    def deryyr2(self, u):
        return self._colop("orc_deryyr2", u)

    def deryr(self, u):
        u = _f64(u); o = np.zeros_like(u)
        lib().orc_deryr(self._h, _p(u), _p(o)); return o

    def deryyr(self, u):
        u = _f64(u); o = np.zeros_like(u)
        lib().orc_deryyr(self._h, _p(u), _p(o)); return o

    def inty8(self, du):
        du = _f64(du); o = np.zeros_like(du)
        lib().orc_inty8(self._h, _p(du), _p(o)); return o

    def inty8_ref(self, du):
        du = _f64(du); o = np.zeros_like(du)
        lib().orc_inty8_ref(self._h, _p(du), _p(o)); return o

    def _solve(self, fn, phi, rK):
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        v = np.zeros_like(phi); dv = np.zeros_like(phi)
        getattr(lib(), fn)(self._h, _p(phi.view(np.float64)), _p(v.view(np.float64)),
                           _p(dv.view(np.float64)), float(rK))
        return v, dv

    def lapvdvhom(self, phi, rK):
        return self._solve("orc_lapvdvhom", phi, rK)

    def solvepsi(self, phi, rK):
        return self._solve("orc_solvepsi", phi, rK)

    def helmholtz(self, f, rK, shift):
        f = np.ascontiguousarray(f, dtype=np.complex128)
        x = np.zeros_like(f)
        lib().orc_helmholtz_dirichlet(self._h, _p(f.view(np.float64)), _p(x.view(np.float64)),
                                      float(rK), float(shift))
        return x

    def visc(self, vorwk, v, dvdy, phiwk, rk1, rk2, dalbe, dtri, dtxi):
        """visc (trunk/mixing/src/laps.f:535) on complex columns (my,): returns
        the six outputs (vorwk = omega, vor, phi, v, dvdy, phiwk = phi)."""
        a = {k: np.array(x, dtype=np.complex128) for k, x in
             (("vorwk", vorwk), ("v", v), ("dvdy", dvdy), ("phiwk", phiwk))}
        vor = np.zeros_like(a["vorwk"])
        phi = np.zeros_like(a["vorwk"])
        lib().orc_visc(self._h, _p(a["vorwk"].view(np.float64)), _p(vor.view(np.float64)),
                       _p(phi.view(np.float64)), _p(a["v"].view(np.float64)),
                       _p(a["dvdy"].view(np.float64)), _p(a["phiwk"].view(np.float64)),
                       float(rk1), float(rk2), float(dalbe), float(dtri), float(dtxi))
        return dict(vorwk=a["vorwk"], vor=vor, phi=phi, v=a["v"], dvdy=a["dvdy"],
                    phiwk=a["phiwk"])

    def rK(self):
        """rK[i,k] = kx^2 + kz^2 in the reference mode order."""
        kx = self.alp * np.arange(self.nplanes)
        k = np.arange(self.mz)
        kz = np.where(k <= self.nz, self.bet * k, -self.bet * (self.mz - k))
        return kx[:, None] ** 2 + kz[None, :] ** 2


def cft(c, iopt):
    """FFTPACK-convention complex transform of rows of c (complex128 2-D)."""
    c = np.array(c, dtype=np.complex128, order="C")
    m, n = c.shape
    lib().orc_cft(_p(c.view(np.float64)), n, m, iopt)
    return c


def rft(c, n, iopt):
    """FFTPACK-convention real transform; c has rows of n+2 doubles."""
    c = np.array(c, dtype=np.float64, order="C")
    lib().orc_rft(_p(c), n, c.shape[0], iopt)
    return c


class State:
    """Oracle time-advance state (cross1 restatement)."""

    def __init__(self, grid: Grid, dt0=1e-4, cfl=0.5, ncfl=1, flags=0):
        self.g = grid
        self.flags = flags
        self._h = lib().orc_state_create(grid._h, dt0, cfl, ncfl, flags)
        self.S = grid.nplanes * grid.mz * grid.my

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_state_destroy(self._h)
            self._h = None

    def _shape(self):
        return (self.g.nplanes, self.g.mz, self.g.my)

    def set(self, vor, phi, psi, scal, u00, v00, w00, z=None):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128).reshape(self._shape())
                for a in (vor, phi, psi, scal)]
        o = [_f64(a) for a in (u00, v00, w00)]
        lib().orc_state_set(self._h, *[_p(a.view(np.float64)) for a in arrs], *[_p(a) for a in o])
        if z is not None:
            zz = np.ascontiguousarray(z, dtype=np.complex128).reshape(self._shape())
            lib().orc_state_set_z(self._h, _p(zz.view(np.float64)))

    def get(self):
        shp = self._shape()
        f = [np.zeros(shp, np.complex128) for _ in range(4)]
        o = [np.zeros(self.g.my) for _ in range(3)]
        lib().orc_state_get(self._h, *[_p(a.view(np.float64)) for a in f], *[_p(a) for a in o])
        out = dict(vor=f[0], phi=f[1], psi=f[2], scal=f[3], u00=o[0], v00=o[1], w00=o[2])
        if self.flags & 2:
            z = np.zeros(shp, np.complex128)
            lib().orc_state_get_z(self._h, _p(z.view(np.float64)))
            out["z"] = z
        return out

    def history(self):
        shp = self._shape()
        t12 = np.zeros(shp, np.complex128); t13 = np.zeros(shp, np.complex128)
        rhst = np.zeros((self.g.my, self.g.mgalz, self.g.mgalx))
        a = np.zeros(self.g.my); b = np.zeros(self.g.my)
        lib().orc_state_history(self._h, _p(t12.view(np.float64)), _p(t13.view(np.float64)),
                                _p(rhst), _p(a), _p(b))
        return dict(ten12=t12, ten13=t13, rhst=rhst, rf0u=a, rf0w=b)

    def step(self, n=1):
        return lib().orc_step(self._h, n)

    def foup2(self, inputs, rkstep, dt):
        """The transform stage alone (cross.f:1282-1430): inputs (8, nplanes,
        mz, my) complex -> outputs (11, nplanes, mz, my), with the state's
        rhst history updated; returns (outputs, dtr) (dtr at rkstep 1)."""
        shp = self._shape()
        a = np.ascontiguousarray(inputs, dtype=np.complex128).reshape((8,) + shp)
        o = np.zeros((11,) + shp, np.complex128)
        d = np.zeros(1)
        lib().orc_foup2(self._h, _p(a.view(np.float64)), _p(o.view(np.float64)), int(rkstep),
                        float(dt), _p(d))
        return o, float(d[0])

    def set_rhst(self, rhst):
        """Set the physical RHS history of T ([y][z][x], the reference's line
        order, cross.f:92)."""
        r = np.ascontiguousarray(rhst, dtype=np.float64)
        assert r.size == self.g.my * self.g.mgalz * self.g.mgalx
        lib().orc_state_set_rhst(self._h, _p(r))

    def history_cf(self):
        """trapz, massu, massv, massw of the last substep (cross.f:1013-1036)."""
        o = np.zeros(4)
        lib().orc_state_cf(self._h, _p(o))
        return dict(trapz=o[0], massu=o[1], massv=o[2], massw=o[3])

    def set_cfl_ranks(self, ranks):
        """Reduce the CFL maxima over `ranks` line slabs, as the reference's
        MPI_ALLREDUCE of per-rank sums does (cross.f:1778-1799)."""
        lib().orc_state_set_cfl_ranks(self._h, int(ranks))

    def substep(self, rk):
        return lib().orc_substep(self._h, rk)

    def stats_config(self, rows):
        """Spectra rows (0-based y indices, jsptot of main.f:698-703) and a
        reset of the statistics window."""
        r = np.ascontiguousarray(rows, dtype=np.int32)
        self._nsp = int(r.size)
        rc = lib().orc_stats_config(self._h, self._nsp,
                                    r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        if rc != 0:
            raise ValueError("spectra row outside the grid")

    def stats_accumulate(self, on=True):
        """istati: accumulate at rkstep 1 of the following steps."""
        lib().orc_stats_accumulate(self._h, 1 if on else 0)

    def stats(self):
        """writestats' outputs: sta (my, 33) .sta columns, spe (nplanes, 12,
        nsp, nz+1) spectra, and nacum, nacumsp, masst, enerdis, energy(9)."""
        g = self.g
        sta = np.zeros((g.my, 33))
        spe = np.zeros((g.nplanes, 12, self._nsp, g.nz + 1))
        info = np.zeros(13)
        lib().orc_stats_get(self._h, _p(sta), _p(spe), _p(info))
        return dict(sta=sta, spe=spe, nacum=int(info[0]), nacumsp=int(info[1]),
                    masst=info[2], enerdis=info[3], energy=info[4:13].copy())

    def stats_reset(self):
        lib().orc_stats_reset(self._h)

    def scalars(self):
        o = np.zeros(4)
        lib().orc_state_scalars(self._h, _p(o))
        return dict(dt=o[0], time=o[1], dtr=o[2], istep=int(o[3]))


def set_threads(n: int):
    lib().orc_set_threads(int(n))
