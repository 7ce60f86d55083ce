"""ctypes binding of libloma_hip.so (the C ABI declared in include/loma_amd.h).

The product path: every call here runs the HIP kernels built for gfx950.
Importing this module loads the library; it raises if the library has not
been built (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB = os.environ.get("LOMA_HIP_LIB") or os.path.join(PKG, "_lib", "libloma_hip.so")

LOMA_OK, LOMA_ERR_ARG, LOMA_ERR_NAN, LOMA_ERR_OOM = 0, 1, 2, 3
LOMA_ERR_DEVICE, LOMA_ERR_STATE, LOMA_ERR_NODEV = 4, 5, 6
LOMA_FP64, LOMA_MIXED = 0, 1
LOMA_FLAG_DT_FP32, LOMA_FLAG_SCALAR_Z, LOMA_FLAG_AUTOTUNE, LOMA_FLAG_HISTORY = 1, 2, 4, 8

# every symbol the header declares (checked by tests on the CPU)
EXPORTS = [
    "loma_version", "loma_last_error", "loma_device_count", "loma_fd_matrices",
    "loma_init", "loma_free", "loma_dims", "loma_upload", "loma_download",
    "loma_upload_z", "loma_download_z", "loma_advance", "loma_substep", "loma_scalars",
    "loma_cross1", "loma_set_stream", "loma_use_graph", "loma_deryr2_dev",
    "loma_deryyr2_dev", "loma_lapvdvhom_dev", "loma_solvepsi_dev", "loma_helmholtz_dev",
    "loma_inty8_dev", "loma_deryr2", "loma_deryyr2", "loma_lapvdvhom", "loma_solvepsi",
    "loma_fourz_dev", "loma_band5_solve_dev", "loma_band5_solve_dev_noalloc",
    "loma_advance_many", "loma_advance_async",
    "loma_tiles", "loma_footprint", "loma_step_begin", "loma_step_poll",
    "loma_visc_dev", "loma_visc", "loma_set_physics", "loma_seed", "loma_thickness",
    "loma_sync", "loma_stream", "loma_profile", "loma_profile_read",
    "loma_deryr", "loma_deryyr", "loma_inty8", "loma_cft", "loma_rft", "loma_fourz",
    "loma_fourx", "loma_history", "loma_snapshot", "loma_restore", "loma_set_cfl",
    "loma_foup2", "loma_stats_config", "loma_stats_accumulate", "loma_stats_read",
    "loma_stats_reset",
]
KERNEL_CLASSES = ["phaseA", "zfft_inv", "xline", "zfft_fwd", "phaseC", "finish", "phaseC2"]

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class LomaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"loma error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(
                f"{LIB} is missing: build it first (python -m loma_phd_amd.build); "
                "there is no CPU fallback")
        L = ctypes.CDLL(LIB)
        i, d = ctypes.c_int, ctypes.c_double
        sig = {
            "loma_version": ([], ctypes.c_char_p),
            "loma_last_error": ([], ctypes.c_char_p),
            "loma_device_count": ([ctypes.POINTER(i)], i),
            "loma_fd_matrices": ([i, _dp, _dp, _dp, _dp, _dp, _dp], i),
            "loma_init": ([ctypes.POINTER(_vp), i, i, i, _dp, d, d, d, d, d, i, d, i, i], i),
            "loma_free": ([_vp], i),
            "loma_dims": ([_vp, ctypes.POINTER(ctypes.c_longlong)], i),
            "loma_upload": ([_vp, _vp, _vp, _vp, _vp, _dp, _dp, _dp, i], i),
            "loma_download": ([_vp, _vp, _vp, _vp, _vp, _dp, _dp, _dp, i], i),
            "loma_upload_z": ([_vp, _vp, i], i),
            "loma_seed": ([_vp, _vp, i, d, d, d, d], i),
            "loma_thickness": ([_vp, _dp], i),
            "loma_history": ([_vp, _dp], i),
            "loma_stats_config": ([_vp, i, ctypes.POINTER(i)], i),
            "loma_stats_accumulate": ([_vp, i], i),
            "loma_stats_read": ([_vp, _dp, _fp, _dp], i),
            "loma_stats_reset": ([_vp], i),
            "loma_snapshot": ([_vp], i),
            "loma_restore": ([_vp], i),
            "loma_set_cfl": ([_vp, ctypes.c_double], i),
            "loma_download_z": ([_vp, _vp, i], i),
            "loma_advance": ([_vp, i], i),
            "loma_substep": ([_vp, i], i),
            "loma_scalars": ([_vp, _dp], i),
            "loma_cross1": ([_vp, _fp, _fp, _fp, _fp, _dp, _dp, _dp, i], i),
            "loma_set_stream": ([_vp, _vp], i),
            "loma_use_graph": ([_vp, i], i),
            "loma_deryr2_dev": ([_vp, _vp, _vp, i], i),
            "loma_deryyr2_dev": ([_vp, _vp, _vp, i], i),
            "loma_lapvdvhom_dev": ([_vp, _vp, _vp, _vp, _vp, i], i),
            "loma_solvepsi_dev": ([_vp, _vp, _vp, _vp, _vp, i], i),
            "loma_helmholtz_dev": ([_vp, _vp, _vp, _vp, d, i], i),
            "loma_inty8_dev": ([_vp, _vp, _vp], i),
            "loma_visc_dev": ([_vp] * 9 + [d, d, d, i], i),
            "loma_visc": ([_vp] + [_fp] * 6 + [d] * 5, i),
            "loma_set_physics": ([_vp, d, d, d, d], i),
            "loma_deryr2": ([_vp, _fp, _fp, i], i),
            "loma_deryyr2": ([_vp, _fp, _fp, i], i),
            "loma_lapvdvhom": ([_vp, _fp, _fp, _fp, d], i),
            "loma_solvepsi": ([_vp, _fp, _fp, _fp, d], i),
            "loma_fourz_dev": ([_vp, _vp, i, i], i),
            "loma_band5_solve_dev": ([_vp, _vp, _vp, i, i, i, _vp], i),
            "loma_band5_solve_dev_noalloc": ([_vp, _vp, _vp, i, i, i, _vp], i),
            "loma_advance_many": ([ctypes.POINTER(_vp), i, i], i),
            "loma_tiles": ([_vp, ctypes.POINTER(i), ctypes.POINTER(i)], i),
            "loma_footprint": ([i, i, i, i, i, ctypes.POINTER(ctypes.c_longlong)], i),
            "loma_step_begin": ([_vp], i),
            "loma_step_poll": ([_vp, ctypes.POINTER(i), _dp, ctypes.POINTER(i)], i),
            "loma_advance_async": ([_vp, i], i),
            "loma_sync": ([_vp], i),
            "loma_stream": ([_vp, ctypes.POINTER(_vp)], i),
            "loma_profile": ([_vp, i], i),
            "loma_profile_read": ([_vp, _dp, ctypes.POINTER(i), i], i),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(code: int):
    if code != LOMA_OK:
        msg = lib().loma_last_error().decode(errors="replace")
        raise LomaError(code, msg)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def fd_matrices(y):
    """Compact-FD bands of a y grid: (dt11, dt12, dt21, dt22, fmap), each band
    (my, nb) row-major (the reference's (nb,my) Fortran arrays)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    my = y.size
    out = [np.zeros((my, 7)), np.zeros((my, 7)), np.zeros((my, 5)), np.zeros((my, 5)),
           np.zeros(my)]
    check(lib().loma_fd_matrices(my, _d(y), *[_d(a) for a in out]))
    return tuple(out)


def band5_solve_dev(band_ptr, rhs_ptr, x_ptr, n, nsys, rhs_fp32=False, stream=None):
    """Batched pentadiagonal solves on device pointers (bandec/banbks,
    cfdiff.f:954-1022): band [nsys][n][5] float64, rhs [nsys][n] complex
    (complex64 when rhs_fp32), x [nsys][n] complex128 or None for in place."""
    check(lib().loma_band5_solve_dev(band_ptr, rhs_ptr, x_ptr, n, nsys, 1 if rhs_fp32 else 0,
                                     stream))


def band5_solve_dev_noalloc(band_ptr, rhs_ptr, x_ptr, n, nsys, rhs_fp32=False, stream=None):
    """band5_solve_dev without device allocation: the scratch is a fixed
    32 MiB array per GPU, so n*5*nsys <= 4194304; overlapping calls on one
    GPU race."""
    check(lib().loma_band5_solve_dev_noalloc(band_ptr, rhs_ptr, x_ptr, n, nsys,
                                             1 if rhs_fp32 else 0, stream))


def footprint(mgalx, my, mgalz, precision=LOMA_FP64, flags=0) -> int:
    """Device bytes a run of this grid allocates (host-only, no GPU needed)."""
    b = ctypes.c_longlong(0)
    check(lib().loma_footprint(mgalx, my, mgalz, precision, flags, ctypes.byref(b)))
    return int(b.value)


def device_count() -> int:
    n = ctypes.c_int(0)
    lib().loma_device_count(ctypes.byref(n))
    return n.value


class Run:
    """One device-resident mixing-layer run (the reference's cross1 state).

    Mirrors the reference's driver contract: initcr-style set-up at
    construction, getfil/escru-style upload/download, cross1 time loop as
    advance().  Arrays use the reference layout [kx plane][kz][y], complex.
    """

    def __init__(self, mgalx, my, mgalz, y, Re, alp, bet, pr=0.7, cfl=0.5, ncfl=1,
                 dt0=1e-4, precision=LOMA_FP64, flags=0):
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        assert self.y.size == my
        h = _vp()
        check(lib().loma_init(ctypes.byref(h), mgalx, my, mgalz, _d(self.y), Re, alp, bet, pr,
                              cfl, ncfl, dt0, precision, flags))
        self._h = h
        dims = (ctypes.c_longlong * 8)()
        check(lib().loma_dims(h, dims))
        self.mx, self.mz, self.nplanes, self.nz, self.S, self.W, self.P, self.bytes = list(dims)
        self.mgalx, self.my, self.mgalz = mgalx, my, mgalz
        self.Re, self.alp, self.bet, self.pr = Re, alp, bet, pr
        self.precision, self.flags = precision, flags

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.loma_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def shape(self):
        return (self.nplanes, self.mz, self.my)

    def upload(self, vor, phi, psi, scal, u00, v00, w00, z=None):
        f = [np.ascontiguousarray(a, dtype=np.complex128).reshape(self.shape())
             for a in (vor, phi, psi, scal)]
        o = [np.ascontiguousarray(a, dtype=np.float64) for a in (u00, v00, w00)]
        check(lib().loma_upload(self._h, *[a.ctypes.data_as(_vp) for a in f],
                                *[_d(a) for a in o], 1))
        if z is not None:
            zz = np.ascontiguousarray(z, dtype=np.complex128).reshape(self.shape())
            check(lib().loma_upload_z(self._h, zz.ctypes.data_as(_vp), 1))

    def seed(self, coef, s, amp, env_width, k0):
        """Seeded initial condition built on the device (loma_seed) in place
        of an upload; coef: complex (3, nplanes, mz, nymax) draws."""
        c = np.ascontiguousarray(coef, dtype=np.complex128)
        if c.ndim != 4 or c.shape[:3] != (3,) + self.shape()[:2]:
            raise ValueError(f"coef shape {c.shape} is not (3, nplanes, mz, nymax)")
        check(lib().loma_seed(self._h, c.ctypes.data_as(_vp), c.shape[3], float(s),
                              float(amp), float(env_width), float(k0)))

    def history(self):
        """The .cf history columns that need no statistics window
        (cross.f:1166-1177): time, dm, dw, Redw, trapz, massu, massv, massw
        of the last step (dm, dw, Redw NaN without LOMA_FLAG_HISTORY)."""
        o = np.zeros(8)
        check(lib().loma_history(self._h, o.ctypes.data_as(_dp)))
        return dict(zip(("time", "dm", "dw", "Redw", "trapz", "massu", "massv", "massw"), o))

    def stats_config(self, jspecy):
        """Open a statistics window: spectra rows jspecy (1-based y indices,
        the reference's ctes3D list; the centre row (my+1)/2 is appended) and
        zeroed accumulators (loma_stats_config)."""
        rows = np.ascontiguousarray(jspecy, dtype=np.int32)
        self._nsp = int(rows.size) + 1
        check(lib().loma_stats_config(self._h, int(rows.size),
                                      rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def stats_accumulate(self, on=True):
        """istati: rkstep 1 of the following steps adds to the window."""
        check(lib().loma_stats_accumulate(self._h, 1 if on else 0))

    def stats_read(self):
        """writestats' outputs: sta (my, 33) .sta record rows, spe (nplanes,
        12, nspec+1, nz+1) float32 spectra, and nacum, nacumsp, masst,
        enerdis, energy(9)."""
        sta = np.zeros((self.my, 33))
        spe = np.zeros((self.nplanes, 12, self._nsp, self.nz + 1), np.float32)
        info = np.zeros(13)
        check(lib().loma_stats_read(self._h, sta.ctypes.data_as(_dp), spe.ctypes.data_as(_fp),
                                    info.ctypes.data_as(_dp)))
        return dict(sta=sta, spe=spe, nacum=int(info[0]), nacumsp=int(info[1]),
                    masst=info[2], enerdis=info[3], energy=info[4:13].copy())

    def stats_reset(self):
        """resetstats: zero the window's accumulators and counters."""
        check(lib().loma_stats_reset(self._h))

    def snapshot(self):
        """Device checkpoint of the time-advance state (loma_snapshot)."""
        check(lib().loma_snapshot(self._h))

    def restore(self):
        """Return to the last device checkpoint (loma_restore)."""
        check(lib().loma_restore(self._h))

    def set_cfl(self, cfl):
        """Change the CFL number of the run, keeping its state (loma_set_cfl)."""
        check(lib().loma_set_cfl(self._h, float(cfl)))

    def thickness(self):
        """(delta_m, delta_w) of a LOMA_FLAG_HISTORY run (loma_thickness)."""
        out = np.zeros(2)
        check(lib().loma_thickness(self._h, _d(out)))
        return float(out[0]), float(out[1])

    def download(self):
        f = [np.zeros(self.shape(), np.complex128) for _ in range(4)]
        o = [np.zeros(self.my) for _ in range(3)]
        check(lib().loma_download(self._h, *[a.ctypes.data_as(_vp) for a in f],
                                  *[_d(a) for a in o], 1))
        out = dict(vor=f[0], phi=f[1], psi=f[2], scal=f[3], u00=o[0], v00=o[1], w00=o[2])
        if self.flags & LOMA_FLAG_SCALAR_Z:
            z = np.zeros(self.shape(), np.complex128)
            check(lib().loma_download_z(self._h, z.ctypes.data_as(_vp), 1))
            out["z"] = z
        return out

    def advance(self, nsteps=1):
        check(lib().loma_advance(self._h, nsteps))

    def advance_async(self, nsteps=1):
        check(lib().loma_advance_async(self._h, nsteps))

    def sync(self):
        check(lib().loma_sync(self._h))

    def stream(self) -> int:
        p = _vp()
        check(lib().loma_stream(self._h, ctypes.byref(p)))
        return p.value or 0

    def profile(self, on=True):
        check(lib().loma_profile(self._h, 1 if on else 0))

    def profile_read(self):
        n = len(KERNEL_CLASSES)
        ms = np.zeros(n)
        cnt = (ctypes.c_int * n)()
        check(lib().loma_profile_read(self._h, _d(ms), cnt, n))
        return {k: (float(ms[q]), int(cnt[q])) for q, k in enumerate(KERNEL_CLASSES)}

    def step_begin(self):
        """Enqueue one RK3 step and return at once (one step in flight)."""
        check(lib().loma_step_begin(self._h))

    def step_poll(self):
        """(done, device_ms, nonfinite) of the step in flight, without
        blocking."""
        done, nf = ctypes.c_int(0), ctypes.c_int(0)
        ms = ctypes.c_double(0.0)
        check(lib().loma_step_poll(self._h, ctypes.byref(done), ctypes.byref(ms),
                                   ctypes.byref(nf)))
        return bool(done.value), float(ms.value), bool(nf.value)

    def substep(self, rk):
        check(lib().loma_substep(self._h, rk))

    def scalars(self):
        o = np.zeros(5)
        check(lib().loma_scalars(self._h, _d(o)))
        return dict(dt=o[0], time=o[1], dtr=o[2], istep=int(o[3]), nan=int(o[4]))

    def set_physics(self, Re, pr=0.7, cfl=0.5, dt0=1e-4):
        """New Re, Pr, CFL and first dt on the same grid and buffers; upload a
        state before advancing."""
        check(lib().loma_set_physics(self._h, float(Re), float(pr), float(cfl), float(dt0)))
        self.Re, self.pr = Re, pr

    def use_graph(self, on=True):
        check(lib().loma_use_graph(self._h, 1 if on else 0))

    def set_stream(self, stream_ptr: int):
        check(lib().loma_set_stream(self._h, ctypes.c_void_p(stream_ptr)))

    # per-operator entry points on device pointers (ints)
    def deryr2_dev(self, u_ptr, du_ptr, ncols):
        check(lib().loma_deryr2_dev(self._h, u_ptr, du_ptr, ncols))

    def deryyr2_dev(self, u_ptr, du_ptr, ncols):
        check(lib().loma_deryyr2_dev(self._h, u_ptr, du_ptr, ncols))

    def lapvdvhom_dev(self, phi, v, dv, rK, ncols):
        check(lib().loma_lapvdvhom_dev(self._h, phi, v, dv, rK, ncols))

    def solvepsi_dev(self, phi, v, dv, rK, ncols):
        check(lib().loma_solvepsi_dev(self._h, phi, v, dv, rK, ncols))

    def helmholtz_dev(self, f, x, rK, shift, ncols):
        check(lib().loma_helmholtz_dev(self._h, f, x, rK, float(shift), ncols))

    def visc_dev(self, vorwk, vor, phi, v, dvdy, phiwk, rk1, rk2, dalbe, dtri, dtxi, ncols):
        """visc (trunk/mixing/src/laps.f:535) on ncols device columns; rk1, rk2
        device double arrays."""
        check(lib().loma_visc_dev(self._h, vorwk, vor, phi, v, dvdy, phiwk, rk1, rk2,
                                  float(dalbe), float(dtri), float(dtxi), ncols))

    def visc(self, vorwk, v, dvdy, phiwk, rk1, rk2, dalbe, dtri, dtxi):
        """visc with the reference's real(4) (2,my) columns on the host;
        returns dict(vorwk, vor, phi, v, dvdy, phiwk) as complex64."""
        a = {k: np.ascontiguousarray(x, dtype=np.complex64).copy() for k, x in
             (("vorwk", vorwk), ("v", v), ("dvdy", dvdy), ("phiwk", phiwk))}
        vor = np.zeros_like(a["vorwk"])
        phi = np.zeros_like(a["vorwk"])
        fp = lambda x: x.view(np.float32).ctypes.data_as(_fp)
        check(lib().loma_visc(self._h, fp(a["vorwk"]), fp(vor), fp(phi), fp(a["v"]),
                              fp(a["dvdy"]), fp(a["phiwk"]), float(rk1), float(rk2),
                              float(dalbe), float(dtri), float(dtxi)))
        return dict(vorwk=a["vorwk"], vor=vor, phi=phi, v=a["v"], dvdy=a["dvdy"],
                    phiwk=a["phiwk"])

    def inty8_dev(self, du, u):
        check(lib().loma_inty8_dev(self._h, du, u))

    def tiles(self):
        """(J, L): y entries per z-transform workgroup, lines per x-line
        workgroup."""
        J, L = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().loma_tiles(self._h, ctypes.byref(J), ctypes.byref(L)))
        return J.value, L.value

    def fourz_dev(self, c_ptr, nrows, iopt):
        check(lib().loma_fourz_dev(self._h, c_ptr, nrows, iopt))

    # host float versions with the reference's shapes
    def deryr2(self, u):
        u = np.ascontiguousarray(u, dtype=np.complex64)
        du = np.zeros_like(u)
        m = u.size // self.my
        check(lib().loma_deryr2(self._h, u.view(np.float32).ctypes.data_as(_fp),
                                du.view(np.float32).ctypes.data_as(_fp), m))
        return du

    def deryyr2(self, u):
        u = np.ascontiguousarray(u, dtype=np.complex64)
        du = np.zeros_like(u)
        m = u.size // self.my
        check(lib().loma_deryyr2(self._h, u.view(np.float32).ctypes.data_as(_fp),
                                 du.view(np.float32).ctypes.data_as(_fp), m))
        return du


def advance_many(runs, nsteps):
    arr = (_vp * len(runs))(*[r.handle for r in runs])
    check(lib().loma_advance_many(arr, len(runs), nsteps))
