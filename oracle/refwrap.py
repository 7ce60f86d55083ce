"""ctypes wrapper of oracle/_ref/libloma_ref.so: the reference's own Fortran
routines (trunk/lomacte3/src) built in place by oracle/Makefile.ref.

TEST INFRASTRUCTURE ONLY.  The library exists only where the reference tree
was present at build time; tests that use it skip otherwise.  Its grid is the
reference's compile-time one (ctes3D:19-21: mgalx=192, my=453, mgalz=96).
"""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libloma_ref.so")
REFSRC = "/root/reference/trunk/lomacte3/src"
MESH453 = "/root/reference/trunk/initloma3/data/meshLy172my453.txt"
_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def build() -> bool:
    """Build the reference library when the reference tree and flang exist."""
    if not os.path.isdir(REFSRC) or not os.path.exists("/opt/rocm/llvm/bin/flang"):
        return False
    repo = os.path.dirname(HERE)
    subprocess.check_call(["make", "-s", "-f", "oracle/Makefile.ref"], cwd=repo)
    return os.path.exists(LIB)


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB)
        L.ref_dims.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.ref_setup.argtypes = [_dp]
        L.ref_matrices.argtypes = [_dp] * 8
        L.ref_deryr2.argtypes = [_fp, _fp, ctypes.c_int]
        L.ref_deryyr2.argtypes = [_fp, _fp, ctypes.c_int]
        L.ref_deryr.argtypes = [_dp, _dp]
        L.ref_deryyr.argtypes = [_dp, _dp]
        L.ref_inty8.argtypes = [_dp, _dp]
        L.ref_lapvdvhom.argtypes = [_fp, _fp, _fp, ctypes.c_double]
        L.ref_solvepsi.argtypes = [_fp, _fp, _fp, ctypes.c_double]
        L.ref_bandec.argtypes = [_dp, ctypes.c_int]
        L.ref_banbks.argtypes = [_dp, ctypes.c_int, _dp]
        L.ref_bandec7.argtypes = [_dp, ctypes.c_int]
        L.ref_banbks7.argtypes = [_dp, ctypes.c_int, _dp]
        L.ref_fft_init.argtypes = []
        L.ref_fourz.argtypes = [_fp, ctypes.c_int]
        L.ref_fourx.argtypes = [_fp, _dp, ctypes.c_int]
        L.ref_localyz2zy.argtypes = [_fp, _fp]
        L.ref_localzy2yz.argtypes = [_fp, _fp]
        _lib = L
    return _lib


def dims():
    d = (ctypes.c_int * 3)()
    lib().ref_dims(d)
    return tuple(d)


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def setup(y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    lib().ref_setup(_d(y))


def matrices():
    my = dims()[1]
    p1, t11, t12 = (np.zeros((my, 7)) for _ in range(3))
    p3, t21, t22 = (np.zeros((my, 5)) for _ in range(3))
    fm, ed = np.zeros(my), np.zeros(my)
    lib().ref_matrices(*[_d(a) for a in (p1, t11, t12, p3, t21, t22, fm, ed)])
    return dict(prem1=p1, dt11=t11, dt12=t12, prem3=p3, dt21=t21, dt22=t22, fmap=fm,
                endder=ed)


def colop(name, u):
    """u: (m, my) complex -> complex64 columns through the FP32 interface"""
    u = np.ascontiguousarray(u, dtype=np.complex64)
    out = np.zeros_like(u)
    getattr(lib(), name)(_f(u.view(np.float32)), _f(out.view(np.float32)), u.shape[0])
    return out


def realop(name, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    o = np.zeros_like(u)
    getattr(lib(), name)(_d(u), _d(o))
    return o


def solve(name, phi, rK):
    phi = np.ascontiguousarray(phi, dtype=np.complex64)
    v = np.zeros_like(phi); dv = np.zeros_like(phi)
    getattr(lib(), name)(_f(phi.view(np.float32)), _f(v.view(np.float32)),
                         _f(dv.view(np.float32)), float(rK))
    return v, dv


# ---- the reference's visc (trunk/mixing/src/laps.f:535), its own library ----
MIXLIB = os.path.join(HERE, "_ref", "libloma_ref_mix.so")
_mix = None


def mix_available() -> bool:
    return os.path.exists(MIXLIB)


def mix_lib():
    global _mix
    if _mix is None:
        L = ctypes.CDLL(MIXLIB)
        L.ref_mix_dims.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.ref_mix_set.argtypes = [_dp] * 5
        L.ref_mix_visc.argtypes = [_fp] * 6 + [ctypes.c_double] * 5
        _mix = L
    return _mix


def mix_my() -> int:
    d = (ctypes.c_int * 1)()
    mix_lib().ref_mix_dims(d)
    return d[0]


def mix_set(prem1, dt12, dt21, dt22, fmap):
    """Matrices in (my, nb) row-major form (the reference's (nb, my))."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (prem1, dt12, dt21, dt22, fmap)]
    mix_lib().ref_mix_set(*[_d(a) for a in arrs])


def mix_visc(vorwk, v, dvdy, phiwk, rk1, rk2, dalbe, dtri, dtxi):
    """visc on complex64 columns (my,); returns dict of the six outputs."""
    a = {k: np.ascontiguousarray(x, dtype=np.complex64).copy()
         for k, x in (("vorwk", vorwk), ("v", v), ("dvdy", dvdy), ("phiwk", phiwk))}
    vor = np.zeros_like(a["vorwk"])
    phi = np.zeros_like(a["vorwk"])
    mix_lib().ref_mix_visc(_f(a["vorwk"].view(np.float32)), _f(vor.view(np.float32)),
                           _f(phi.view(np.float32)), _f(a["v"].view(np.float32)),
                           _f(a["dvdy"].view(np.float32)), _f(a["phiwk"].view(np.float32)),
                           float(rk1), float(rk2), float(dalbe), float(dtri), float(dtxi))
    return dict(vorwk=a["vorwk"], vor=vor, phi=phi, v=a["v"], dvdy=a["dvdy"],
                phiwk=a["phiwk"])
