"""The C ABI as a drop-in for the reference's own subroutines, called with
the reference's argument shapes (real(4) columns, complex*8 planes, the
cross1 argument order) through ctypes.

* loma_cross1 (cross.f:24) from the FP32 restart the reference solver was
  run from (tests/golden/ref_cross1_192x453x96.npz, written by
  tests/golden/make_ref_cross1.py): after one step the GPU's state matches
  the reference's own image at the reference's FP32 floor, with the bars of
  the oracle's pin (tests/test_oracle_cross1.py).  Only the first step is
  compared: from the second step on the reference's time step carries its
  8-rank CFL reduction, the library the one-rank one.
* loma_deryr2, loma_deryyr2, loma_lapvdvhom, loma_solvepsi (float columns)
  and loma_deryr, loma_deryyr, loma_inty8 (double columns) on the
  reference's own mesh meshLy172my453 against the outputs of the
  reference's routines (tests/golden/ref_operators_my453.npz).
* loma_cft, loma_fourz, loma_rft, loma_fourx against the oracle's FFTPACK
  restatement, itself pinned to the reference's FFT test vectors.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)


def _f(a):
    return a.ctypes.data_as(_fp)


def _d(a):
    return a.ctypes.data_as(_dp)


def _bind(L, name, args):
    fn = getattr(L, name)
    fn.argtypes = args
    fn.restype = ctypes.c_int
    return fn


def test_cross1_one_step_matches_reference(gpu):
    from loma_phd_amd import capi, grid, ic
    with np.load(os.path.join(GOLD, "ref_cross1_192x453x96.npz"), allow_pickle=False) as z:
        ref = {k: z[k] for k in z.files}
    g = [int(v) for v in ref["grid"]]
    Re, s, alp, bet, pr, cfl, amp, seed, dt0 = [float(v) for v in ref["physics"]]
    cfg = grid.MixingLayerConfig(mgalx=g[0], my=g[1], mgalz=g[2], Re=Re, s=s, amp=amp,
                                 seed=int(seed), cfl=cfl, dt0=dt0, pr=pr)
    print("cross1 drop-in at", g, flush=True)
    st = ic.initial_condition(cfg)
    # the reference's complex*8 fields, (2, my, mz, nplanes) in Fortran order
    f32 = {f: np.ascontiguousarray(np.asarray(st[f]).astype(np.complex64))
           for f in ("vor", "phi", "psi", "scal")}
    m00 = {f: np.asarray(st[f]).astype(np.float32).astype(np.float64)
           for f in ("u00", "w00", "v00")}
    start = {f: f32[f].astype(np.complex128) for f in f32}
    run = capi.Run(cfg.mgalx, cfg.my, cfg.mgalz, st["y"], cfg.Re, cfg.alp, cfg.bet, cfg.pr,
                   cfg.cfl, cfg.ncfl, cfg.dt0, capi.LOMA_FP64,
                   capi.LOMA_FLAG_DT_FP32 | capi.LOMA_FLAG_HISTORY)
    L = capi.lib()
    cross1 = _bind(L, "loma_cross1", [ctypes.c_void_p] + [_fp] * 4 + [_dp] * 3 + [ctypes.c_int])
    fv = [f32[f].view(np.float32) for f in ("vor", "phi", "psi", "scal")]
    rc = cross1(run._h, *[_f(a) for a in fv], _d(m00["u00"]), _d(m00["w00"]), _d(m00["v00"]), 1)
    assert rc == 0, L.loma_last_error()
    t = run.scalars()["time"]
    hist = run.history()
    run.close()
    assert abs(t - float(ref["time_1"])) <= 1e-6 * float(ref["time_1"])
    # the .cf history line of the step (cross.f:1166-1177); the reference
    # sums it in FP32 from its FP32 state
    cf = ref["cf"][0]
    for name, col, tol in (("dm", 1, 1e-6), ("dw", 2, 1e-5), ("trapz", 4, 5e-8),
                           ("massu", 16, 2e-5), ("massv", 17, 1e-6), ("massw", 18, 1e-12)):
        err = abs(hist[name] - cf[col])
        print(f"  .cf {name}: {hist[name]:.9e} vs {cf[col]:.9e}", flush=True)
        assert err <= tol, (name, hist[name], cf[col])
    for f, key in (("u00", "u00_1"), ("v00", "v00_1"), ("w00", "w00_1")):
        err = np.abs(m00[f] - ref[key]).max()
        print(f"  {f}: max abs err {err:.2e}", flush=True)
        assert err <= 1e-6, (f, err)
    for f in ("vor", "phi", "psi", "scal"):
        m = ref[f"modes_{f}"]
        got = f32[f].astype(np.complex128)
        cr = ref[f"cols_{f}_1"].astype(np.complex128)
        ca = got[m[:, 0], m[:, 1]]
        c0 = start[f][m[:, 0], m[:, 1]]
        e_state = np.linalg.norm(ca - cr) / np.linalg.norm(cr)
        e_inc = np.linalg.norm(ca - cr) / np.linalg.norm(cr - c0)
        nr = ref[f"norms_{f}_1"].astype(np.float64)
        na = np.sqrt(np.sum(np.abs(got) ** 2, axis=2))
        e_norm = np.linalg.norm(na - nr) / np.linalg.norm(nr)
        print(f"  {f}: state {e_state:.2e} increment {e_inc:.2e} norms {e_norm:.2e}", flush=True)
        assert e_state <= 2e-5 and e_norm <= 1e-3, (f, e_state, e_norm)
        if f != "scal":
            assert e_inc <= 1e-2, (f, e_inc)


@pytest.fixture(scope="module")
def run453(gpu):
    from loma_phd_amd import capi
    with np.load(os.path.join(GOLD, "ref_operators_my453.npz"), allow_pickle=False) as z:
        ref = {k: z[k] for k in z.files}
    run = capi.Run(192, ref["y"].size, 96, ref["y"], 160.0, 0.5, 0.5)
    yield run, ref
    run.close()


def test_float_column_ops_match_reference(run453):
    from loma_phd_amd import capi
    run, ref = run453
    L = capi.lib()
    u = np.ascontiguousarray(ref["u"])
    m = u.shape[0]
    for name in ("deryr2", "deryyr2"):
        fn = _bind(L, "loma_" + name, [ctypes.c_void_p, _fp, _fp, ctypes.c_int])
        out = np.zeros_like(u)
        assert fn(run._h, _f(u.view(np.float32)), _f(out.view(np.float32)), m) == 0
        e = rel_l2(out, ref[name])
        print(name, f"{e:.2e}", flush=True)
        assert e < 2e-7, (name, e)
    for name in ("lapvdvhom", "solvepsi"):
        fn = _bind(L, "loma_" + name, [ctypes.c_void_p, _fp, _fp, _fp, ctypes.c_double])
        for c, rk in enumerate(ref["rK"]):
            phi = np.ascontiguousarray(u[0])
            v = np.zeros_like(phi)
            dv = np.zeros_like(phi)
            assert fn(run._h, _f(phi.view(np.float32)), _f(v.view(np.float32)),
                      _f(dv.view(np.float32)), float(rk)) == 0
            for what, got in (("v", v), ("dv", dv)):
                want = ref[f"{name}_{what}"][c]
                e = rel_l2(got, want) if np.linalg.norm(want) > 0 else np.abs(got).max()
                print(name, what, rk, f"{e:.2e}", flush=True)
                assert e < 2e-6, (name, what, rk, e)


def test_double_column_ops_match_reference(run453):
    from loma_phd_amd import capi
    run, ref = run453
    L = capi.lib()
    f = np.ascontiguousarray(ref["f"])
    for name, tol in (("deryr", 1e-10), ("deryyr", 1e-10), ("inty8", 2e-6)):
        fn = _bind(L, "loma_" + name, [ctypes.c_void_p, _dp, _dp])
        out = np.zeros_like(f)
        assert fn(run._h, _d(f), _d(out)) == 0
        e = rel_l2(out, ref[name])
        print(name, f"{e:.2e}", flush=True)
        # inty8: the library solves the well-posed (gauge-row) form; the
        # reference's singular form carries its own 1e-7-level noise
        assert e < tol, (name, e)


def test_transforms_match_fftpack_restatement(run453, oracle_mod):
    from loma_phd_amd import capi
    run, ref = run453
    L = capi.lib()
    O = oracle_mod
    rng = np.random.default_rng(3)
    mgalx, my, mgalz = 192, ref["y"].size, 96
    # cft / fourz: a plane of my rows of mgalz complex numbers
    plane = (rng.standard_normal((my, mgalz)) + 1j * rng.standard_normal((my, mgalz))
             ).astype(np.complex64)
    fourz = _bind(L, "loma_fourz", [ctypes.c_void_p, _fp, ctypes.c_int])
    for iopt in (-1, 1):
        got = plane.copy()
        assert fourz(run._h, _f(got.view(np.float32)), iopt) == 0
        want = O.cft(plane.astype(np.complex128), iopt)
        e = rel_l2(got, want)
        print("fourz", iopt, f"{e:.2e}", flush=True)
        assert e < 3e-7
    # cft with strides: 5 vectors, element stride 6 floats, vector stride 1000
    cft = _bind(L, "loma_cft", [ctypes.c_void_p, _fp, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int])
    buf = rng.standard_normal(5000).astype(np.float32)
    got = buf.copy()
    assert cft(run._h, _f(got), 6, 1000, 5, -1) == 0
    idx = np.arange(5)[:, None] * 1000 + np.arange(mgalz)[None, :] * 6
    vec = buf[idx] + 1j * buf[idx + 1]
    want = O.cft(vec.astype(np.complex128), -1)
    assert rel_l2(got[idx] + 1j * got[idx + 1], want) < 3e-7
    untouched = np.ones(5000, bool)
    untouched[idx] = False
    untouched[idx + 1] = False
    assert np.array_equal(got[untouched], buf[untouched])
    # rft: 3 real vectors of mgalx points at stride mgalx + 2
    rft = _bind(L, "loma_rft", [ctypes.c_void_p, _fp, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int])
    x = np.zeros((3, mgalx + 2), np.float32)
    x[:, :mgalx] = rng.standard_normal((3, mgalx))
    got = x.copy()
    assert rft(run._h, _f(got), mgalx + 2, 3, -1) == 0
    want = O.rft(x.astype(np.float64), mgalx, -1)
    assert rel_l2(got, want) < 3e-7
    back = got.copy()
    assert rft(run._h, _f(back), mgalx + 2, 3, 1) == 0
    assert rel_l2(back[:, :mgalx], x[:, :mgalx]) < 3e-7
    # fourx: one line, inverse to FP64 physical values, then forward again
    fourx = _bind(L, "loma_fourx", [ctypes.c_void_p, _fp, _dp, ctypes.c_int])
    line = np.zeros(mgalx + 2, np.float32)
    mx = 2 * (mgalx // 3)
    line[:mx] = rng.standard_normal(mx)
    line[1] = 0.0
    p8 = np.zeros(mgalx + 2)
    ph = line.copy()
    assert fourx(run._h, _f(ph), _d(p8), 1) == 0
    want = O.rft(np.where(np.arange(mgalx + 2) < mx, line, 0.0)[None, :].astype(np.float64),
                 mgalx, 1)[0]
    assert rel_l2(p8[:mgalx], want[:mgalx]) < 1e-12
    assert assert_close(ph[:mgalx], p8[:mgalx])
    fwd = ph.copy()
    assert fourx(run._h, _f(fwd), _d(p8), -1) == 0
    assert rel_l2(fwd[:mx], line[:mx]) < 3e-7


def assert_close(a, b):
    return np.allclose(a, b, rtol=1e-6, atol=1e-6 * np.abs(b).max())
