"""Writes the golden fixtures used by the CPU pin tests.

1. fftpack_fixtures.npz -- the reference's own FFT test vectors
   (branches/reactml/benchmark_fft/{cfttest,rffttest}_{single,double}.txt,
   written by benchmark_fft/main.f:118-178 with format 5E14.7): input column
   and FFTPACK-convention output of cft (complex, n=384) and rft (real,
   n=768).
2. ref_operators_my453.npz -- outputs of the reference's own Fortran routines
   (trunk/lomacte3/src/cfdiff.f, laps.f built in place by oracle/Makefile.ref)
   on the reference mesh trunk/initloma3/data/meshLy172my453.txt for seeded
   inputs: FD matrices (derivadas), deryr2/deryyr2/Lapvdvhom/solvepsi on
   float32 columns, inty8/deryr/deryyr on float64 columns.

Run in the container that holds the reference tree:
    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, REPO)


def _read_fixed(path, width=14):
    vals = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            for p in range(0, len(line), width):
                t = line[p:p + width].strip()
                if t:
                    vals.append(float(t))
    return np.array(vals)


def fft_fixtures():
    base = os.path.join(REF, "branches/reactml/benchmark_fft")
    out = {}
    for prec in ("single", "double"):
        c = _read_fixed(os.path.join(base, f"cfttest_{prec}.txt")).reshape(-1, 3)
        out[f"cft_{prec}_in"] = c[:, 0]
        out[f"cft_{prec}_out"] = c[:, 1] + 1j * c[:, 2]
        r = _read_fixed(os.path.join(base, f"rffttest_{prec}.txt")).reshape(-1, 2)
        out[f"rft_{prec}_in"] = r[:, 0]
        out[f"rft_{prec}_out"] = r[:, 1]
    np.savez(os.path.join(HERE, "fftpack_fixtures.npz"), **out)


def ref_operator_fixtures():
    from oracle import refwrap as R
    R.build()
    y = np.loadtxt(R.MESH453)
    R.setup(y)
    M = R.matrices()
    rng = np.random.default_rng(20241202)
    my = y.size
    u = (rng.standard_normal((6, my)) + 1j * rng.standard_normal((6, my))).astype(np.complex64)
    f = np.exp(-(y / 12.0) ** 2) * (1.0 + 0.3 * np.sin(y / 5.0))
    out = dict(y=y, u=u, f=f, rK=np.array([0.0, 0.04, 1.7, 250.0]))
    out.update({k: v for k, v in M.items()})
    out["deryr2"] = R.colop("ref_deryr2", u)
    out["deryyr2"] = R.colop("ref_deryyr2", u)
    for name in ("lapvdvhom", "solvepsi"):
        vs, dvs = [], []
        for rk in out["rK"]:
            v, dv = R.solve("ref_" + name, u[0], rk)
            vs.append(v)
            dvs.append(dv)
        out[name + "_v"] = np.array(vs)
        out[name + "_dv"] = np.array(dvs)
    out["inty8"] = R.realop("ref_inty8", f)
    out["deryr"] = R.realop("ref_deryr", f)
    out["deryyr"] = R.realop("ref_deryyr", f)
    np.savez(os.path.join(HERE, "ref_operators_my453.npz"), **out)


if __name__ == "__main__":
    fft_fixtures()
    ref_operator_fixtures()
    print("fixtures written to", HERE)
