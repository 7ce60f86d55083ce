"""The Fortran binding of INTEGRATION.md compiles and calls the library.

The module block of INTEGRATION.md (`module loma_amd`, ISO_C_BINDING
interfaces of the C ABI) is extracted as it stands and compiled with flang,
together with a small driver that calls two host-only entry points through
it: loma_fd_matrices (the compact-FD set-up, derivadas cfdiff.f:869) and
loma_footprint.  The FD matrices that come back through Fortran equal the
ones the Python binding returns, and the footprint equals capi.footprint.
Skipped where flang is absent (the GPU box does not run CPU tests)."""
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/llvm/bin/flang"

DRIVER = """
program call_loma
  use iso_c_binding
  use loma_amd
  implicit none
  integer, parameter :: my = 65
  real(c_double) :: y(my), d11(7*my), d12(7*my), d21(5*my), d22(5*my), fm(my)
  integer(c_long_long) :: bytes
  integer(c_int) :: ierr
  integer :: j
  do j = 1, my
    y(j) = -10d0 + 20d0*dble(j-1)/dble(my-1) + 0.3d0*sin(dble(j))
  end do
  ierr = loma_fd_matrices(my, y, d11, d12, d21, d22, fm)
  write(*,'(a,i3)') 'fd ', ierr
  write(*,'(es26.17e3)') y, d11, d12, d21, d22, fm
  ierr = loma_footprint(48, 97, 36, LOMA_MIXED, LOMA_FLAG_SCALAR_Z, bytes)
  write(*,'(a,i3,i20)') 'footprint ', ierr, bytes
end program
"""


def _module_source():
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    blocks = re.findall(r"```fortran\n(.*?)```", text, re.S)
    mods = [b for b in blocks if re.search(r"^module loma_amd\b", b, re.M)]
    assert len(mods) == 1, "INTEGRATION.md must hold exactly one module loma_amd block"
    return mods[0]


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_integration_module_compiles_and_calls(tmp_path):
    from loma_phd_amd import capi
    lib = capi.LIB
    assert os.path.exists(lib)
    (tmp_path / "loma_amd_mod.f90").write_text(_module_source())
    (tmp_path / "call_loma.f90").write_text(DRIVER)
    libdir = os.path.dirname(lib)
    subprocess.run([FLANG, "-c", "loma_amd_mod.f90"], cwd=tmp_path, check=True)
    subprocess.run([FLANG, "-o", "call_loma", "call_loma.f90", "loma_amd_mod.o",
                    f"-L{libdir}", "-lloma_hip", f"-Wl,-rpath,{libdir}"],
                   cwd=tmp_path, check=True)
    r = subprocess.run([str(tmp_path / "call_loma")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    assert lines[0].split() == ["fd", "0"]
    my = 65
    vals = np.array([float(v) for v in lines[1:1 + my * 26]])
    y = vals[:my]
    got = np.split(vals[my:], np.cumsum([7 * my, 7 * my, 5 * my, 5 * my]))
    want = capi.fd_matrices(y)
    for g, w in zip(got, want):
        w = np.asarray(w).ravel()
        assert np.allclose(g, w, rtol=1e-15, atol=1e-15 * np.abs(w).max()), "fd matrices"
    fp = [ln for ln in lines if ln.startswith("footprint")][0].split()
    assert fp[1] == "0"
    assert int(fp[2]) == capi.footprint(48, 97, 36, capi.LOMA_MIXED, capi.LOMA_FLAG_SCALAR_Z)
