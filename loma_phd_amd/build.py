"""Build libloma_hip.so (HIP kernels for gfx950 + the C ABI) in-tree.

The library lands in loma_phd_amd/_lib/libloma_hip.so; object files go to
loma_phd_amd/_build/.  Every translation unit is compiled in parallel with
hipcc --offload-arch=gfx950.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "libloma_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LOMA_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions are disabled for all device code.  Multiply-adds
# are fused only where one source expression holds both (-ffp-contract=on,
# not the default "fast", under which the backend fuses across statements
# depending on the surrounding code): a kernel body compiled into the
# single-run and the batched kernel then rounds identically in both.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-ffp-contract=on", "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17"]


def _sources():
    hip = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    cpp = sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    return hip, cpp


def _compile(src: str, force: bool) -> str:
    base = os.path.splitext(os.path.basename(src))[0]
    obj = os.path.join(OBJ, base + ".o")
    headers = glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(CSRC, "*.inc")) + [
        os.path.join(os.path.dirname(PKG), "include", "loma_amd.h")]
    newest = max([os.path.getmtime(src)] + [os.path.getmtime(h) for h in headers])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return obj
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    else:
        cmd = [HIPCC] + CXX_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
    return obj


def build(force: bool = True, jobs: int | None = None, verbose: bool = False) -> str:
    """Compile every source for gfx950 and link libloma_hip.so."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hip, cpp = _sources()
    srcs = hip + cpp
    jobs = jobs or min(16, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
