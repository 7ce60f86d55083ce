"""Grids and workload configurations of the temporal mixing layer.

Mesh generator follows the reference's tanh-stretched 1-D mesh recipe
(trunk/initloma3/scripts/writefiles.py:1-28, create_mesh1D): a spacing
function built from a tanh step, integrated with the trapezoid rule and
scaled to [-Ly, Ly].  Here the fine spacing sits in the shear-layer centre,
as in the reference's production meshes (trunk/initloma3/data/
meshLy172my453.txt: dy = 0.2 at the centre, 0.91 at the edges).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


def tanh_mesh(my: int, Ly: float, ratio: float = 3.0, eta0: float = 0.45,
              width: float = 0.15) -> np.ndarray:
    """Symmetric y grid on [-Ly, Ly]: spacing ~1 in |eta| < eta0 rising
    smoothly (tanh) to `ratio` near the boundaries."""
    eta = np.linspace(-1.0, 1.0, my)
    dy = 1.0 + (ratio - 1.0) * 0.5 * (1.0 + np.tanh((np.abs(eta) - eta0) / width))
    # trapezoid cumulative integral
    y = np.concatenate([[0.0], np.cumsum(0.5 * (dy[1:] + dy[:-1]) * np.diff(eta))])
    y = 2.0 * y * Ly / y[-1] - Ly
    y = 0.5 * (y - y[::-1])  # exact antisymmetry
    return y


@dataclasses.dataclass
class MixingLayerConfig:
    """One run: grid, physics and initial-condition parameters."""
    mgalx: int = 64
    my: int = 129
    mgalz: int = 64
    Lx: float = 8 * math.pi      # streamwise box (delta_w0 = 1)
    Lz: float = 8 * math.pi
    Ly: float = 20.0             # y in [-Ly, Ly]
    Re: float = 160.0
    pr: float = 0.7              # Prandtl = Schmidt
    s: float = 2.0               # density ratio rho_bot / rho_top
    amp: float = 1e-3            # perturbation rms at the centre plane
    env_width: float = 2.0       # Gaussian envelope width of the perturbation
    seed: int = 0
    cfl: float = 0.5
    ncfl: int = 1
    dt0: float = 1e-4
    scalar_z: bool = True
    mesh_ratio: float = 3.0

    @property
    def alp(self) -> float:
        return 2 * math.pi / self.Lx

    @property
    def bet(self) -> float:
        return 2 * math.pi / self.Lz

    @property
    def mx(self):
        return 2 * (self.mgalx // 3)

    @property
    def nplanes(self):
        return self.mx // 2

    @property
    def mz(self):
        return 2 * (self.mgalz // 3) - 1

    def y(self) -> np.ndarray:
        return tanh_mesh(self.my, self.Ly, self.mesh_ratio)

    @property
    def points(self) -> int:
        return self.mgalx * self.my * self.mgalz

    def spectral_size(self) -> int:
        return self.nplanes * self.mz * self.my


# the five workloads (configs A-E)
CONFIG_A = MixingLayerConfig(mgalx=64, my=129, mgalz=64, Re=160.0, s=2.0)
CONFIG_B = MixingLayerConfig(mgalx=512, my=513, mgalz=256, Re=800.0, s=4.0)

# ensemble batch: 32 runs, grids drawn by seed from these, density ratios
# U[1,8], Re U[160,800], independent seeds, mixed precision
ENSEMBLE_GRIDS = [(128, 129, 64), (192, 193, 96), (256, 257, 128)]


def ensemble_configs(nruns: int = 32, seed: int = 7, grids=ENSEMBLE_GRIDS,
                     ratio=(1.0, 8.0), re=(160.0, 800.0)):
    """Seeded draw of the ensemble batch's run configurations."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(nruns):
        g = grids[int(rng.integers(len(grids)))]
        out.append(MixingLayerConfig(mgalx=g[0], my=g[1], mgalz=g[2],
                                     s=float(rng.uniform(*ratio)),
                                     Re=float(rng.uniform(*re)),
                                     seed=int(rng.integers(2**31))))
    return out
ENSEMBLE_GRIDS = [(128, 129, 64), (192, 193, 96), (256, 257, 128)]
SERVE_GRIDS = [(64, 129, 64), (128, 129, 64), (192, 193, 96), (256, 257, 128),
               (384, 385, 192)]
