"""Physics constants of the SPH step.

In the reference they are module globals of config.py (config.py:18-23) that numba
freezes into the kernels at JIT time.  Here they are a runtime value passed to each
strategy, with the same defaults.
"""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class SPHConstants:
    mass: float = 1.0   # MASS  (config.py:18)
    rho0: float = 1.0   # RHO_0 (config.py:19)
    h: float = 2.0      # INF_R (config.py:20), smoothing radius
    visc: float = 0.5   # VISC  (config.py:21)
    k: float = 10.0     # K     (config.py:22), p = K (rho - rho0)
    damp: float = 0.7   # DAMP  (config.py:23)

    # derived forms (config.py:24-29)
    @property
    def w_const(self) -> float:
        return 315.0 / (64.0 * math.pi * self.h ** 9)

    @property
    def grad_w_const(self) -> float:
        return -45.0 / (math.pi * self.h ** 6)

    @property
    def lap_w_const(self) -> float:
        return 45.0 / (math.pi * self.h ** 6)


MAX_NEIGHBOURS = 32             # config.py:30, VOXEL formulation cap
NEIGHBOURING_VOXELS_COUNT = 27  # config.py:31
RNG_SEED = 16435234             # abstract_sph_strategy.py:27
