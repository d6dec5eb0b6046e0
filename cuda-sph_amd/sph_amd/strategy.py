"""Drop-in SPH step strategies.

SPHStrategy mirrors AbstractSPHStrategy (sim/src/sph/strategies/abstract_sph_strategy.py:15-132):
it is constructed from SimulationParameters and advances a SimulationState with
compute_next_state.  NaiveSPHStrategy and VoxelSPHStrategy select the two reference
formulations (naive_sph_strategy.py, voxel_sph_strategy.py).  The work runs on the GPU
through libsph_amd.so; the state stays resident on the device between calls.

Differences in behaviour at the boundary:
  - physics constants are an SPHConstants value (defaults = config.py:18-23) instead of
    module globals, and the collision mode ('BOX' / 'PIPE') is a constructor argument
    instead of config.SIM_MODE;
  - arrays of the returned state are read-only.  Passing the returned state back skips
    the host-to-device copy while the device still holds it (no other call has changed
    the resident state since); any other state object is uploaded;
  - a step that leaves a particle position NaN or infinite raises NonFiniteStateError
    (a RuntimeError) instead of returning that state (check_finite=False restores the
    reference's silent behaviour).
"""
from __future__ import annotations

import logging
from typing import Optional, Tuple

import numpy as np
import torch

from . import _native as N
from .constants import RNG_SEED, SPHConstants
from .data_classes import SimulationParameters, SimulationState

logger = logging.getLogger(__name__)

_MODES = {"BOX": N.SPH_MODE_BOX, "PIPE": N.SPH_MODE_PIPE}


class SPHStrategy:
    """Template of one step (abstract_sph_strategy.py:31-46).  Subclasses set formulation."""

    formulation = N.SPH_FORM_NAIVE
    # the reference's template stages, in order; on the GPU the pressure, viscosity,
    # integration and collision stages run as one fused kernel
    STAGES = ("_initialize_computation", "_compute_density", "_compute_pressure", "_compute_viscosity",
              "_integrate", "_collide", "_finalize_computation")

    def __init__(self, params: SimulationParameters, constants: Optional[SPHConstants] = None, mode: str = "BOX",
                 device="cuda:0", keep_terms: bool = False, seed: int = RNG_SEED, grid_voxel: Optional[float] = None,
                 launch: Optional[Tuple[int, int]] = None, lanes: int = 1, check_finite: bool = True,
                 slabs: bool = True):
        """grid_voxel: NAIVE formulation only, the edge of the search grid's cells in place
        of params.voxel_size (the NAIVE sums do not depend on it; see autotune.py).
        launch: (density, force) threads per workgroup.  lanes: NAIVE only, lanes of a
        wavefront per particle in the density and force passes.  check_finite: every step
        call waits for its steps and raises NonFiniteStateError if a position turned
        NaN or infinite.  slabs: NAIVE only, whether the search cuts its cells into x slabs
        (False: whole cells; see autotune.py)."""
        if mode not in _MODES:
            raise Exception(f"Wrong simulation mode! ({mode})")
        if grid_voxel is not None and self.formulation != N.SPH_FORM_NAIVE:
            raise ValueError("grid_voxel applies to the NAIVE formulation; VOXEL uses params.voxel_size")
        self.params = params
        self.constants = constants or SPHConstants()
        self.mode = mode
        self.dt = 1 / self.params.fps
        self.old_state: Optional[SimulationState] = None
        self.new_state: Optional[SimulationState] = None
        n = int(params.particle_count)
        self.result_force = np.zeros((n, 3), dtype=np.float64)
        table = params.pipe.to_numpy() if mode == "PIPE" else None
        voxel = params.voxel_size if grid_voxel is None else [float(grid_voxel)] * 3
        flags = (N.SPH_FLAG_KEEP_TERMS if keep_terms else 0) | (N.SPH_FLAG_CHECK_FINITE if check_finite else 0)
        self._params_c = N.make_params(n, params.external_force, self.dt, params.space_size, voxel,
                                       _MODES[mode], self.formulation, table, flags, seed)
        self.keep_terms = keep_terms
        self.ctx = N.NativeContext(self._params_c, N.make_consts(self.constants), device)
        self.device = self.ctx.device
        self._launch = (256, 256)
        self._lanes = 1
        self._returned_generation = -1  # ctx.generation when new_state was downloaded
        if launch:
            self.ctx.set_launch(*launch)
            self._launch = (int(launch[0]), int(launch[1]))
        if lanes != 1:
            self.ctx.set_lanes(lanes)
            self._lanes = int(lanes)
        self.slabs = bool(slabs) if self.formulation == N.SPH_FORM_NAIVE else False
        if self.formulation == N.SPH_FORM_NAIVE and not slabs:
            self.ctx.set_slabs(False)
        logger.info("SPH strategy on %s: %s", self.device, self.ctx.grid_info())

    # -- reference entry point -----------------------------------------------------------
    def compute_next_state(self, old_state: SimulationState) -> SimulationState:
        self.old_state = old_state
        self._initialize_computation()
        self.ctx.step(1)
        self._finalize_computation()
        return self.new_state

    def advance(self, state: SimulationState, nsteps: int) -> SimulationState:
        """nsteps steps with one upload and one download."""
        self.old_state = state
        self._initialize_computation()
        self.ctx.step(int(nsteps))
        self._finalize_computation()
        return self.new_state

    def capture(self, nsteps: int) -> torch.cuda.CUDAGraph:
        """A HIP graph of nsteps steps of the resident state (recorded, not run).  The step
        records no memsets and no host work, so each replay advances the resident state by
        nsteps steps exactly as ctx.step(nsteps) does.  Replays change the resident state
        behind this object's back, so the last returned state is re-uploaded when it is
        passed back; the finite check of a captured step is left to the next step call
        outside the graph."""
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize(self.device)
        with torch.cuda.graph(g):
            self.ctx.step(int(nsteps))
        self.new_state = None
        return g

    # -- stages --------------------------------------------------------------------------
    def _initialize_computation(self):
        """Host-to-device copy of the state (abstract_sph_strategy.py:64-74), skipped when the
        state is the one this strategy returned last."""
        st = self.old_state
        if st is not None and st is self.new_state and self.ctx.generation == self._returned_generation:
            return
        pos = torch.from_numpy(np.ascontiguousarray(st.position, dtype=np.float64)).to(self.device)
        vel = torch.from_numpy(np.ascontiguousarray(st.velocity, dtype=np.float64)).to(self.device)
        if pos.shape != (self.ctx.n, 3) or vel.shape != (self.ctx.n, 3):
            raise ValueError(f"state arrays must be ({self.ctx.n}, 3), got {tuple(pos.shape)} / {tuple(vel.shape)}")
        self.ctx.upload(pos, vel)

    def _finalize_computation(self):
        """Device-to-host copy into a new SimulationState (abstract_sph_strategy.py:76-84)."""
        out = self.ctx.download(force=self.keep_terms)
        host = [t.cpu().numpy() for t in out[:3]]
        for a in host:
            a.setflags(write=False)
        if self.keep_terms:
            self.result_force = out[3].cpu().numpy()
        self.new_state = SimulationState(host[0], host[1], host[2])
        self._returned_generation = self.ctx.generation

    def diagnostics(self) -> dict:
        """FP64 totals of the resident state."""
        d = self.ctx.diagnostics()
        return dict(mass=d[0], momentum=d[1:4], kinetic_energy=d[4], non_finite=int(d[5]), max_speed=d[6],
                    mean_density=d[7])

    # thread_layout.organize (thread_layout.py:27-33) gives the reference's launch shape;
    # sim/src/main.py:20-21 logs it as grid_size / block_size
    @property
    def block_size(self) -> int:
        """Threads per workgroup of the force pass (the step's fused integration launch)."""
        return self._launch[1]

    @property
    def grid_size(self) -> int:
        """Workgroups of the force pass: ceil(N * lanes / block_size)."""
        return -(-int(self.ctx.n) * self._lanes // self.block_size)

    @property
    def voxel_dims(self):
        """Voxels per axis of the grid the step sorts by."""
        return self.ctx.grid_info()["dims"]


class NaiveSPHStrategy(SPHStrategy):
    """naive_sph_strategy.py:13-45 formulation: rho over all j within h (self included),
    P = m sum (p_i + p_j)/(2 rho_j) grad W, V = m mu sum (v_j - v_i)/rho_j lap W."""

    formulation = N.SPH_FORM_NAIVE


class VoxelSPHStrategy(SPHStrategy):
    """voxel_sph_strategy.py:16-116 formulation: 32-neighbour lists from the 27 voxels,
    rho without self, P = sum (p_i/rho_i^2 + p_j/rho_j^2) grad W, V = min(1, m mu/rho_i sum ...)."""

    formulation = N.SPH_FORM_VOXEL


def make_strategy(name: str, params: SimulationParameters, **kw) -> SPHStrategy:
    """config.SIM_STRATEGY selection (state_generator.py:21-23)."""
    return (NaiveSPHStrategy if name == "NAIVE" else VoxelSPHStrategy)(params, **kw)
