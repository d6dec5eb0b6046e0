"""Batched form: many independent box simulations advanced together.

The reference advances one simulation per process (sim/src/state_generator.py:14-37).
BatchedSPH advances B of them, each with its own particle count, voxel grid, h, mu, box
and dt, with one launch per stage for all instances (libsph_amd's sph_create_batch), and
can replay whole runs of steps as one HIP graph.  Every instance computes exactly what a
NaiveSPHStrategy on that instance alone computes (same cell order within the instance,
same neighbour lists, same arithmetic).
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _native as N
from .constants import SPHConstants
from .data_classes import SimulationParameters, SimulationState
from .workloads import voxel_for


def _per_instance(value, count: int, what: str) -> list:
    """One value per instance from a scalar or a sequence of length count."""
    if value is None or np.isscalar(value):
        return [value] * count
    out = list(value)
    if len(out) != count:
        raise ValueError(f"{len(out)} values of {what} for {count} instances")
    return out


class BatchedSPH:
    def __init__(self, instances: Sequence[Tuple[SimulationParameters, SPHConstants]], device="cuda:0",
                 launch: Optional[Tuple[int, int]] = None,
                 cell_scale: Optional[Union[float, Sequence[float]]] = None,
                 lanes: Union[int, Sequence[int]] = 1):
        """launch: (density, force) threads per workgroup.  cell_scale: when given (one value,
        or one per instance), instance b searches a grid of edge voxel_for(h_b) * cell_scale_b
        instead of its voxel_size (the NAIVE sums do not depend on the grid; see autotune.py).
        lanes: lanes of a wavefront per particle in the density and force passes, one value or
        one per instance (1, 2, 4, 8, 16, 32 or 64)."""
        self.instances = list(instances)
        count = len(self.instances)
        self.cell_scale = _per_instance(cell_scale, count, "cell_scale")
        self.lanes = _per_instance(lanes, count, "lanes")

        def grid(b, p, c):
            sc = self.cell_scale[b]
            return p.voxel_size if sc is None else [voxel_for(c.h) * sc] * 3

        params = [N.make_params(p.particle_count, p.external_force, 1.0 / p.fps, p.space_size, grid(b, p, c),
                                N.SPH_MODE_BOX, N.SPH_FORM_NAIVE) for b, (p, c) in enumerate(self.instances)]
        consts = [N.make_consts(c) for _, c in self.instances]
        self.ctx = N.NativeContext(params, consts, device)
        self.device = self.ctx.device
        self.first = self.ctx.first
        self.n = self.ctx.n
        if launch:
            self.ctx.set_launch(*launch)
        if any(L != 1 for L in self.lanes):
            self.ctx.set_batch_lanes(self.lanes)
        self._graphs = {}

    # -- state ---------------------------------------------------------------------------
    def upload(self, states: Sequence[SimulationState]) -> None:
        if len(states) != len(self.instances):
            raise ValueError(f"{len(states)} states for {len(self.instances)} instances")
        pos = np.concatenate([np.asarray(s.position, dtype=np.float64).reshape(-1, 3) for s in states])
        vel = np.concatenate([np.asarray(s.velocity, dtype=np.float64).reshape(-1, 3) for s in states])
        if pos.shape[0] != self.n:
            raise ValueError(f"states hold {pos.shape[0]} particles, the batch has {self.n}")
        self._pos = torch.from_numpy(pos).to(self.device)
        self._vel = torch.from_numpy(vel).to(self.device)
        self.ctx.upload(self._pos, self._vel)

    def ctx_state_tensors(self) -> Tuple[torch.Tensor, torch.Tensor]:
        """The last uploaded state as float64 (N, 3) device tensors."""
        return self._pos, self._vel

    def download(self) -> List[SimulationState]:
        pos, vel, rho = [t.cpu().numpy() for t in self.ctx.download()]
        out = []
        for b in range(len(self.instances)):
            a, e = self.first[b], self.first[b + 1]
            out.append(SimulationState(pos[a:e], vel[a:e], rho[a:e]))
        return out

    # -- stepping ------------------------------------------------------------------------
    def step(self, nsteps: int = 1) -> None:
        """nsteps steps of every instance, launched directly."""
        self.ctx.step(nsteps)

    def capture(self, nsteps: int) -> torch.cuda.CUDAGraph:
        """A HIP graph of nsteps steps.  The step records no memsets and no host work, so
        the graph replays the same launches on the resident state."""
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize(self.device)
        with torch.cuda.graph(g):
            self.ctx.step(nsteps)
        self._graphs[nsteps] = g
        return g

    def replay(self, nsteps: int) -> None:
        g = self._graphs.get(nsteps) or self.capture(nsteps)
        g.replay()
