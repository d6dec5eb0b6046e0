"""Differentiable rollout of the NAIVE box step: gradients of a loss on the state after T
steps with respect to the initial velocities, the initial positions and the viscosity.

The reference has no gradients; the forward map is exactly its step (sim/src/sph/
strategies/naive_sph_strategy.py with the box collision of base_kernels.py:75-98), run in
FP32 by libsph_amd.  The backward pass is hand-written (libsph_amd's sph_adjoint_* kernels):
the rollout keeps the state every `checkpoint_every` steps, and for each segment, last to
first, recomputes the segment's states from its checkpoint and pulls the cotangent back
through the steps one at a time.  Memory: (T / k + k) states of 24 bytes per particle.
When T step snapshots (about 270 bytes per particle each) fit in KEEP_ALL_FRACTION of the
device memory free when the simulator is made, the rollout keeps them all and the backward
pass runs no forward step at all (1M particles over 200 steps: 57 GB).

    sim = DifferentiableSPH(params, constants, steps=200, checkpoint_every=20)
    xT, vT = rollout(sim, x0, v0, mu)          # x0, v0 (N, 3) float32 cuda; mu a tensor
    loss = ((xT - target) ** 2).sum(1).mean()
    loss.backward()                            # fills v0.grad, mu.grad (and x0.grad)
"""
from __future__ import annotations

import math
from dataclasses import replace
from typing import Optional, Tuple

import torch

from . import _native as N
from .constants import SPHConstants
from .data_classes import SimulationParameters


# Share of the free device memory the rollout may spend on keeping every step's snapshot
# (keep_all=None).
KEEP_ALL_FRACTION = 0.5


class Tape:
    """What the backward pass needs of one forward rollout.  generation: which forward call
    made it (the simulator's per-step snapshots belong to its latest forward call only)."""

    def __init__(self, checkpoints: torch.Tensor, visc: float, steps: int, generation: int = 0):
        self.checkpoints = checkpoints  # (n_ck, 6, N) float32: x, y, z, vx, vy, vz rows
        self.visc = visc
        self.steps = steps
        self.generation = generation


class DifferentiableSPH:
    """graphs: replay the forward rollout and each backward segment as HIP graphs (one
    launch instead of hundreds; captured once per viscosity value, since the step's
    constants are kernel arguments).  snapshots: the backward's recompute pass keeps each
    step's cell-ordered state, densities, lists and cell table (sph_step_save), so that
    reversing the step does not sort and recompute it again; it costs checkpoint_every
    snapshots of device memory (about 270 bytes per particle each) and gives the same
    gradients bit for bit.  keep_all: the forward rollout itself keeps all `steps` snapshots
    (steps times that memory), so that the backward of the latest forward call reverses
    every step from them without recomputing any (an older tape falls back to its
    checkpoints); None keeps them when they fit in KEEP_ALL_FRACTION of the free device
    memory.  The choice changes speed and memory only: the gradients are the same bits."""

    def __init__(self, params: SimulationParameters, constants: SPHConstants, steps: int, checkpoint_every: int = 20,
                 device="cuda:0", graphs: bool = True, launch: Optional[Tuple[int, int]] = None, lanes: int = 1,
                 snapshots: bool = True, keep_all: Optional[bool] = None):
        if steps < 1 or checkpoint_every < 1:
            raise ValueError("steps and checkpoint_every must be positive")
        self.params = params
        self.constants = constants
        self.steps = int(steps)
        self.every = int(checkpoint_every)
        self.graphs = bool(graphs)
        p = N.make_params(params.particle_count, params.external_force, 1.0 / params.fps, params.space_size,
                          params.voxel_size, N.SPH_MODE_BOX, N.SPH_FORM_NAIVE, flags=N.SPH_FLAG_ADJOINT)
        self.ctx = N.NativeContext(p, N.make_consts(constants), device)
        self.device = self.ctx.device
        self.n = self.ctx.n
        if launch:
            self.ctx.set_launch(*launch)
        if lanes != 1:
            self.ctx.set_lanes(lanes)
        f32 = dict(dtype=torch.float32, device=self.device)
        self.snapshots = bool(snapshots)
        # states of the segment being reversed (with snapshots, only its first)
        self._seg = torch.empty((1 if self.snapshots else self.every, 6, self.n), **f32)
        nbytes = self.ctx.adjoint_snapshot_bytes()
        if nbytes <= 0:
            raise RuntimeError("libsph_amd gives no step snapshots for this context")
        if keep_all is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            keep_all = self.snapshots and self.steps * nbytes <= KEEP_ALL_FRACTION * free
        self.keep_all = bool(keep_all)
        if self.snapshots:
            self._snap = torch.empty((self.every, nbytes), dtype=torch.uint8, device=self.device)
        if self.keep_all:
            self._all = torch.empty((self.steps, nbytes), dtype=torch.uint8, device=self.device)
        self._generation = 0  # forward calls so far; _all holds the latest one's snapshots
        self._in = torch.empty((6, self.n), **f32)                    # rollout input (x0, v0 rows)
        self._ck = torch.empty((self.num_checkpoints, 6, self.n), **f32)
        self._out = torch.empty((6, self.n), **f32)
        self._fwd_graphs = {}
        self._seg_graphs = {}
        self._rev_graphs = {}
        self._graph_visc = {}  # id(cache) -> viscosity of that cache's previous call

    def set_keep_all(self, keep_all: bool) -> None:
        """Switches between keeping every step's snapshot (allocated at construction) and the
        checkpoint-only backward; drops the recorded graphs, which hold the other scheme."""
        keep_all = bool(keep_all)
        if keep_all and not hasattr(self, "_all"):
            raise ValueError("keep_all needs the per-step snapshots allocated at construction")
        self.keep_all = keep_all
        self._generation += 1  # no tape recorded so far is reversed from snapshots any more
        for cache in (self._fwd_graphs, self._seg_graphs, self._rev_graphs):
            cache.clear()
        self._graph_visc.clear()

    @property
    def num_checkpoints(self) -> int:
        return math.ceil(self.steps / self.every)

    def set_viscosity(self, mu: float) -> None:
        self.ctx.set_constants(N.make_consts(replace(self.constants, visc=float(mu))))

    # -- state rows --------------------------------------------------------------------
    def _save(self, rows: torch.Tensor) -> None:
        self.ctx.download_f32([rows[0], rows[1], rows[2], rows[3], rows[4], rows[5], None])

    def _load(self, rows: torch.Tensor) -> None:
        self.ctx.upload_f32(rows[0], rows[1], rows[2], rows[3], rows[4], rows[5])

    def _rows(self, x: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
        x = x.detach().to(self.device, torch.float32).reshape(self.n, 3)
        v = v.detach().to(self.device, torch.float32).reshape(self.n, 3)
        return torch.cat([x.t(), v.t()]).contiguous()

    # -- passes ------------------------------------------------------------------------
    def _forward_body(self) -> None:
        """_in -> _ck (every `every` steps, _ck[0] = _in), _out and, with keep_all, the
        snapshot of every step in _all."""
        self._load(self._in)
        t = 0
        while t < self.steps:
            if self.keep_all:
                self.ctx.step_save(self._all[t])
                t += 1
            else:
                # up to the next checkpoint in one call (bitwise the same as single steps,
                # and its sorts after the first start from counts the force passes made)
                k = min(self.every - t % self.every, self.steps - t)
                self.ctx.step(k)
                t += k
            if t % self.every == 0 and t < self.steps:
                self._save(self._ck[t // self.every])
        self._save(self._out)

    def _reverse_all_body(self) -> None:
        """Pulls the cotangent back through every step of the latest rollout, last to first,
        from its snapshots."""
        for t in reversed(range(self.steps)):
            self.ctx.adjoint_step_from(self._all[t])

    def _segment_body(self, length: int) -> None:
        """Reverses one segment whose first state is in _seg[0]: recomputes its states, then
        pulls the cotangent back through its steps, last to first."""
        self._load(self._seg[0])
        if self.snapshots:
            for a in range(length):
                self.ctx.step_save(self._snap[a])
            for a in reversed(range(length)):
                self.ctx.adjoint_step_from(self._snap[a])
            return
        for a in range(1, length):
            self.ctx.step(1)
            self._save(self._seg[a])
        for a in reversed(range(length)):
            self._load(self._seg[a])
            self.ctx.adjoint_step()

    def _run(self, cache: dict, key, body) -> None:
        """Runs body, from a HIP graph when graphs are on.  A graph bakes the viscosity into
        its kernel arguments, so graphs are kept for one viscosity only (the key's first
        element): a pass at a viscosity not seen on the previous call of this cache runs
        directly, the second consecutive call at one viscosity records its graph, and graphs
        of any other viscosity are dropped.  An optimisation loop that changes mu every
        iteration therefore never records a graph it would replay once."""
        if not self.graphs:
            body()
            return
        visc = key[0] if isinstance(key, tuple) else key
        last = self._graph_visc.get(id(cache))
        self._graph_visc[id(cache)] = visc
        g = cache.get(key)
        if g is None:
            for k in [k for k in cache if (k[0] if isinstance(k, tuple) else k) != visc]:
                del cache[k]
            if last != visc:
                body()
                return
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize(self.device)
            with torch.cuda.graph(g):
                body()
            cache[key] = g
        g.replay()

    def forward(self, x0: torch.Tensor, v0: torch.Tensor, mu: float) -> Tuple[torch.Tensor, torch.Tensor, Tape]:
        """Runs the rollout; returns x_T, v_T (float32 (N, 3), particle order) and the tape."""
        mu = float(mu)
        self.set_viscosity(mu)
        self._in.copy_(self._rows(x0, v0))
        self._run(self._fwd_graphs, mu, self._forward_body)
        self._generation += 1
        ck = self._ck.clone()
        ck[0] = self._in
        out = self._out
        return out[:3].t().contiguous(), out[3:].t().contiguous(), Tape(ck, mu, self.steps, self._generation)

    def backward(self, tape: Tape, gx: Optional[torch.Tensor], gv: Optional[torch.Tensor]):
        """Cotangents of x_T, v_T (float32 (N, 3) or None) -> (dL/dx0, dL/dv0, dL/dmu)."""
        self.set_viscosity(tape.visc)
        f32 = dict(dtype=torch.float32, device=self.device)
        gx = None if gx is None else gx.detach().to(**f32).reshape(self.n, 3).contiguous()
        gv = None if gv is None else gv.detach().to(**f32).reshape(self.n, 3).contiguous()
        self.ctx.adjoint_seed(gx, gv)
        if self.keep_all and tape.generation == self._generation:
            self._run(self._rev_graphs, tape.visc, self._reverse_all_body)
            return self.ctx.adjoint_read()
        for seg in reversed(range(self.num_checkpoints)):
            length = min(self.every, tape.steps - seg * self.every)
            self._seg[0].copy_(tape.checkpoints[seg])
            self._run(self._seg_graphs, (tape.visc, length), lambda: self._segment_body(length))
        return self.ctx.adjoint_read()


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, mu, sim):
        xT, vT, tape = sim.forward(x0, v0, float(mu))
        ctx.sim, ctx.tape = sim, tape
        ctx.meta = [(t.dtype, t.device, t.shape) for t in (x0, v0, mu)]
        return xT, vT

    @staticmethod
    def backward(ctx, gxT, gvT):
        gx0, gv0, gmu = ctx.sim.backward(ctx.tape, gxT, gvT)
        (dx, devx, sx), (dv, devv, sv), (dm, devm, sm) = ctx.meta
        return (gx0.to(devx, dx).reshape(sx), gv0.to(devv, dv).reshape(sv), gmu.to(devm, dm).reshape(sm), None)


def rollout(sim: DifferentiableSPH, x0: torch.Tensor, v0: torch.Tensor, mu: torch.Tensor):
    """x_T, v_T after sim.steps steps from (x0, v0) at viscosity mu, differentiable with
    respect to x0, v0 and mu (a tensor; a 0-d or one-element float tensor)."""
    if not torch.is_tensor(mu):
        mu = torch.tensor(float(mu), dtype=torch.float64)
    return _Rollout.apply(x0, v0, mu, sim)
