"""The reference's run configuration (config.py) as seeded functions.

config.py:13-43 fixes the run (mode, strategy, 20,000 particles, 15 s at 20 fps, the
physics constants, space sizes, forces) and builds the start state with unseeded draws
(numpy's global generator and random.random, config.py:79-120).  Here the same
constructions take a numpy Generator seed, so a run is reproducible; every draw has the
reference's distribution and bounds:

  box wall   x ~ U[0, 0.1 Sx), y ~ U[0, Sy), z ~ U[0, Sz); v = (1.5, -5, -5) + U[-0.5, 0.5)
             per component; density 0                               (config.py:79-98)
  pipe fill  x ~ U[0, pipe length), r = sqrt(u) R(x) 0.98, theta ~ U[0, 2 pi), centred on
             the first segment's axis; v = 0; density 0                (config.py:100-120)

R(x) is Pipe.radius_at(x) (data_classes.py:59-63), including its quirk that an x in no
segment indexes segments[-1].  The presets return (params, constants, start state, mode)
for config.SIM_MODE 'BOX' and 'PIPE'; run(...) is sim/src/main.py's frame loop on the GPU.
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from .constants import SPHConstants
from .data_classes import Pipe, Segment, SimulationParameters, SimulationState
from .pipe_builder import PipeBuilder

# config.py:15-43
DURATION = 15
FPS = 20
PARTICLE_COUNT = 20_000
INF_R = 2.0
VOXEL_SIZE = [INF_R, INF_R, INF_R]
BOX_SPACE_SIZE = [20 * INF_R, 20 * INF_R, 20 * INF_R]
PIPE_SPACE_SIZE = [20 * INF_R, 3 * INF_R, 3 * INF_R]
GRAVITY = [0.0, -2.0, 0.0]
HORIZONTAL_FORCE = [2.0, 0.0, 0.0]
OUT_DIRNAME = "out"


def start_state_box_wall(particle_count: int = PARTICLE_COUNT, space_size=BOX_SPACE_SIZE,
                         seed: Optional[int] = 0) -> SimulationState:
    """config.py:79-98: a wall of particles in the first tenth of the box along x, thrown
    towards -y and -z."""
    rng = np.random.default_rng(seed)
    n = int(particle_count)
    pos = rng.random((n, 3))
    pos[:, 0] *= space_size[0] * 0.1
    pos[:, 1] *= space_size[1]
    pos[:, 2] *= space_size[2]
    base = np.array([1.5, -5.0, -5.0])
    offset_range = 1.0
    vel = base + (rng.random((n, 3)) - 0.5) * offset_range
    return SimulationState(pos, vel, np.zeros(n))


def start_state_inside_pipe(pipe: Pipe, particle_count: int = PARTICLE_COUNT,
                            seed: Optional[int] = 0) -> SimulationState:
    """config.py:100-120: particles spread over the pipe's length, inside 0.98 of the wall
    radius around the first segment's axis, at rest."""
    rng = np.random.default_rng(seed)
    n = int(particle_count)
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    offset = pipe.segments[0].start_point
    length = pipe.get_length()
    for i in range(n):
        x = rng.random() * length
        r_max = pipe.radius_at(x)
        r = math.sqrt(rng.random()) * r_max * 0.98
        theta = rng.random() * 2.0 * math.pi
        pos[i, 0] = x
        pos[i, 1] = offset[1] + r * math.cos(theta)
        pos[i, 2] = offset[2] + r * math.sin(theta)
    return SimulationState(pos, vel, np.zeros(n))


def config_pipe() -> Pipe:
    """config.py:68-76: roller, widening, roller, narrowing, roller, rescaled to the pipe
    space."""
    return (PipeBuilder().with_starting_radius(1)
            .add_roller_segment(1)
            .add_increasing_segment(1, 1.2)
            .add_roller_segment(1)
            .add_lessening_segment(1, 1.2)
            .add_roller_segment(1)
            .transform(PIPE_SPACE_SIZE[0], PIPE_SPACE_SIZE[1])
            .get_result())


def preset(mode: str = "BOX", particle_count: int = PARTICLE_COUNT, seed: Optional[int] = 0,
           duration: int = DURATION, fps: int = FPS) -> Tuple[SimulationParameters, SPHConstants, SimulationState, str]:
    """config.py:44-66 and 123-134: parameters, constants (config.py:18-23, the SPHConstants defaults),
    start state and mode of the reference's default run in BOX or PIPE mode."""
    if mode == "BOX":
        pipe = Pipe([Segment()])
        params = SimulationParameters(particle_count=particle_count, external_force=np.asarray(GRAVITY),
                                      duration=duration, fps=fps, pipe=pipe, space_size=np.asarray(BOX_SPACE_SIZE),
                                      voxel_size=np.asarray(VOXEL_SIZE))
        state = start_state_box_wall(particle_count, BOX_SPACE_SIZE, seed)
    elif mode == "PIPE":
        pipe = config_pipe()
        params = SimulationParameters(particle_count=particle_count, external_force=np.asarray(HORIZONTAL_FORCE),
                                      duration=duration, fps=fps, pipe=pipe, space_size=np.asarray(PIPE_SPACE_SIZE),
                                      voxel_size=np.asarray(VOXEL_SIZE))
        state = start_state_inside_pipe(pipe, particle_count, seed)
    else:
        raise Exception(f"Wrong simulation mode! ({mode})")
    return params, SPHConstants(), state, mode


def run(mode: str = "BOX", strategy: str = "NAIVE", out_dirname: str = OUT_DIRNAME, root: Optional[str] = None,
        particle_count: int = PARTICLE_COUNT, seed: Optional[int] = 0, frames: Optional[int] = None,
        device="cuda:0") -> int:
    """sim/src/main.py:10-26: the preset run for duration * fps frames (or `frames`), every
    frame saved in the Saver format.  Returns the number of frames written."""
    from .serializer import run_to_disk
    from .strategy import make_strategy

    params, consts, state, mode = preset(mode, particle_count, seed)
    strat = make_strategy(strategy, params, constants=consts, mode=mode, device=device)
    return run_to_disk(strat, state, params, out_dirname, frames=frames, root=root)
