"""Seeded synthetic workloads (numpy.random.default_rng) for tests and benchmarks.

  box_4k        4,096 particles uniform in [0.1, 0.5]^3 of the unit box, at rest; m 0.02,
                h 0.0457, rho0 1000, K 3, mu 3.5, damp 0.5, f (0, -9.81, 0), dt 1e-3.
  dam_break_1m  1,048,576 particles on a 64 x 128 x 128 lattice filling [0,0.4]x[0,0.8]x[0,1],
                jittered by U(+-0.1 spacing), box [0,2]x[0,1]x[0,1]; same constants with
                h = 0.01298 (about 30 neighbours at the mean spacing).
  pipe_512k     524,288 particles in an 8-segment pipe (segments 0.5 long, radii U[0.1, 0.3])
                along +x with its axis at y = z = 0.3, seeded uniformly inside the first two
                segments with velocity (1, 0, 0) + N(0, 0.05^2); f (2, 0, 0) (the reference's
                pipe-mode force, config.py:36); h chosen for about 30 neighbours.
  batch_64      64 box instances, N_b log-uniform in [2048, 65536], mu U[1, 8], h U[0.03, 0.06],
                box edge U[0.5, 2]; particles uniform in [0.1, 0.5]^3 * edge.
  adjoint_16k   16,384 particles as box_4k.
Voxel edges are the largest power of two not above h (exact keys in FP32; each axis
searches ceil(h / voxel) = 2 cells on each side, which tests fewer candidates than cells
of edge >= h).  The NAIVE formulation's results do not depend on the voxel edge.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .constants import SPHConstants
from .data_classes import Pipe, Segment, SimulationParameters, SimulationState

BOX_CONSTANTS = dict(mass=0.02, rho0=1000.0, k=3.0, visc=3.5, damp=0.5)


def pow2_at_least(x: float) -> float:
    """Smallest power of two >= x."""
    return 2.0 ** math.ceil(math.log2(x))


def voxel_for(h: float) -> float:
    """Largest power of two <= h: the voxel edge used by the workloads."""
    return 2.0 ** math.floor(math.log2(h))


@dataclass
class Workload:
    name: str
    params: SimulationParameters
    constants: SPHConstants
    state: SimulationState
    steps: int
    mode: str = "BOX"
    notes: str = ""


def _params(n, ext, fps, space, voxel, pipe=None, duration=1) -> SimulationParameters:
    return SimulationParameters(particle_count=n, external_force=np.asarray(ext, dtype=np.float64), duration=duration,
                                fps=fps, pipe=pipe if pipe is not None else Pipe([Segment()]),
                                space_size=np.asarray(space, dtype=np.float64),
                                voxel_size=np.asarray(voxel, dtype=np.float64))


def box_uniform(n: int, seed: int = 0, h: float = 0.0457, edge: float = 1.0, visc: float = 3.5,
                steps: int = 100) -> Workload:
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.1 * edge, 0.5 * edge, size=(n, 3))
    vel = np.zeros((n, 3))
    c = SPHConstants(h=h, **{**BOX_CONSTANTS, "visc": visc})
    vx = voxel_for(h)
    p = _params(n, (0.0, -9.81, 0.0), 1000, (edge, edge, edge), (vx, vx, vx))
    return Workload(f"box_uniform_n{n}_seed{seed}", p, c, SimulationState(pos, vel, np.zeros(n)), steps)


def box_4k(seed: int = 0) -> Workload:
    w = box_uniform(4096, seed=seed)
    w.name = "box_4k"
    return w


def dam_break(nx: int = 64, ny: int = 128, nz: int = 128, seed: int = 1, steps: int = 500,
              h: Optional[float] = None) -> Workload:
    sx, sy, sz = 0.4 / nx, 0.8 / ny, 1.0 / nz
    rng = np.random.default_rng(seed)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = np.stack([(ix.ravel() + 0.5) * sx, (iy.ravel() + 0.5) * sy, (iz.ravel() + 0.5) * sz], axis=1)
    jitter = rng.uniform(-0.1, 0.1, size=base.shape) * np.array([sx, sy, sz])
    pos = base + jitter
    n = pos.shape[0]
    if h is None:
        spacing = (sx * sy * sz) ** (1.0 / 3.0)
        h = spacing * (30.0 / (4.0 / 3.0 * math.pi)) ** (1.0 / 3.0)
    vx = voxel_for(h)
    c = SPHConstants(h=h, **BOX_CONSTANTS)
    p = _params(n, (0.0, -9.81, 0.0), 1000, (2.0, 1.0, 1.0), (vx, vx, vx))
    return Workload(f"dam_break_{n}", p, c, SimulationState(pos, np.zeros((n, 3)), np.zeros(n)), steps)


def dam_break_1m(seed: int = 1) -> Workload:
    w = dam_break(64, 128, 128, seed=seed, h=0.01298)
    w.name = "dam_break_1m"
    return w


def pipe_radius_np(tbl: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Wall radius at x with the reference's table semantics (util_kernels.py:56-77)."""
    xs = tbl[:, 0]
    seg = np.clip(np.searchsorted(xs, x, side="right") - 1, 0, len(xs) - 2)
    r0, r1 = tbl[seg, 3], tbl[seg + 1, 3]
    t = (x - xs[seg]) / tbl[seg, 4]
    return r0 + (r1 - r0) * t


def pipe_flow(n: int = 524288, seed: int = 2, steps: int = 1000, segments: int = 8, seg_len: float = 0.5,
              axis_yz: float = 0.3, neighbours: float = 30.0) -> Workload:
    rng = np.random.default_rng(seed)
    radii = rng.uniform(0.1, 0.3, size=segments)
    segs = [Segment(start_point=(k * seg_len, axis_yz, axis_yz), start_radius=float(r), end_radius=float(r),
                    length=seg_len) for k, r in enumerate(radii)]
    pipe = Pipe(segs)
    tbl = pipe.to_numpy()
    # volume of the first two segments under the table's wall, by midpoint quadrature
    xs = (np.arange(20000) + 0.5) / 20000 * 2 * seg_len
    vol = float(np.mean(math.pi * pipe_radius_np(tbl, xs) ** 2) * 2 * seg_len)
    # x with density proportional to the cross-section (rejection), then r = R sqrt(u)
    rmax = float(np.max(radii[:2]))
    xs_out = np.empty(0)
    while xs_out.size < n:
        cand = rng.uniform(0.0, 2 * seg_len, size=2 * n)
        keep = rng.uniform(0, 1, size=cand.size) < (pipe_radius_np(tbl, cand) / rmax) ** 2
        xs_out = np.concatenate([xs_out, cand[keep]])
    x = xs_out[:n]
    R = pipe_radius_np(tbl, x)
    r = R * np.sqrt(rng.uniform(0, 1, size=n))
    th = rng.uniform(0, 2 * math.pi, size=n)
    pos = np.stack([x, axis_yz + r * np.cos(th), axis_yz + r * np.sin(th)], axis=1)
    vel = np.zeros((n, 3))
    vel[:, 0] = 1.0
    vel += rng.normal(0.0, 0.05, size=(n, 3))
    spacing = (vol / n) ** (1.0 / 3.0)
    h = spacing * (neighbours / (4.0 / 3.0 * math.pi)) ** (1.0 / 3.0)
    vx = voxel_for(h)
    c = SPHConstants(h=h, **BOX_CONSTANTS)
    p = _params(n, (2.0, 0.0, 0.0), 1000, (segments * seg_len, 2 * axis_yz, 2 * axis_yz), (vx, vx, vx), pipe=pipe)
    return Workload(f"pipe_{n}", p, c, SimulationState(pos, vel, np.zeros(n)), steps, mode="PIPE",
                    notes=f"h={h:.5f}, radii={np.round(radii, 4).tolist()}")


@dataclass
class BatchWorkload:
    name: str
    instances: List[Workload] = field(default_factory=list)
    steps: int = 300

    @property
    def total_particles(self) -> int:
        return sum(w.params.particle_count for w in self.instances)


def batch_64(seed: int = 3, count: int = 64, steps: int = 300, n_lo: int = 2048, n_hi: int = 65536) -> BatchWorkload:
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), size=count))).astype(int)
    viscs = rng.uniform(1.0, 8.0, size=count)
    hs = rng.uniform(0.03, 0.06, size=count)
    edges = rng.uniform(0.5, 2.0, size=count)
    inst = []
    for b in range(count):
        w = box_uniform(int(sizes[b]), seed=seed * 1000 + b, h=float(hs[b]), edge=float(edges[b]),
                        visc=float(viscs[b]), steps=steps)
        w.name = f"batch{b}"
        inst.append(w)
    return BatchWorkload("batch_64", inst, steps)


def adjoint_16k(seed: int = 4) -> Workload:
    w = box_uniform(16384, seed=seed, steps=200)
    w.name = "adjoint_16k"
    return w
