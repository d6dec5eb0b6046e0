"""Autotuner for the neighbour kernels' launch shapes and the grid they walk.

Three knobs decide how the step maps onto the GPU:
  - the workgroup sizes of the density and force passes (64, 128 or 256 threads: one to
    four wavefronts over consecutive cell-ordered particles);
  - the lanes per particle of those passes (1 to 16 lanes of a wavefront share one
    particle's candidates: small or dense problems gain parallelism); and
  - the grid cell edge, which sets how many cells a workgroup's particles cover ("cells
    per tile") and how many candidates each particle tests.  Smaller cells test fewer
    candidates outside the h-sphere but walk more cell rows.  Only the NAIVE formulation
    may change it: its sums run over every j within h whatever the grid (results equal up
    to FP32 summation order), while in the VOXEL formulation the voxel edge is part of the
    reference's semantics (voxel_sph_strategy.py:109-116), so only block sizes are tuned.

Every candidate is timed on the GPU from the same initial state over the same steps
(HIP events on the stream the step runs on), the fastest wins, and the result is cached
in a JSON file keyed by the workload's signature and the device name, so a later run
with the same configuration skips the search.  Candidate grid edges are powers of two
(voxel_for(h) times 2, 1, 1/2, 1/4), for which cell keys are exact in FP32.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from dataclasses import asdict, dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .constants import SPHConstants
from .data_classes import SimulationParameters, SimulationState
from .workloads import voxel_for

BLOCKS = (64, 128, 256)
CELL_SCALES = (2.0, 1.0, 0.5, 0.25)
LANES = (1, 2, 4, 8, 16)


@dataclass(frozen=True)
class TuneResult:
    cell_scale: Optional[float]   # grid edge = voxel_for(h) * cell_scale; None = the params' voxel_size
    density_block: int
    force_block: int
    ms_per_step: float
    trials: Tuple[Tuple[Optional[float], int, int, int, float], ...] = field(default=())  # scale, blocks, lanes, ms
    seconds: float = 0.0          # time the search took (0 when read from the cache)
    cached: bool = False
    lanes: int = 1                # lanes per particle of the density and force passes

    @property
    def launch(self) -> Tuple[int, int]:
        return self.density_block, self.force_block

    def grid_voxel(self, h: float) -> Optional[float]:
        return None if self.cell_scale is None else voxel_for(h) * self.cell_scale

    def to_json(self) -> dict:
        d = asdict(self)
        d["trials"] = [list(t) for t in self.trials]
        return d

    @staticmethod
    def from_json(d: dict) -> "TuneResult":
        return TuneResult(d["cell_scale"], int(d["density_block"]), int(d["force_block"]), float(d["ms_per_step"]),
                          tuple(tuple(t) for t in d.get("trials", ())), 0.0, True, int(d.get("lanes", 1)))


def default_cache_path() -> str:
    return os.environ.get("SPH_AMD_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "sph_amd",
                                                                "tune.json")


class TuneCache:
    """JSON file {signature: TuneResult}; writes are atomic (rename)."""

    def __init__(self, path: Optional[str] = None):
        self.path = path or default_cache_path()

    def _load(self) -> Dict[str, dict]:
        try:
            with open(self.path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}

    def get(self, key: str) -> Optional[TuneResult]:
        d = self._load().get(key)
        return TuneResult.from_json(d) if d else None

    def put(self, key: str, res: TuneResult) -> None:
        data = self._load()
        data[key] = res.to_json()
        os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
        tmp = self.path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def _device_name(device) -> str:
    try:
        return torch.cuda.get_device_name(torch.device(device))
    except Exception:  # pragma: no cover - no device
        return "unknown"


def _fingerprint(parts: Iterable) -> str:
    h = hashlib.sha1()
    for p in parts:
        h.update(repr(p).encode())
    return h.hexdigest()[:16]


def _instance_key(p: SimulationParameters, c: SPHConstants, state: SimulationState) -> tuple:
    pos = np.asarray(state.position)
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    return (int(p.particle_count), round(float(c.h), 9), tuple(np.round(np.asarray(p.space_size, float), 9)),
            tuple(np.round(np.asarray(p.voxel_size, float), 9)), tuple(np.round(lo, 4)), tuple(np.round(hi, 4)))


def signature(kind: str, instances: Sequence[Tuple[SimulationParameters, SPHConstants, SimulationState]],
              mode: str, formulation: str, steps: int, device) -> str:
    parts = [kind, mode, formulation, steps, _device_name(device)]
    parts += [_instance_key(p, c, s) for p, c, s in instances]
    return f"{kind}:{formulation}:{mode}:{len(instances)}:{_fingerprint(parts)}"


def _time_steps(step, reset, steps: int, warmup: int) -> float:
    reset()
    step(warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    step(steps)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def _search(make, scales: Sequence[Optional[float]], blocks: Sequence[int], steps: int, warmup: int,
            log=None, lanes: Sequence[int] = (1,), finalists: int = 3) -> TuneResult:
    """make(scale) -> (configure(db, fb, lanes), step(k), reset(), close()) for one grid.
    First every (grid, lanes) pair is timed with the largest workgroups, then every block
    pair on the fastest of them; the best `finalists` are timed once more and the fastest of
    that second timing wins (the first timing of a run carries warm-up effects).  The very
    first timing is repeated and discarded for the same reason."""
    t0 = time.perf_counter()
    trials: List[Tuple[Optional[float], int, int, int, float]] = []
    warmed = [False]

    def run(scale, pairs, record):
        try:
            configure, step, reset, close = make(scale)
        except RuntimeError as e:  # e.g. a grid too fine for the planner's limits
            if log:
                log(f"autotune: cell scale {scale} rejected: {e}")
            return []
        out = []
        try:
            for db, fb, ln in pairs:
                try:
                    configure(db, fb, ln)
                except RuntimeError as e:  # e.g. too many threads for this lane count
                    if log:
                        log(f"autotune: lanes {ln} rejected: {e}")
                    continue
                if not warmed[0]:
                    _time_steps(step, reset, steps, warmup)
                    warmed[0] = True
                ms = _time_steps(step, reset, steps, warmup)
                out.append((scale, db, fb, ln, round(ms, 5)))
                if log:
                    log(f"autotune: scale {scale} blocks {db}/{fb} lanes {ln}: {ms:.4f} ms/step")
        finally:
            close()
        if record:
            trials.extend(out)
        return out

    top = max(blocks)
    for scale in scales:
        run(scale, [(top, top, ln) for ln in lanes], True)
    if not trials:
        raise RuntimeError("autotune: no candidate could be planned")
    s_best, _, _, l_best, _ = min(trials, key=lambda t: t[4])
    run(s_best, [(db, fb, l_best) for db in blocks for fb in blocks if (db, fb) != (top, top)], True)
    ranked = sorted(trials, key=lambda t: t[4])[:max(1, finalists)]
    final = []
    for scale in dict.fromkeys(t[0] for t in ranked):
        final += run(scale, [t[1:4] for t in ranked if t[0] == scale], False)
    best = min(final or ranked, key=lambda t: t[4])
    return TuneResult(best[0], best[1], best[2], best[4], tuple(trials), time.perf_counter() - t0, False, best[3])


def tune_strategy(params: SimulationParameters, constants: SPHConstants, state: SimulationState, mode: str = "BOX",
                  formulation: str = "NAIVE", steps: int = 20, warmup: int = 3,
                  scales: Sequence[float] = (1.0, 0.5), blocks: Sequence[int] = BLOCKS,
                  lanes: Sequence[int] = LANES, cache: Optional[TuneCache] = None, device="cuda:0",
                  log=None) -> TuneResult:
    """Tunes the plain step for one simulation.  VOXEL: block sizes only."""
    from .strategy import NaiveSPHStrategy, VoxelSPHStrategy

    key = signature("plain", [(params, constants, state)], mode, formulation, steps, device)
    if cache is not None:
        hit = cache.get(key)
        if hit is not None:
            return hit
    cls = NaiveSPHStrategy if formulation == "NAIVE" else VoxelSPHStrategy
    grid_scales: Sequence[Optional[float]] = list(scales) if formulation == "NAIVE" else [None]
    pos = torch.from_numpy(np.ascontiguousarray(state.position, dtype=np.float64)).to(device)
    vel = torch.from_numpy(np.ascontiguousarray(state.velocity, dtype=np.float64)).to(device)

    def make(scale):
        gv = None if scale is None else voxel_for(constants.h) * scale
        strat = cls(params, constants=constants, mode=mode, device=device, grid_voxel=gv)

        def configure(db, fb, ln):
            strat.ctx.set_launch(db, fb)
            strat.ctx.set_lanes(ln)
        return configure, strat.ctx.step, lambda: strat.ctx.upload(pos, vel), strat.ctx.close

    res = _search(make, grid_scales, blocks, steps, warmup, log, lanes if formulation == "NAIVE" else (1,))
    if cache is not None:
        cache.put(key, res)
    return res


def tune_batch(instances: Sequence[Tuple[SimulationParameters, SPHConstants]], states: Sequence[SimulationState],
               steps: int = 20, warmup: int = 3, scales: Sequence[float] = CELL_SCALES[:3],
               blocks: Sequence[int] = BLOCKS, lanes: Sequence[int] = LANES, cache: Optional[TuneCache] = None,
               device="cuda:0", log=None) -> TuneResult:
    """Tunes the batched form: one cell scale for all instances (each instance's grid edge
    is voxel_for(h_b) times it) and one block pair for the shared launches."""
    from .batched import BatchedSPH

    key = signature("batch", [(p, c, s) for (p, c), s in zip(instances, states)], "BOX", "NAIVE", steps, device)
    if cache is not None:
        hit = cache.get(key)
        if hit is not None:
            return hit

    def make(scale):
        b = BatchedSPH(instances, device=device, cell_scale=scale)
        b.upload(states)
        pos, vel = b.ctx_state_tensors()

        def configure(db, fb, ln):
            b.ctx.set_launch(db, fb)
            b.ctx.set_lanes(ln)
        return configure, b.step, lambda: b.ctx.upload(pos, vel), b.ctx.close

    res = _search(make, list(scales), blocks, steps, warmup, log, lanes)
    if cache is not None:
        cache.put(key, res)
    return res
