"""Autotuner for the neighbour kernels' launch shapes and the grid they walk.

Three knobs decide how the step maps onto the GPU:
  - the workgroup sizes of the density and force passes (64, 128 or 256 threads: one to
    four wavefronts over consecutive cell-ordered particles);
  - the lanes per particle of those passes (1 to 64 lanes of a wavefront share one
    particle's candidates: small or dense problems gain parallelism); and
  - the grid cell edge, which sets how many cells a workgroup's particles cover ("cells
    per tile") and how many candidates each particle tests.  Smaller cells test fewer
    candidates outside the h-sphere but walk more cell rows.  Only the NAIVE formulation
    may change it: its sums run over every j within h whatever the grid (results equal up
    to FP32 summation order), while in the VOXEL formulation the voxel edge is part of the
    reference's semantics (voxel_sph_strategy.py:109-116), so only block sizes are tuned.

The batched form tunes the grid and the lanes per instance: instances differ by orders of
magnitude in neighbours per particle (a dense one wants many lanes and finer cells, a sparse
one a single lane and coarse cells, whose table the counting sort scans every step).  Each
instance's neighbour load is estimated from one short run (the density over the particle
mass is the number density, times the volume of the h-sphere), and the candidates are
rules over that load: lanes = the load over a divisor (a power of two in 1..16), finer
cells above a load threshold.

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
from dataclasses import asdict, dataclass, field, replace
from typing import Dict, Iterable, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from .constants import SPHConstants
from .data_classes import SimulationParameters, SimulationState
from .workloads import voxel_for

BLOCKS = (64, 128, 256)
CELL_SCALES = (2.0, 1.0, 0.5, 0.25)
LANES = (1, 2, 4, 8, 16, 32, 64)


@dataclass(frozen=True)
class TuneResult:
    # grid edge = voxel_for(h) * cell_scale; None = the params' voxel_size.  Batched form: one
    # value per instance.
    cell_scale: Union[None, float, Tuple[float, ...]]
    density_block: int
    force_block: int
    ms_per_step: float
    # (scale, density block, force block, lanes, ms); the batched form records its rules
    # ("load>=T" for the grid, "load/D" for the lanes) in place of scale and lanes
    trials: Tuple[tuple, ...] = field(default=())
    seconds: float = 0.0          # time the search took (0 when read from the cache)
    cached: bool = False
    # lanes per particle of the density and force passes (batched form: one per instance)
    lanes: Union[int, Tuple[int, ...]] = 1
    rule: Tuple[str, ...] = ()    # batched form: the winning (grid, lanes) rules
    slabs: bool = True            # plain NAIVE step: cells cut into x slabs (sph_set_slabs)

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
        scale, lanes = d["cell_scale"], d.get("lanes", 1)
        scale = tuple(float(x) for x in scale) if isinstance(scale, list) else scale
        lanes = tuple(int(x) for x in lanes) if isinstance(lanes, list) else int(lanes)
        return TuneResult(scale, int(d["density_block"]), int(d["force_block"]), float(d["ms_per_step"]),
                          tuple(tuple(t) for t in d.get("trials", ())), 0.0, True, lanes,
                          tuple(d.get("rule", ())), bool(d.get("slabs", True)))


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


_LIBRARY_ID: Optional[str] = None


def _library_id() -> str:
    """Content hash of the built library: a rebuilt kernel set invalidates cached choices."""
    global _LIBRARY_ID
    if _LIBRARY_ID is None:
        from ._build import LIB_PATH

        try:
            with open(LIB_PATH, "rb") as f:
                _LIBRARY_ID = hashlib.sha1(f.read()).hexdigest()[:16]
        except OSError:
            _LIBRARY_ID = "missing"
    return _LIBRARY_ID


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
    parts = [kind, mode, formulation, steps, _device_name(device), _library_id()]
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


def _search(make, scales: Sequence, blocks: Sequence[int], steps: int, warmup: int,
            log=None, lanes: Sequence = (1,), finalists: int = 3, cross: bool = True,
            final_steps: Optional[int] = None) -> TuneResult:
    """make(scale) -> (configure(db, fb, lanes), step(k), reset(), close()) for one grid;
    scale and lanes are opaque keys.  With cross, every (grid, lanes) pair is timed with the
    largest workgroups first; otherwise every grid with the middle lanes key, then every
    lanes key on the fastest grid.  Then every block pair is timed on the fastest of those;
    the fastest trial of each of the `finalists` best lanes keys (and the runner-up block pair
    of the best) is timed once more, over final_steps steps (default: steps; a longer window
    than the search's when a workload's cost changes as it evolves), and the fastest of that
    second timing wins (the first
    timing of a run also carries warm-up effects).  The very first timing is repeated and
    discarded for the same reason."""
    t0 = time.perf_counter()
    trials: List[tuple] = []
    warmed = [False]

    def run(scale, pairs, record, nsteps=steps):
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
                    _time_steps(step, reset, nsteps, warmup)
                    warmed[0] = True
                ms = _time_steps(step, reset, nsteps, warmup)
                out.append((scale, db, fb, ln, round(ms, 5)))
                if log:
                    log(f"autotune: scale {scale} blocks {db}/{fb} lanes {ln}: {ms:.4f} ms/step")
        finally:
            close()
        if record:
            trials.extend(out)
        return out

    top = max(blocks)
    if cross:
        for scale in scales:
            run(scale, [(top, top, ln) for ln in lanes], True)
    else:
        mid = lanes[len(lanes) // 2]
        for scale in scales:
            run(scale, [(top, top, mid)], True)
        if trials:
            s_best = min(trials, key=lambda t: t[4])[0]
            run(s_best, [(top, top, ln) for ln in lanes if ln != mid], True)
    if not trials:
        raise RuntimeError("autotune: no candidate could be planned")
    s_best, _, _, l_best, _ = min(trials, key=lambda t: t[4])
    run(s_best, [(db, fb, l_best) for db in blocks for fb in blocks if (db, fb) != (top, top)], True)
    ranked = []
    for t in sorted(trials, key=lambda t: t[4]):
        if len(ranked) < max(1, finalists) and all(r[3] != t[3] for r in ranked):
            ranked.append(t)
    # the best lanes key's runner-up block pair too: block pairs a few tenths of a percent
    # apart in the short search trade places over the whole run
    runner = [t for t in sorted(trials, key=lambda t: t[4])
              if t[0] == ranked[0][0] and t[3] == ranked[0][3] and t[1:3] != ranked[0][1:3]]
    if runner:
        ranked.append(runner[0])
    final = []
    for scale in dict.fromkeys(t[0] for t in ranked):
        final += run(scale, [t[1:4] for t in ranked if t[0] == scale], False, final_steps or steps)
    best = min(final or ranked, key=lambda t: t[4])
    return TuneResult(best[0], best[1], best[2], best[4], tuple(trials), time.perf_counter() - t0, False, best[3])


def _choose_slabs(make, res: TuneResult, steps: int, warmup: int, log=None) -> TuneResult:
    """The winning configuration timed again with the cells cut into x slabs (the default)
    and searched whole; the faster is kept.  Cells of many particles (coarse grids, dense
    problems) gain little from the slabs and pay for ordering them."""
    configure, step, reset, close = make(res.cell_scale)
    ms = {}
    try:
        for slabs in (True, False):
            configure(res.density_block, res.force_block, res.lanes, slabs)
            ms[slabs] = _time_steps(step, reset, steps, warmup)
            if log:
                log(f"autotune: scale {res.cell_scale} blocks {res.density_block}/{res.force_block} "
                    f"lanes {res.lanes} slabs {'on' if slabs else 'off'}: {ms[slabs]:.4f} ms/step")
    finally:
        close()
    best = min(ms, key=ms.get)
    return replace(res, slabs=best, ms_per_step=round(ms[best], 5))


def tune_strategy(params: SimulationParameters, constants: SPHConstants, state: SimulationState, mode: str = "BOX",
                  formulation: str = "NAIVE", steps: int = 20, warmup: int = 3,
                  scales: Sequence[float] = (1.0, 0.5), blocks: Sequence[int] = BLOCKS,
                  lanes: Sequence[int] = LANES, cache: Optional[TuneCache] = None, device="cuda:0",
                  log=None, final_steps: Optional[int] = None, choose_slabs: bool = True) -> TuneResult:
    """Tunes the plain step for one simulation.  VOXEL: block sizes only.  final_steps: the
    window over which the finalists are re-timed (see _search).  choose_slabs: NAIVE, also
    decide whether the cells are cut into x slabs (False: they are, as in a context that
    cannot change it, one with SPH_FLAG_ADJOINT)."""
    from .strategy import NaiveSPHStrategy, VoxelSPHStrategy

    key = signature("plain", [(params, constants, state)], mode, formulation, (steps, final_steps, choose_slabs),
                    device)
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

        def configure(db, fb, ln, slabs=True):
            strat.ctx.set_launch(db, fb)
            strat.ctx.set_lanes(ln)
            if formulation == "NAIVE":
                strat.ctx.set_slabs(slabs)
        return configure, strat.ctx.step, lambda: strat.ctx.upload(pos, vel), strat.ctx.close

    res = _search(make, grid_scales, blocks, steps, warmup, log, lanes if formulation == "NAIVE" else (1,),
                  final_steps=final_steps)
    if formulation == "NAIVE" and choose_slabs:
        res = _choose_slabs(make, res, final_steps or steps, warmup, log)
    if cache is not None:
        cache.put(key, res)
    return res


# Batched-form rules over the neighbour load (mean neighbours within h per particle).
LOAD_THRESHOLDS = (0.0, 128.0, 512.0, 2048.0, float("inf"))  # finer cells from this load on
LOAD_DIVISORS = (4.0, 8.0, 16.0, 32.0, 64.0, 128.0)        # lanes = load / divisor
BATCH_SCALES = (1.0, 0.5)                                   # (coarse, fine) grid scales


def neighbour_load(instances: Sequence[Tuple[SimulationParameters, SPHConstants]],
                   states: Sequence[SimulationState], steps: int = 3, device="cuda:0") -> np.ndarray:
    """Mean neighbours within h per particle of each instance after `steps` batched steps,
    from the density: rho / m is the number density of the poly6 estimate, and a particle's
    neighbours are that times the volume of the sphere of radius h (itself included)."""
    from .batched import BatchedSPH

    b = BatchedSPH(instances, device=device)
    try:
        b.upload(states)
        b.step(steps)
        out = b.download()
    finally:
        b.ctx.close()
    load = []
    for st, (_, c) in zip(out, instances):
        rho = np.asarray(st.density, dtype=np.float64)
        rho = rho[np.isfinite(rho)]
        load.append(float(rho.mean()) / c.mass * (4.0 / 3.0) * np.pi * c.h ** 3 if rho.size else 0.0)
    return np.asarray(load)


def lanes_for_load(load: np.ndarray, divisor: float, cap: int = 16) -> Tuple[int, ...]:
    """Largest power of two in 1..cap not above load / divisor, per instance.  Beyond 16 lanes
    a group's lanes outnumber the candidates of most of its rows (measured on batch_64 over its
    first 100 steps: 64 lanes for the densest instances, 1.76 against 1.42 ms per step)."""
    out = []
    for x in np.asarray(load, dtype=np.float64):
        q = x / divisor if np.isfinite(x) else 1.0
        out.append(int(min(cap, 2 ** int(np.floor(np.log2(q))))) if q >= 1.0 else 1)
    return tuple(out)


def scales_for_load(load: np.ndarray, threshold: float, coarse: float = BATCH_SCALES[0],
                    fine: float = BATCH_SCALES[1]) -> Tuple[float, ...]:
    return tuple(fine if x >= threshold else coarse for x in np.asarray(load, dtype=np.float64))


def tune_batch(instances: Sequence[Tuple[SimulationParameters, SPHConstants]], states: Sequence[SimulationState],
               steps: int = 20, warmup: int = 3, thresholds: Sequence[float] = LOAD_THRESHOLDS,
               divisors: Sequence[float] = LOAD_DIVISORS, blocks: Sequence[int] = BLOCKS,
               cache: Optional[TuneCache] = None, device="cuda:0", log=None) -> TuneResult:
    """Tunes the batched form: per instance a grid scale and a lane count, both set by rules
    over the instance's neighbour load (finer cells from a load threshold on, lanes = load /
    divisor), and one block pair for the shared launches.  Thresholds are searched first (at
    the middle divisor), then divisors, then block pairs."""
    from .batched import BatchedSPH

    key = signature("batch", [(p, c, s) for (p, c), s in zip(instances, states)], "BOX", "NAIVE", steps, device)
    if cache is not None:
        hit = cache.get(key)
        if hit is not None:
            return hit
    t0 = time.perf_counter()
    load = neighbour_load(instances, states, device=device)
    if log:
        log(f"autotune: neighbour load per instance: min {load.min():.1f}, median {np.median(load):.1f}, "
            f"max {load.max():.1f}")
    scale_rule = {f"load>={t:g}": scales_for_load(load, t) for t in thresholds}
    lane_rule = {f"load/{d:g}": lanes_for_load(load, d) for d in divisors}

    def make(rule):
        b = BatchedSPH(instances, device=device, cell_scale=scale_rule[rule])
        b.upload(states)
        pos, vel = b.ctx_state_tensors()

        def configure(db, fb, ln):
            b.ctx.set_launch(db, fb)
            b.ctx.set_batch_lanes(lane_rule[ln])
        return configure, b.step, lambda: b.ctx.upload(pos, vel), b.ctx.close

    res = _search(make, list(scale_rule), blocks, steps, warmup, log, list(lane_rule), cross=False)
    res = TuneResult(scale_rule[res.cell_scale], res.density_block, res.force_block, res.ms_per_step, res.trials,
                     time.perf_counter() - t0, False, lane_rule[res.lanes], (res.cell_scale, res.lanes))
    if cache is not None:
        cache.put(key, res)
    return res
