#!/usr/bin/env python3
"""Throughput of the plain SPH step on one MI355X: particle-updates/s.

Workload: the 1,048,576-particle dam break (64 x 128 x 128 jittered lattice in a
[0,2]x[0,1]x[0,1] box, h = 0.01298, NAIVE formulation, box walls), state resident in HBM.
A step is one full advance: voxel hash, counting sort by cell (which builds the cell table
and gathers the state into cell order), density, force, integration, collision.  W
warm-up steps, then K timed steps (launched one by one, or replayed as one HIP graph with
--graph on) bracketed by device synchronisation and HIP events on the stream the kernels
run on.

Also measured in the same run:
  tuning        the autotuner's choice of grid cell edge and block sizes (before timing)
  stage timing  the same K steps again in one call, with HIP events on the kernel stream
                between the sort, density and force stages (sph_step_timed), so the
                stages are timed with the kernels and launch sequence of the timed region;
  roofline      algorithmic bytes per launch / event time of the dominant kernel (and of
                the force kernel, and per stage and per step in "kernels"/"step_roofline");
  repeats       the timed region repeated five times, and the median;
  pipe          the pipe workload (524,288 particles, 1000 steps);
  voxel_formulation  the reference's VOXEL formulation on the same dam break (500 steps);
  box_4k        the reference's CPU-sized case (4,096 particles, 100 steps) on the GPU;
  frames        frame output to disk in the reference's Saver format, synchronous
                (the reference's loop) against the asynchronous writer;
  batched       64 simulations per launch (HIP graph) against running them one by one;
  adjoint       the differentiable form: forward + backward of a 200-step rollout;
  cpu_baseline  the FP64 CPU oracle (OpenMP over the host cores) on a bounded sample of
                the same workload (the first steps until about 10 s have elapsed), and on
                box_4k with the O(N^2) pair loops for all 100 steps.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "cuda-sph_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak

# Algorithmic (compulsory) bytes per particle of each kernel, FP32 state:
#   force kernel (fused pressure + viscosity sums, integration, collision and the voxel hash
#   of the new position): read x, v (24 B), rho and p (8 B), particle id (4 B); write x, v
#   (24 B), rho (4 B), id (4 B), sort key and slot (8 B) -> 76 B.
FORCE_BYTES_PER_PARTICLE = 76
#   density kernel: read x (12 B), write rho and p (8 B) -> 20 B (1/rho is derived).
DENSITY_BYTES_PER_PARTICLE = 20
#   radix sort, per 8-bit pass: histogram read of the key (4 B), key + slot read and written
#   (16 B) -> 20 B; reorder: slot (4 B), x, v and id gathered (28 B) and written (28 B) -> 60 B;
#   cell table: sorted key (4 B) plus 4 B per cell of the table.
SORT_BYTES_PER_PASS = 20
REORDER_BYTES_PER_PARTICLE = 60
# FP32 operations per density candidate: 3 sub, 3 for r^2, 2 for (1 - r^2/h^2) clamped,
# 2 for the cube, 1 add.
DENSITY_OPS_PER_CANDIDATE = 11
# FP32 operations per list entry walked by the force kernel (3 sub, 3 for r^2, rsqrt, the
# pressure and viscosity coefficient chains, 6 accumulations).
FLOPS_PER_PAIR = 40


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dam_break_1m", choices=["dam_break_1m", "box_4k", "pipe_512k"])
    ap.add_argument("--voxel", type=float, nargs="+", default=None,
                    help="override the voxel edges of the workload (one value or x y z)")
    ap.add_argument("--block", type=int, nargs=2, default=None, metavar=("DENSITY", "FORCE"),
                    help="threads per block of the two neighbour kernels")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched", choices=["on", "off"], default="on",
                    help="also measure the batched form (64 box simulations, 300 steps, HIP graph)")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--tune", choices=["on", "off"], default="on",
                    help="autotune grid cell edge and block sizes before timing (skipped with --voxel/--block)")
    ap.add_argument("--frames", choices=["on", "off"], default="on",
                    help="also measure frame output to disk in the reference's format")
    ap.add_argument("--pipe", choices=["on", "off"], default="on",
                    help="also measure the pipe workload (524,288 particles, 1000 steps)")
    ap.add_argument("--voxel-form", choices=["on", "off"], default="on",
                    help="also measure the VOXEL formulation on the dam break (500 steps)")
    ap.add_argument("--small", choices=["on", "off"], default="on",
                    help="also measure box_4k (4,096 particles, 100 steps) on the GPU")
    ap.add_argument("--graph", choices=["on", "off"], default="off",
                    help="replay the timed steps of the plain and pipe legs as one HIP graph (default: launch "
                         "them one by one, which measures the same on these workloads)")
    ap.add_argument("--adjoint", choices=["on", "off"], default="on",
                    help="also measure the differentiable form (16,384 particles, 200 steps, forward + backward)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain device copy of known size (for counter calibration)")
    return ap.parse_args()


def tune_workload(w, mode="BOX", scales=(2.0, 1.0, 0.5), enabled=True, steps=None, choose_slabs=True):
    """Autotuned (params with the chosen grid, (density, force) block sizes, lanes per
    particle, summary with the slab choice) for a workload; the untuned defaults when
    disabled.  The grid edge is
    tried at 2, 1 and 1/2 times the largest power of two <= h (the coarse grid suits small
    problems with many lanes per particle: fewer cells to sort, fewer rows).  Candidates are
    timed over the first min(100, run length) steps and the best of each of the three best
    lane counts again over the whole run: the cost of a step changes as the workload evolves
    (a 20-step window picked 4 lanes for the pipe flow, which is slower than one over its
    1000 steps)."""
    import dataclasses

    import numpy as np

    if not enabled:
        return w.params, None, 1, None
    from sph_amd.autotune import tune_strategy

    steps = steps or min(100, int(w.steps))
    res = tune_strategy(w.params, w.constants, w.state, mode=mode, scales=scales, steps=steps,
                        final_steps=int(w.steps), log=log, choose_slabs=choose_slabs)
    gv = res.grid_voxel(w.constants.h)
    params = w.params if gv is None else dataclasses.replace(w.params, voxel_size=np.asarray([gv] * 3))
    info = {"cell_scale": res.cell_scale, "grid_voxel": gv, "density_block": res.density_block,
            "force_block": res.force_block, "lanes": res.lanes, "slabs": res.slabs, "trials": len(res.trials),
            "seconds": round(res.seconds, 2),
            "candidate_ms_per_step": [min(t[4] for t in res.trials), max(t[4] for t in res.trials)]}
    log(f"tuned {w.name}: {info}")
    return params, res.launch, res.lanes, info


def make_workload(name):
    from sph_amd import workloads

    return {"dam_break_1m": workloads.dam_break_1m, "box_4k": workloads.box_4k,
            "pipe_512k": workloads.pipe_flow}[name]()


def neighbour_stats(w, sample=2048, seed=0, slabs=4, field_max=6):
    """Per particle, from a sample (host): candidates the density pass tests, cell rows it
    walks and list entries the force pass walks (pairs within h, the particle included).
    The enumeration is the fused step's: for each row of cells (y, z) whose slab comes within
    h, the x slabs (`slabs` per cell) overlapping [x - e, x + e], e = sqrt(h^2 - dy^2 - dz^2),
    boundary cells extending to infinity; a cell at either end of the row whose slab offsets
    the table does not keep (more than field_max particles below its last slab) is searched
    whole."""
    import math

    import numpy as np

    pos = np.asarray(w.state.position, dtype=np.float64)
    h = float(w.constants.h)
    hs = h * (1.0 + 1e-5) + 4e-7 * float(np.max(w.params.space_size))
    v = np.asarray(w.params.voxel_size, dtype=np.float64)
    dims = np.maximum(np.ceil(np.asarray(w.params.space_size) / v).astype(np.int64), 1)
    cell = np.clip(np.floor(pos / v), 0, dims - 1).astype(np.int64)
    key = cell[:, 0] + dims[0] * (cell[:, 1] + dims[1] * cell[:, 2])
    sub = np.clip(np.floor(pos[:, 0] / v[0] * slabs) - slabs * cell[:, 0], 0, slabs - 1).astype(np.int64)
    gkey = key * slabs + sub  # the order of the fused step's sort, as one global slab index
    order = np.argsort(gkey, kind="stable")
    sg, spos = gkey[order], pos[order]
    below_last = np.bincount(key[sub < slabs - 1], minlength=int(np.prod(dims)))
    whole = below_last > field_max  # cells whose slab offsets are not kept

    def gc(x, d):
        return int(min(max(math.floor(x / v[d]), 0), dims[d] - 1))

    def gs(x):  # global x slab index, clamped to the row
        return int(min(max(math.floor(x / v[0] * slabs), 0), dims[0] * slabs - 1))

    def slab(p, c, d):
        lo = -math.inf if c == 0 else c * v[d]
        hi = math.inf if c == dims[d] - 1 else (c + 1) * v[d]
        return max(lo - p, p - hi, 0.0)

    rng = np.random.default_rng(seed)
    idx = rng.choice(len(pos), size=min(sample, len(pos)), replace=False)
    cand = rows = pairs = 0
    for i in idx:
        px, py, pz = pos[i]
        for cz in range(gc(pz - hs, 2), gc(pz + hs, 2) + 1):
            ez = slab(pz, cz, 2)
            for cy in range(gc(py - hs, 1), gc(py + hs, 1) + 1):
                ey = slab(py, cy, 1)
                rem = hs * hs - ez * ez - ey * ey
                if rem < 0:
                    continue
                e = math.sqrt(rem)
                row = dims[0] * (cy + dims[1] * cz)
                q0, q1 = gs(px - e), gs(px + e) + 1
                c0, c1 = row + q0 // slabs, row + (q1 - 1) // slabs
                lo = c0 * slabs if whole[c0] else row * slabs + q0
                hi = (c1 + 1) * slabs if whole[c1] else row * slabs + q1
                a, b = np.searchsorted(sg, [lo, hi])
                cand += b - a
                rows += 1
                pairs += int((((spos[a:b] - pos[i]) ** 2).sum(-1) <= h * h).sum())
    return cand / len(idx), rows / len(idx), pairs / len(idx)


def cpu_baseline(w, budget_s: float):
    import oracle

    oracle.build()
    threads = oracle.num_threads()
    pos, vel = w.state.position, w.state.velocity
    steps = 0
    t0 = time.perf_counter()
    while True:
        pos, vel, _, _, _ = oracle.step(pos, vel, consts=w.constants, ext=w.params.external_force,
                                        dt=1.0 / w.params.fps, space=w.params.space_size, voxel=w.params.voxel_size,
                                        use_grid=True)
        steps += 1
        el = time.perf_counter() - t0
        log(f"cpu baseline: step {steps} at {el:.1f} s")
        if el >= budget_s or steps >= 200:
            break
    n = len(pos)
    out = {"value": n * steps / el, "unit": "particle-updates/s", "cores": threads, "kind": "port",
           "sample": f"first {steps} steps of {w.name} ({n} particles), FP64 C restatement of the reference "
                     f"step (NAIVE formulation through a grid search), OpenMP, {el:.1f} s"}
    # the reference's own CPU-sized case (box_4k, 100 steps) with the O(N^2) pair loops of
    # naive_kernels.py, every step
    from sph_amd import workloads

    s = workloads.box_4k()
    p, v = s.state.position, s.state.velocity
    t0 = time.perf_counter()
    for _ in range(s.steps):
        p, v, _, _, _ = oracle.step(p, v, consts=s.constants, ext=s.params.external_force, dt=1.0 / s.params.fps,
                                    space=s.params.space_size, voxel=s.params.voxel_size, use_grid=False)
    el = time.perf_counter() - t0
    out["box_4k"] = {"value": round(len(p) * s.steps / el, 1), "unit": "particle-updates/s", "cores": threads,
                     "kind": "port", "sample": f"all {s.steps} steps of box_4k, O(N^2) pair loops, {el:.2f} s"}
    log(f"cpu baseline box_4k: {el:.2f} s")
    return out


def bench_frames(n: int = 65536, frames: int = 50):
    """Frame output in the reference's file format (a Saver folder of .npy frames), every
    step saved: the reference's loop (StateGenerator + Saver: one synchronous step, copy
    and write per frame) against run_to_disk (frames gathered on the GPU, copied on a
    second stream and written by a thread while the GPU steps on)."""
    import itertools
    import shutil
    import tempfile

    from sph_amd import workloads
    from sph_amd.serializer import Saver, run_to_disk
    from sph_amd.state_generator import StateGenerator
    from sph_amd.strategy import NaiveSPHStrategy

    w = workloads.box_uniform(n, seed=5)
    params = w.params  # dt = 1 / fps = 1e-3; the first `frames` frames are written
    root = tempfile.mkdtemp(prefix="sph_frames_")
    try:
        gen = StateGenerator(w.state, params, constants=w.constants)
        saver = Saver("sync", params, root=root)
        t0 = time.perf_counter()
        for state in itertools.islice(gen, frames):
            saver.save_next_state(state)
        t_sync = time.perf_counter() - t0
        strat = NaiveSPHStrategy(params, constants=w.constants, check_finite=False)
        t0 = time.perf_counter()
        run_to_disk(strat, w.state, params, "async", root=root, frames=frames)
        t_async = time.perf_counter() - t0
    finally:
        shutil.rmtree(root, ignore_errors=True)
    mb = n * 56 / 1e6
    log(f"frames: sync {t_sync:.2f} s, async {t_async:.2f} s for {frames} frames of {mb:.1f} MB")
    return {"workload": f"box_uniform_{n}", "particles": n, "frames": frames, "mb_per_frame": round(mb, 2),
            "sync_frames_per_s": round(frames / t_sync, 1), "async_frames_per_s": round(frames / t_async, 1),
            "speedup": round(t_sync / t_async, 2)}


def bench_small(warmup: int = 3, tune: bool = True):
    """The reference's CPU-sized case (box_4k: 4,096 particles, 100 steps) on the GPU."""
    import numpy as np
    import torch

    from sph_amd import workloads
    from sph_amd.strategy import NaiveSPHStrategy

    w = workloads.box_4k()
    n = w.params.particle_count
    params, launch, lanes, tuning = tune_workload(w, enabled=tune)
    strat = NaiveSPHStrategy(params, constants=w.constants, launch=launch, lanes=lanes, check_finite=False,
                             slabs=tuning["slabs"] if tuning else True)
    pos = torch.from_numpy(np.ascontiguousarray(w.state.position)).cuda()
    vel = torch.from_numpy(np.ascontiguousarray(w.state.velocity)).cuda()
    strat.ctx.upload(pos, vel)
    strat.ctx.step(warmup)
    graph = strat.capture(w.steps)  # the 100 steps as one HIP graph (recorded, not run)
    strat.ctx.upload(pos, vel)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    graph.replay()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b)
    return {"workload": "box_4k", "particles": n, "steps": w.steps, "value": round(n * w.steps / (ms / 1e3), 1),
            "unit": "particle-updates/s", "ms_per_step": round(ms / w.steps, 5), "graph": True, "tuning": tuning}


def bench_pipe(steps: int = 1000, warmup: int = 5, tune: bool = True, graph: bool = True):
    """The pipe workload (pipe_512k: 8 segments, inflow, re-injection through the seeded
    xoroshiro128+ streams): particle-updates/s over its 1000 steps (one HIP graph with
    graph)."""
    import numpy as np
    import torch

    from sph_amd import workloads
    from sph_amd.strategy import NaiveSPHStrategy

    w = workloads.pipe_flow()
    n = w.params.particle_count
    params, launch, lanes, tuning = tune_workload(w, "PIPE", enabled=tune)
    strat = NaiveSPHStrategy(params, constants=w.constants, mode="PIPE", launch=launch, lanes=lanes,
                             check_finite=False, slabs=tuning["slabs"] if tuning else True)
    pos = torch.from_numpy(np.ascontiguousarray(w.state.position)).cuda()
    vel = torch.from_numpy(np.ascontiguousarray(w.state.velocity)).cuda()
    run = timed_steps(strat, steps, graph, pos, vel)
    strat.ctx.upload(pos, vel)
    strat.ctx.step(warmup)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    run()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b)
    d = strat.diagnostics()
    log(f"pipe: {steps} steps of {n} particles in {ms:.1f} ms")
    return {"workload": "pipe_512k", "particles": n, "steps": steps, "value": round(n * steps / (ms / 1e3), 1),
            "unit": "particle-updates/s", "ms_per_step": round(ms / steps, 5), "h": w.constants.h,
            "non_finite": d["non_finite"], "graph": graph, "tuning": tuning}


def bench_voxel_form(steps: int = 500, warmup: int = 5):
    """The reference's VOXEL formulation (voxel_sph_strategy.py: stable radix sort of the
    voxel keys, neighbour lists of at most 32 from the 27 voxels, its own density and force
    formulas) on dam_break_1m with voxels of 2^-6 (128 x 64 x 64): its 27-voxel search
    reaches h only when the voxel edge is at least h (0.01298 here).  Particle-updates/s over
    500 steps launched one by one."""
    import dataclasses

    import numpy as np
    import torch

    from sph_amd import workloads
    from sph_amd.strategy import VoxelSPHStrategy

    w = workloads.dam_break_1m()
    w.params = dataclasses.replace(w.params, voxel_size=np.asarray([2.0 ** -6] * 3))
    n = w.params.particle_count
    strat = VoxelSPHStrategy(w.params, constants=w.constants, check_finite=False)
    pos = torch.from_numpy(np.ascontiguousarray(w.state.position)).cuda()
    vel = torch.from_numpy(np.ascontiguousarray(w.state.velocity)).cuda()
    strat.ctx.upload(pos, vel)
    strat.ctx.step(warmup)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    strat.ctx.step(steps)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b)
    d = strat.diagnostics()
    log(f"voxel formulation: {steps} steps of {n} particles in {ms:.1f} ms")
    out = {"workload": "dam_break_1m", "formulation": "VOXEL", "particles": n, "steps": steps,
           "voxel": [float(x) for x in w.params.voxel_size], "value": round(n * steps / (ms / 1e3), 1),
           "unit": "particle-updates/s", "ms_per_step": round(ms / steps, 5), "non_finite": d["non_finite"],
           "state_went_non_finite": d["non_finite"] > 0}
    if d["non_finite"]:
        # the reference's VOXEL density leaves the particle itself out (voxel_kernels.py:127-130):
        # a particle that ends with no neighbour has rho = 0 and the integration divides by it
        out["note"] = (f"{d['non_finite']} particles ended NaN or infinite (an isolated particle has rho = 0 in this "
                       "formulation); the rate is that of a run whose state went non-finite")
    return out


def launches_per_step(ctx) -> int:
    """Kernel launches of one step of a context: the nodes of a HIP graph recorded from
    ctx.step(1) (recorded, never run)."""
    import ctypes

    import torch

    g = torch.cuda.CUDAGraph(keep_graph=True)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        ctx.step(1)
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime torch has loaded (same soname)
    count = ctypes.c_size_t(0)
    rc = hip.hipGraphGetNodes(ctypes.c_void_p(g.raw_cuda_graph()), None, ctypes.byref(count))
    del g
    return int(count.value) if rc == 0 else -1


def timed_steps(strat, steps: int, graph: bool, pos, vel):
    """A callable that runs `steps` steps of the resident state: one HIP graph replay (the
    graph is recorded here from the uploaded state, not run) or direct launches."""
    if not graph:
        return lambda: strat.ctx.step(steps)
    strat.ctx.upload(pos, vel)
    g = strat.capture(steps)
    return g.replay


def bench_batched(steps: int, warmup: int = 3, tune: bool = True):
    """Aggregate particle-updates/s of the batch_64 workload: one launch per stage for all 64
    instances, the timed steps replayed as one HIP graph; against the same 64 simulations
    advanced one after another with the plain step (one context each, the best of the
    default, the batch's and the instance's own autotuned configuration, each launched step
    by step and as one HIP graph).  With tune, each instance's grid scale and lanes per
    particle and the shared block sizes come from the autotuner."""
    import numpy as np
    import torch

    from sph_amd import workloads
    from sph_amd.batched import BatchedSPH
    from sph_amd.strategy import NaiveSPHStrategy

    bw = workloads.batch_64(steps=steps)
    total = bw.total_particles
    states = [w.state for w in bw.instances]
    inst = [(w.params, w.constants) for w in bw.instances]
    count = len(bw.instances)
    scale, launch, lanes, tuning = [None] * count, None, [1] * count, None
    if tune:
        from collections import Counter

        from sph_amd.autotune import tune_batch

        log("autotuning the batch")
        # candidates are timed over the whole run: its cost falls several-fold as the
        # instances' dense starting blocks spread, so a short window misjudges the rules
        res = tune_batch(inst, states, steps=steps, log=log)
        scale, launch, lanes = list(res.cell_scale), res.launch, list(res.lanes)
        tuning = {"rule": list(res.rule), "density_block": res.density_block, "force_block": res.force_block,
                  "instances_per_lanes": {str(k): v for k, v in sorted(Counter(lanes).items())},
                  "instances_per_cell_scale": {str(k): v for k, v in sorted(Counter(scale).items())},
                  "trials": len(res.trials), "seconds": round(res.seconds, 2),
                  "candidate_ms_per_step": [min(t[4] for t in res.trials), max(t[4] for t in res.trials)]}
        log(f"batch tuned: {tuning}")
    batch = BatchedSPH(inst, cell_scale=scale, launch=launch, lanes=lanes)
    batch.upload(states)
    launches = launches_per_step(batch.ctx)
    log(f"batched step: {launches} kernel launches for {count} instances")
    graph = batch.capture(steps)  # recorded, not run
    batch.upload(states)
    batch.step(warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    graph.replay()
    e1.record()
    torch.cuda.synchronize()
    t_batch = e0.elapsed_time(e1) / 1e3
    log(f"batched: {steps} steps of {total} particles in {t_batch * 1e3:.1f} ms")
    # the baseline gets its best case: each instance alone, with the plain step's default grid
    # and blocks, with the batch's choices for that instance and (with tune) with the
    # instance's own autotuned plain configuration, its steps launched one by one and replayed
    # as one HIP graph; the fastest counted
    t_seq = 0.0
    per_instance = []
    plain_launches = None
    for b, w in enumerate(bw.instances):
        configs = [(None, None, 1, True)]
        if tune:
            from sph_amd.autotune import TuneCache, tune_strategy

            own = tune_strategy(w.params, w.constants, w.state, cache=TuneCache())
            configs += [(scale[b], launch, lanes[b], True), (own.cell_scale, own.launch, own.lanes, own.slabs)]
        pos = torch.from_numpy(np.ascontiguousarray(w.state.position)).cuda()
        vel = torch.from_numpy(np.ascontiguousarray(w.state.velocity)).cuda()
        best = None
        for sc, la, ln, sl in dict.fromkeys(configs):
            gv = None if sc is None else workloads.voxel_for(w.constants.h) * sc
            strat = NaiveSPHStrategy(w.params, constants=w.constants, grid_voxel=gv, launch=la, lanes=ln,
                                     check_finite=False, slabs=sl)
            if plain_launches is None:
                strat.ctx.upload(pos, vel)
                plain_launches = launches_per_step(strat.ctx)
            for graph in (False, True):
                run = timed_steps(strat, steps, graph, pos, vel)
                strat.ctx.upload(pos, vel)
                strat.ctx.step(warmup)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                run()
                e1.record()
                torch.cuda.synchronize()
                t = e0.elapsed_time(e1) / 1e3
                best = t if best is None else min(best, t)
                del run
            del strat
        t_seq += best
        per_instance.append((best * 1e3 / steps, w.params.particle_count, w.name))
    log(f"sequential plain runs: {t_seq * 1e3:.1f} ms")
    slowest = sorted(per_instance, reverse=True)[:5]
    log("slowest instances alone (ms/step, particles): " + ", ".join(f"{ms:.3f} ({n})" for ms, n, _ in slowest))
    return {"workload": "batch_64", "instances": len(bw.instances), "particles": total, "steps": steps,
            "value": round(total * steps / t_batch, 1), "unit": "particle-updates/s",
            "ms_per_step": round(t_batch * 1e3 / steps, 4), "graph": True,
            "sequential_plain_value": round(total * steps / t_seq, 1),
            "speedup_vs_sequential": round(t_seq / t_batch, 3), "tuning": tuning,
            "launches_per_step": launches, "sequential_launches_per_step": count * plain_launches,
            "slowest_instance_alone_ms_per_step": round(slowest[0][0], 4)}


def bench_adjoint(steps: int = 200, every: int = 20, tune: bool = True):
    """The differentiable form on the adjoint_16k workload: loss = mean |x_T - target|^2 with
    target = the rollout at mu = 5 plus N(0, 0.01^2) noise (both from the GPU step), gradients
    with respect to the initial velocities and mu through the adjoint kernels, checkpoints
    every `every` steps.  Rate = N * steps / (forward + backward time)."""
    import numpy as np
    import torch

    from sph_amd import workloads
    from sph_amd.differentiable import DifferentiableSPH

    w = workloads.adjoint_16k()
    n = w.params.particle_count
    # the grid stays the workload's; an adjoint context keeps its cells cut into slabs
    _, launch, lanes, tuning = tune_workload(w, scales=(1.0,), enabled=tune, choose_slabs=False)
    sim = DifferentiableSPH(w.params, w.constants, steps=steps, checkpoint_every=every, launch=launch, lanes=lanes)
    x0 = torch.from_numpy(w.state.position.astype(np.float32)).cuda()
    v0 = torch.from_numpy(w.state.velocity.astype(np.float32)).cuda()
    x5, _, _ = sim.forward(x0, v0, 5.0)
    noise = torch.from_numpy(np.random.default_rng(44).normal(0.0, 0.01, size=(n, 3)).astype(np.float32)).cuda()
    target = x5 + noise
    mu = float(w.constants.visc)

    def run():
        xT, _, tape = sim.forward(x0, v0, mu)
        b.record()
        gx = 2.0 * (xT - target) / n
        loss = ((xT.double() - target.double()) ** 2).sum(1).mean()
        c.record()
        grads = sim.backward(tape, gx, None)
        d.record()
        return loss, grads

    a, b, c, d = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def timed():
        run()  # warm-up: direct launches (a new viscosity)
        run()  # warm-up: records the HIP graphs of this viscosity
        torch.cuda.synchronize()
        a.record()
        out = run()
        torch.cuda.synchronize()
        return out, a.elapsed_time(b) / 1e3, c.elapsed_time(d) / 1e3

    loss, (gx0, gv0, gmu) = timed()[0]
    _, t_f, t_b = timed()
    log(f"adjoint: forward {t_f * 1e3:.1f} ms, backward {t_b * 1e3:.1f} ms for {steps} steps of {n} particles"
        f" (every step's snapshot kept: {sim.keep_all})")
    # the configuration as written: checkpoints every `every` steps, each segment recomputed
    # from its checkpoint in the backward pass
    keep_all = sim.keep_all
    sim.set_keep_all(False)
    (_, (_, gv0c, gmuc)), t_fc, t_bc = timed()
    sim.set_keep_all(keep_all)
    log(f"adjoint, checkpoints every {every}: forward {t_fc * 1e3:.1f} ms, backward {t_bc * 1e3:.1f} ms")
    same = bool(torch.equal(gv0c, gv0) and torch.equal(gmuc, gmu))
    return {"workload": "adjoint_16k", "particles": n, "steps": steps, "checkpoint_every": every,
            "keeps_every_step": keep_all, "snapshot_mb": round(sim.ctx.adjoint_snapshot_bytes() / 1e6, 2),
            "value": round(n * steps / (t_f + t_b), 1), "unit": "particle-updates/s",
            "ms_forward": round(t_f * 1e3, 3), "ms_backward": round(t_b * 1e3, 3),
            "ms_per_adjoint_step": round(t_b * 1e3 / steps, 4),
            "checkpointed_value": round(n * steps / (t_fc + t_bc), 1),
            "checkpointed_ms_forward": round(t_fc * 1e3, 3), "checkpointed_ms_backward": round(t_bc * 1e3, 3),
            "checkpointed_grads_bitwise_equal": same,
            "loss": float(loss), "grad_mu": float(gmu.item()), "grad_v0_norm": float(gv0.double().norm()),
            "finite": bool(torch.isfinite(gv0).all() and torch.isfinite(gmu).all()), "tuning": tuning}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the SPH step has no "
              "multi-GPU path in this build)", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from sph_amd.strategy import NaiveSPHStrategy

    dev = torch.device("cuda:0")
    log(f"building workload {args.workload}")
    w = make_workload(args.workload)
    if args.voxel:
        import dataclasses

        vox = args.voxel * 3 if len(args.voxel) == 1 else args.voxel
        w.params = dataclasses.replace(w.params, voxel_size=np.asarray(vox, dtype=np.float64))
    n = w.params.particle_count
    tune = args.tune == "on" and not args.voxel and not args.block
    w.params, launch, lanes, tuning = tune_workload(w, w.mode, enabled=tune)
    if args.block:
        launch = tuple(args.block)
    strat = NaiveSPHStrategy(w.params, constants=w.constants, mode=w.mode, device=dev, launch=launch, lanes=lanes,
                             check_finite=False, slabs=tuning["slabs"] if tuning else True)
    pos = torch.from_numpy(np.ascontiguousarray(w.state.position)).to(dev)
    vel = torch.from_numpy(np.ascontiguousarray(w.state.velocity)).to(dev)
    graph = args.graph == "on"
    run = timed_steps(strat, args.steps, graph, pos, vel)
    strat.ctx.upload(pos, vel)
    log(f"warmup {args.warmup} steps")
    strat.ctx.step(args.warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    run()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    diag = strat.diagnostics()
    log(f"timed: {gpu_ms:.2f} ms on the GPU, {wall * 1e3:.2f} ms wall; diagnostics {diag}")
    ms_per_step = gpu_ms / max(args.steps, 1)
    value = n * args.steps / (gpu_ms / 1e3)

    # per-kernel timing over the same trajectory: the step is deterministic, so replaying
    # the upload, the warm-up and the K steps (one sph_step_timed call: the timed region's
    # kernels, with HIP events between the stages) revisits exactly the timed states
    log(f"stage timing over the same {args.steps} steps")
    strat.ctx.upload(pos, vel)
    strat.ctx.step(args.warmup)
    st = strat.ctx.step_timed(args.steps)
    cand, rows, pairs = neighbour_stats(w)
    grid = strat.ctx.grid_info()
    n_cells = int(np.prod(grid["dims"]))
    # the sort stage: counting-sort passes, the cell table and the gather into cell order
    per_particle = {"sort": SORT_BYTES_PER_PASS * grid["passes"] + 4 + 4 * n_cells / n + REORDER_BYTES_PER_PARTICLE,
                    "density": DENSITY_BYTES_PER_PARTICLE, "force": FORCE_BYTES_PER_PARTICLE}
    kernels = {}
    for k, ms in st.items():
        gbs = per_particle[k] * n / (ms / 1e3) / 1e9
        kernels[k] = {"ms": round(ms, 4), "bytes_per_particle": round(per_particle[k], 2),
                      "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    step_bytes = sum(per_particle.values())
    step_gbs = step_bytes * n / (ms_per_step / 1e3) / 1e9
    step_roofline = {"bytes_per_particle": round(step_bytes, 2), "ms": round(ms_per_step, 5),
                     "achieved_gbs": round(step_gbs, 1), "frac": round(step_gbs / HBM_PEAK_GBS, 4),
                     "hbm_floor_ms": round(step_bytes * n / HBM_PEAK_GBS / 1e6, 4)}
    traffic = {}
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        traffic = json.load(open(traffic_file)).get(args.workload, {})
    valu_ops = {"density": n * cand * DENSITY_OPS_PER_CANDIDATE, "force": n * pairs * FLOPS_PER_PAIR}

    # the timed force passes count the next sort's keys (all but the last of the call)
    traffic_key = {"naive_density_kernel": "naive_density_kernel", "naive_force_kernel": "naive_force_kernel_counting"}

    def kernel_roofline(stage, name):
        ms = st[stage]
        achieved = per_particle[stage] * n / (ms / 1e3) / 1e9
        return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": (traffic.get(traffic_key[name]) or {}).get("bytes_per_launch"),
                "bytes_per_particle": per_particle[stage], "kernel_ms": round(ms, 4),
                "valu": {"achieved_tflops": round(valu_ops[stage] / (ms / 1e3) / 1e12, 2),
                         "peak_tflops": FP32_VECTOR_PEAK_TFLOPS, "candidates_per_particle": round(cand, 1),
                         "rows_per_particle": round(rows, 1), "list_entries_per_particle": round(pairs, 1)}}

    # the roofline object is the dominant (longest) kernel's; the force kernel is reported
    # beside it as well
    dominant = max(("density", "force"), key=lambda k: st[k])
    names = {"density": "naive_density_kernel", "force": "naive_force_kernel"}
    roofline = kernel_roofline(dominant, names[dominant])
    roofline_force = kernel_roofline("force", "naive_force_kernel")

    # repeats of the timed region (upload, warm-up, K steps), for the median
    repeats = []
    for _ in range(5):
        strat.ctx.upload(pos, vel)
        strat.ctx.step(args.warmup)
        torch.cuda.synchronize()
        e0.record(stream)
        run()
        e1.record(stream)
        torch.cuda.synchronize()
        repeats.append(round(e0.elapsed_time(e1) / max(args.steps, 1), 5))
    log(f"repeats (ms/step): {repeats}")

    if args.calibrate_copy:
        src = torch.empty(256 * 2 ** 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 5 copies of 256 MiB")

    pipe = bench_pipe(tune=args.tune == "on", graph=graph) if args.pipe == "on" else None
    voxel_form = bench_voxel_form() if args.voxel_form == "on" else None
    frames = bench_frames() if args.frames == "on" else None
    small = bench_small(tune=args.tune == "on") if args.small == "on" else None
    batched = bench_batched(args.batch_steps, tune=args.tune == "on") if args.batched == "on" else None
    adjoint = bench_adjoint(tune=args.tune == "on") if args.adjoint == "on" else None

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(w, args.cpu_seconds)

    line = {
        "metric": "particle-updates/s/GPU", "value": round(value, 1), "unit": "particle-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": args.workload, "particles": n, "form": "plain", "formulation": "NAIVE",
                   "mode": w.mode, "h": w.constants.h, "voxel": [float(v) for v in np.asarray(w.params.voxel_size)],
                   "parallelism": "single GPU", "graph": graph},
        "roofline": roofline,
        "roofline_force": roofline_force,
        "kernels": kernels,
        "step_roofline": step_roofline,
        "repeat_ms_per_step": repeats,
        "median_ms_per_step": float(np.median(repeats)),
        "cpu_baseline": cpu,
        "tuning": tuning,
        "pipe": pipe,
        "voxel_formulation": voxel_form,
        "box_4k": small,
        "frames": frames,
        "stages_ms": {k: round(v, 4) for k, v in st.items()},
        "batched": batched,
        "adjoint": adjoint,
        "wall_ms_per_step": round(wall * 1e3 / max(args.steps, 1), 5),
        "final_state": {"non_finite": diag["non_finite"], "max_speed": float(diag["max_speed"]),
                        "mean_density": float(diag["mean_density"])},
    }
    print(json.dumps(line), flush=True)
    return 0


def reap_children() -> None:
    """Ends and waits for any child process this run left (none is started on purpose), so
    that nothing outlives the bench."""
    try:
        import psutil
    except ImportError:
        return
    kids = psutil.Process().children(recursive=True)
    for k in kids:
        try:
            k.terminate()
        except psutil.Error:
            pass
    gone, alive = psutil.wait_procs(kids, timeout=5)
    for k in alive:
        try:
            k.kill()
        except psutil.Error:
            pass
    if kids:
        log(f"reaped {len(kids)} child process(es)")


if __name__ == "__main__":
    try:
        rc = main()
    finally:
        reap_children()
    sys.exit(rc)
