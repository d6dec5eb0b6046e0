#!/usr/bin/env python3
"""Throughput of the plain SPH step on one MI355X: particle-updates/s.

Workload: the 1,048,576-particle dam break (64 x 128 x 128 jittered lattice in a
[0,2]x[0,1]x[0,1] box, h = 0.01298, NAIVE formulation, box walls), state resident in HBM.
A step is one full advance: voxel hash, radix sort, cell table, density, force,
integration, collision.  W warm-up steps, then K timed steps bracketed by device
synchronisation and HIP events on the stream the kernels run on.

Also measured in the same run:
  stage timing  events around each stage of a few staged steps (the force kernel is one
                launch, so its event time is that kernel's duration);
  roofline      algorithmic bytes of the force kernel per launch / its event time;
  cpu_baseline  the FP64 CPU oracle (grid search, OpenMP over the host cores) on a bounded
                sample of the same workload: the first steps until about 10 s have elapsed.
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
#   density kernel: read x (12 B), write rho, p, 1/rho (12 B) -> 24 B.
DENSITY_BYTES_PER_PARTICLE = 24
# FP32 operations per candidate pair test (3 sub, 3 mul/fma, compare) and per interacting
# pair in the force kernel (rsqrt, 2 pressure / viscosity coefficient chains, 6 fma, 3 sub).
FLOPS_PER_CANDIDATE = 9
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
    ap.add_argument("--adjoint", choices=["on", "off"], default="on",
                    help="also measure the differentiable form (16,384 particles, 200 steps, forward + backward)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain device copy of known size (for counter calibration)")
    return ap.parse_args()


def make_workload(name):
    from sph_amd import workloads

    return {"dam_break_1m": workloads.dam_break_1m, "box_4k": workloads.box_4k,
            "pipe_512k": workloads.pipe_flow}[name]()


def stage_times(strat, nsteps: int):
    """Mean milliseconds of each stage over nsteps staged steps (events on the kernel stream)."""
    import torch

    from sph_amd import _native as N

    names = {N.STAGE_SORT: "sort", N.STAGE_TABLE: "cell_table", N.STAGE_REORDER: "reorder",
             N.STAGE_DENSITY: "density", N.STAGE_FORCE: "force"}
    acc = {v: 0.0 for v in names.values()}
    stream = torch.cuda.current_stream()
    done = 0
    while done < nsteps:
        chunk = min(50, nsteps - done)
        evs = []
        for _ in range(chunk):
            for st in names:
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                strat.ctx.run_stage(st)
                b.record(stream)
                evs.append((names[st], a, b))
        torch.cuda.synchronize()
        for name, a, b in evs:
            acc[name] += a.elapsed_time(b)
        done += chunk
    return {k: v / max(nsteps, 1) for k, v in acc.items()}


def neighbour_stats(w, sample=4096, seed=0):
    """Mean candidates tested and interacting pairs per particle, from a sample (host)."""
    import numpy as np

    pos = np.asarray(w.state.position)
    h = w.constants.h
    vox = float(np.asarray(w.params.voxel_size)[0])
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(pos), size=min(sample, len(pos)), replace=False)
    cell = np.floor(pos / vox).astype(np.int64)
    key = cell[:, 0] + (1 << 20) * (cell[:, 1] + (1 << 20) * cell[:, 2])
    order = np.argsort(key, kind="stable")
    skey = key[order]
    cand = pairs = 0
    R = int(np.ceil(h / vox))
    for i in idx:
        c = cell[i]
        for dz in range(-R, R + 1):
            for dy in range(-R, R + 1):
                k0 = (c[0] - R) + (1 << 20) * ((c[1] + dy) + (1 << 20) * (c[2] + dz))
                a, b = np.searchsorted(skey, [k0, k0 + 2 * R + 1])
                if b > a:
                    cand += b - a
                    d2 = ((pos[order[a:b]] - pos[i]) ** 2).sum(-1)
                    pairs += int((d2 <= h * h).sum())
        pairs -= 1  # self
    return cand / len(idx), pairs / len(idx)


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
    return {"value": n * steps / el, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"first {steps} steps of {w.name} ({n} particles), FP64 C restatement of the reference "
                      f"step (NAIVE formulation through a grid search), OpenMP, {el:.1f} s"}


def bench_batched(steps: int, warmup: int = 3):
    """Aggregate particle-updates/s of the batch_64 workload: one launch per stage for all 64
    instances, the timed steps replayed as one HIP graph; against the same 64 simulations
    advanced one after another with the plain step (same kernels, one context each)."""
    import numpy as np
    import torch

    from sph_amd import workloads
    from sph_amd.batched import BatchedSPH
    from sph_amd.strategy import NaiveSPHStrategy

    bw = workloads.batch_64(steps=steps)
    total = bw.total_particles
    states = [w.state for w in bw.instances]
    batch = BatchedSPH([(w.params, w.constants) for w in bw.instances])
    batch.upload(states)
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
    t_seq = 0.0
    for w in bw.instances:
        strat = NaiveSPHStrategy(w.params, constants=w.constants)
        strat.ctx.upload(torch.from_numpy(np.ascontiguousarray(w.state.position)).cuda(),
                         torch.from_numpy(np.ascontiguousarray(w.state.velocity)).cuda())
        strat.ctx.step(warmup)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        strat.ctx.step(steps)
        b.record()
        torch.cuda.synchronize()
        t_seq += a.elapsed_time(b) / 1e3
        del strat
    log(f"sequential plain runs: {t_seq * 1e3:.1f} ms")
    return {"workload": "batch_64", "instances": len(bw.instances), "particles": total, "steps": steps,
            "value": round(total * steps / t_batch, 1), "unit": "particle-updates/s",
            "ms_per_step": round(t_batch * 1e3 / steps, 4), "graph": True,
            "sequential_plain_value": round(total * steps / t_seq, 1),
            "speedup_vs_sequential": round(t_seq / t_batch, 3)}


def bench_adjoint(steps: int = 200, every: int = 20):
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
    sim = DifferentiableSPH(w.params, w.constants, steps=steps, checkpoint_every=every)
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
    run()  # warm-up
    torch.cuda.synchronize()
    a.record()
    loss, (gx0, gv0, gmu) = run()
    torch.cuda.synchronize()
    t_f, t_b = a.elapsed_time(b) / 1e3, c.elapsed_time(d) / 1e3
    log(f"adjoint: forward {t_f * 1e3:.1f} ms, backward {t_b * 1e3:.1f} ms for {steps} steps of {n} particles")
    return {"workload": "adjoint_16k", "particles": n, "steps": steps, "checkpoint_every": every,
            "value": round(n * steps / (t_f + t_b), 1), "unit": "particle-updates/s",
            "ms_forward": round(t_f * 1e3, 3), "ms_backward": round(t_b * 1e3, 3),
            "ms_per_adjoint_step": round(t_b * 1e3 / steps, 4),
            "loss": float(loss), "grad_mu": float(gmu.item()), "grad_v0_norm": float(gv0.double().norm()),
            "finite": bool(torch.isfinite(gv0).all() and torch.isfinite(gmu).all())}


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
    strat = NaiveSPHStrategy(w.params, constants=w.constants, mode=w.mode, device=dev)
    pos = torch.from_numpy(np.ascontiguousarray(w.state.position)).to(dev)
    vel = torch.from_numpy(np.ascontiguousarray(w.state.velocity)).to(dev)
    if args.block:
        strat.ctx.set_launch(*args.block)
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
    strat.ctx.step(args.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    diag = strat.diagnostics()
    log(f"timed: {gpu_ms:.2f} ms on the GPU, {wall * 1e3:.2f} ms wall; diagnostics {diag}")
    ms_per_step = gpu_ms / max(args.steps, 1)
    value = n * args.steps / (gpu_ms / 1e3)

    # per-kernel timing over the same trajectory: the step is deterministic, so replaying
    # the upload, the warm-up and K staged steps revisits exactly the timed states
    log(f"stage timing over the same {args.steps} steps")
    strat.ctx.upload(pos, vel)
    strat.ctx.step(args.warmup)
    st = stage_times(strat, args.steps)
    cand, pairs = neighbour_stats(w)
    force_s = st["force"] / 1e3
    force_bytes = FORCE_BYTES_PER_PARTICLE * n
    achieved = force_bytes / force_s / 1e9
    flops = n * (cand * FLOPS_PER_CANDIDATE + pairs * FLOPS_PER_PAIR)
    roofline = {"bound": "hbm", "kernel": "naive_force_kernel", "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None, "bytes_per_particle": FORCE_BYTES_PER_PARTICLE,
                "kernel_ms": round(st["force"], 4),
                "valu": {"achieved_tflops": round(flops / force_s / 1e12, 2), "peak_tflops": FP32_VECTOR_PEAK_TFLOPS,
                         "candidates_per_particle": round(cand, 1), "pairs_per_particle": round(pairs, 1)}}
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        tr = json.load(open(traffic_file)).get(args.workload, {}).get("naive_force_kernel")
        if tr:
            roofline["traffic"] = tr.get("bytes_per_launch")

    if args.calibrate_copy:
        src = torch.empty(256 * 2 ** 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 5 copies of 256 MiB")

    batched = bench_batched(args.batch_steps) if args.batched == "on" else None
    adjoint = bench_adjoint() if args.adjoint == "on" else None

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
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages_ms": {k: round(v, 4) for k, v in st.items()},
        "batched": batched,
        "adjoint": adjoint,
        "wall_ms_per_step": round(wall * 1e3 / max(args.steps, 1), 5),
        "final_state": {"non_finite": diag["non_finite"], "max_speed": float(diag["max_speed"]),
                        "mean_density": float(diag["mean_density"])},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
