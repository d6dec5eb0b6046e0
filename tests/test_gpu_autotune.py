"""Autotuner on the GPU: it times every candidate, caches the winner, and a tuned grid
leaves the NAIVE step's results within the parity bar of the FP64 oracle (the grid only
changes which cells are walked and the FP32 summation order)."""
import numpy as np
import pytest
import torch

import oracle
from sph_amd import autotune as T
from sph_amd.batched import BatchedSPH
from sph_amd.data_classes import SimulationState
from sph_amd.strategy import NaiveSPHStrategy
from sph_amd.workloads import box_uniform, voxel_for

pytestmark = pytest.mark.gpu


def test_tune_strategy_caches_and_keeps_parity(gpu, tmp_path):
    w = box_uniform(4096, seed=51)
    cache = T.TuneCache(str(tmp_path / "tune.json"))
    res = T.tune_strategy(w.params, w.constants, w.state, steps=4, warmup=1, scales=(1.0, 0.5, 0.25),
                          blocks=(64, 256), lanes=(1, 4), cache=cache, device=gpu)
    assert len(res.trials) == 3 * 2 + 3 and not res.cached
    assert res.ms_per_step == min(t[4] for t in res.trials)
    again = T.tune_strategy(w.params, w.constants, w.state, steps=4, warmup=1, scales=(1.0, 0.5, 0.25),
                            blocks=(64, 256), lanes=(1, 4), cache=cache, device=gpu)
    assert again.cached and again.launch == res.launch and again.cell_scale == res.cell_scale
    assert again.lanes == res.lanes
    # every candidate grid gives the oracle's step within the bar
    pos = w.state.position.astype(np.float32).astype(np.float64)
    vel = np.random.default_rng(52).normal(0, 0.2, size=pos.shape).astype(np.float32).astype(np.float64)
    ref = oracle.step(pos, vel, consts=w.constants, ext=w.params.external_force, dt=1.0 / w.params.fps,
                      space=w.params.space_size, voxel=w.params.voxel_size)
    for scale in (1.0, 0.5, 0.25):
        s = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu,
                             grid_voxel=voxel_for(w.constants.h) * scale, launch=res.launch, lanes=res.lanes)
        out = s.compute_next_state(SimulationState(pos, vel, np.zeros(len(pos))))
        assert np.max(np.abs(out.density - ref[2]) / ref[2]) < 1e-5
        assert np.max(np.abs(out.position - ref[0])) < 1e-5


def test_tune_batch_and_bitwise_plain_equivalence(gpu, tmp_path):
    rng = np.random.default_rng(53)
    ws = [box_uniform(int(rng.integers(800, 3000)), seed=60 + b, h=float(rng.uniform(0.03, 0.06)),
                      edge=float(rng.uniform(0.5, 2.0))) for b in range(4)]
    inst = [(w.params, w.constants) for w in ws]
    res = T.tune_batch(inst, [w.state for w in ws], steps=3, warmup=1, blocks=(64, 256), lanes=(1, 8),
                       cache=T.TuneCache(str(tmp_path / "t.json")), device=gpu)
    assert len(res.trials) == 3 * 2 + 3
    batch = BatchedSPH(inst, device=gpu, cell_scale=res.cell_scale, launch=res.launch, lanes=res.lanes)
    batch.upload([w.state for w in ws])
    batch.step(5)
    got = batch.download()
    for b, w in enumerate(ws):
        plain = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu,
                                 grid_voxel=voxel_for(w.constants.h) * res.cell_scale, launch=res.launch,
                                 lanes=res.lanes)
        ref = plain.advance(w.state, 5)
        assert np.array_equal(got[b].position, ref.position), b
        assert np.array_equal(got[b].velocity, ref.velocity), b
