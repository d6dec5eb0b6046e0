"""Batched form against the plain step, through the C ABI.

Each instance of a batch must compute exactly what the plain step computes on that
instance alone: the instance keeps its own cell order, neighbour lists and arithmetic, so
the comparison is bitwise.  The plain step is itself checked against the FP64 oracle in
test_gpu_step.py.
"""
import numpy as np
import pytest
import torch

from sph_amd.batched import BatchedSPH
from sph_amd.data_classes import SimulationState
from sph_amd.strategy import NaiveSPHStrategy
from sph_amd.workloads import box_uniform, voxel_for

pytestmark = pytest.mark.gpu


def _instances(seed=0, count=6):
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        n = int(rng.integers(500, 3000)) if b != 2 else 1  # one single-particle instance
        w = box_uniform(n, seed=100 + b, h=float(rng.uniform(0.03, 0.06)), edge=float(rng.uniform(0.5, 2.0)),
                        visc=float(rng.uniform(1, 8)))
        vel = rng.normal(0, 0.3, size=(n, 3))
        w.state = SimulationState(w.state.position.astype(np.float32).astype(np.float64),
                                  vel.astype(np.float32).astype(np.float64), np.zeros(n))
        out.append(w)
    return out


@pytest.mark.parametrize("lanes", [1, 8])
def test_batch_equals_plain_per_instance(gpu, lanes):
    ws = _instances()
    batch = BatchedSPH([(w.params, w.constants) for w in ws], device=gpu)
    batch.ctx.set_lanes(lanes)
    batch.upload([w.state for w in ws])
    batch.step(10)
    got = batch.download()
    for b, w in enumerate(ws):
        plain = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu)
        plain.ctx.set_lanes(lanes)
        ref = plain.advance(w.state, 10)
        assert np.array_equal(got[b].position, ref.position), b
        assert np.array_equal(got[b].velocity, ref.velocity), b
        assert np.array_equal(got[b].density, ref.density), b


def test_batch_mixed_lanes_and_grids_equal_plain(gpu):
    """Each instance with its own lane count and grid: the batch's thread ranges start on
    wavefront boundaries, so an instance computes what the plain step with its lanes and grid
    computes.  One dense instance (hundreds of neighbours, past the 96-slot list) takes the
    rescan path with 16 lanes."""
    ws = _instances(seed=4, count=7)
    dense = box_uniform(6000, seed=77, h=0.058, edge=0.6, visc=3.0)
    ws.append(dense)
    lanes = [1, 2, 4, 8, 16, 32, 2, 64]
    scales = [1.0, 0.5, 1.0, 0.5, 1.0, 0.5, 1.0, 0.5]
    batch = BatchedSPH([(w.params, w.constants) for w in ws], device=gpu, cell_scale=scales, lanes=lanes)
    batch.upload([w.state for w in ws])
    batch.step(6)
    got = batch.download()
    for b, w in enumerate(ws):
        plain = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu,
                                 grid_voxel=voxel_for(w.constants.h) * scales[b], lanes=lanes[b])
        ref = plain.advance(w.state, 6)
        assert np.array_equal(got[b].position, ref.position), b
        assert np.array_equal(got[b].velocity, ref.velocity), b
        assert np.array_equal(got[b].density, ref.density), b
    # lanes changed on a bound batch take effect from the next step, and the step is then the
    # plain step with the new lanes
    new = [16, 1, 8, 1, 2, 4, 1, 4]
    batch.ctx.set_batch_lanes(new)
    batch.upload([w.state for w in ws])
    batch.step(3)
    got = batch.download()
    for b, w in enumerate(ws):
        plain = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu,
                                 grid_voxel=voxel_for(w.constants.h) * scales[b], lanes=new[b])
        ref = plain.advance(w.state, 3)
        assert np.array_equal(got[b].position, ref.position), b


def test_set_batch_lanes_rejects(gpu):
    ws = _instances(seed=5, count=3)
    batch = BatchedSPH([(w.params, w.constants) for w in ws], device=gpu)
    with pytest.raises(RuntimeError, match="2 lane counts for 3"):
        batch.ctx.set_batch_lanes([1, 2])
    with pytest.raises(RuntimeError, match="1, 2, 4, 8, 16, 32 or 64"):
        batch.ctx.set_batch_lanes([1, 3, 1])
    plain = NaiveSPHStrategy(ws[0].params, constants=ws[0].constants, device=gpu)
    with pytest.raises(RuntimeError, match="not a batched context"):
        plain.ctx.set_batch_lanes([1])
    # a rejected call leaves the previous lanes in place
    batch.upload([w.state for w in ws])
    batch.step(2)
    got = batch.download()
    for b, w in enumerate(ws):
        ref = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu).advance(w.state, 2)
        assert np.array_equal(got[b].position, ref.position), b


def test_batch_graph_replays_match_eager(gpu):
    ws = _instances(seed=1, count=5)
    inst = [(w.params, w.constants) for w in ws]
    eager = BatchedSPH(inst, device=gpu)
    graph = BatchedSPH(inst, device=gpu)
    eager.upload([w.state for w in ws])
    graph.upload([w.state for w in ws])
    g = graph.capture(4)
    for r in range(3):
        eager.step(4)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        a, b = eager.download(), graph.download()
        for x, y in zip(a, b):
            assert np.array_equal(x.position, y.position), r
            assert np.array_equal(x.velocity, y.velocity), r


def test_batch_rejects_pipe_and_voxel(gpu):
    from sph_amd import _native as N

    w = box_uniform(100)
    p = N.make_params(100, w.params.external_force, 1e-3, w.params.space_size, w.params.voxel_size,
                      N.SPH_MODE_BOX, N.SPH_FORM_VOXEL)
    with pytest.raises(RuntimeError, match="NAIVE"):
        N.NativeContext([p], [N.make_consts(w.constants)], gpu)
