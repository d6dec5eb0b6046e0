"""Cells cut into x slabs (the fused step of the NAIVE formulation) and stage timing.

The fused step orders the particles of each cell by (x slab, id), with four slabs per cell
edge, and keeps the slab offsets in the cell table, so that the row walk searches slabs
rather than whole cells; the staged sort keeps the reference's (cell, id) order
(voxel_sph_strategy.py:84-88).  Both must search every pair within h: the staged path
and the fused one sum the same pairs in other orders, so they agree to FP32 rounding, and
both stay within the oracle's bars (the other GPU parity tests run the fused step)."""
import numpy as np
import pytest
import torch

import oracle
from sph_amd import _native as N
from sph_amd.strategy import NaiveSPHStrategy
from sph_amd.workloads import box_uniform

pytestmark = pytest.mark.gpu


def f32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def upload(strat, pos, vel):
    dev = strat.device
    strat.ctx.upload(torch.from_numpy(np.ascontiguousarray(pos, dtype=np.float64)).to(dev),
                     torch.from_numpy(np.ascontiguousarray(vel, dtype=np.float64)).to(dev))


def read(strat, which):
    return strat.ctx.read_buffer(which).cpu().numpy()


def host_order(pos, edge, dims, slabs=4):
    """(cell, x slab, id) order of positions inside the grid: the slab is
    floor(slabs * x / edge) - slabs * cell_x, exact for a power-of-two edge."""
    p = np.asarray(pos, dtype=np.float32)
    c = np.clip(np.floor(p / np.float32(edge)), 0, np.asarray(dims) - 1).astype(np.int64)
    key = c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2])
    q = np.floor(p[:, 0] * np.float32(1.0 / edge) * np.float32(slabs)) - slabs * c[:, 0]
    slab = np.clip(q, 0, slabs - 1).astype(np.int64)
    return np.lexsort((np.arange(len(p)), slab, key)), key


@pytest.mark.parametrize("lanes", [1, 4])
def test_fused_order_is_cell_slab_id(gpu, lanes):
    w = box_uniform(20000, seed=91)
    pos = f32(w.state.position)
    vel = f32(np.random.default_rng(91).normal(0.0, 0.5, size=pos.shape))
    edge = 2.0 ** -5
    strat = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, grid_voxel=edge, lanes=lanes)
    upload(strat, pos, vel)
    strat.ctx.step(1)
    dims = np.asarray(strat.ctx.grid_info()["dims"])
    order, key = host_order(pos, edge, dims)
    assert np.array_equal(read(strat, N.BUF_PERM), order)
    # the table read back is the plain cell table (the slab offsets are masked off)
    begin = oracle.voxel_begin(key[order], int(np.prod(dims)))
    assert np.array_equal(read(strat, N.BUF_CELL_BEGIN), begin)


@pytest.mark.parametrize("n,edge", [(20000, 2.0 ** -5), (3000, 2.0 ** -3)])
def test_staged_and_fused_steps_agree(gpu, n, edge):
    """Same pairs, other summation order: the staged step (reference order, no slab
    offsets for occupied cells: every cell searched whole) and the fused step agree to
    FP32 rounding.  3,000 particles on 2^-3 cells hold up to 20 particles in a cell's first
    three slabs, more than the 3-bit offset fields keep: those cells are searched whole."""
    w = box_uniform(n, seed=92, h=0.05 if n > 10000 else 0.2)
    pos = f32(w.state.position)
    vel = f32(np.random.default_rng(92).normal(0.0, 0.5, size=pos.shape))

    def make():
        s = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, grid_voxel=edge)
        upload(s, pos, vel)
        return s

    a, b = make(), make()
    a.ctx.step(1)
    for st in (N.STAGE_SORT, N.STAGE_TABLE, N.STAGE_REORDER, N.STAGE_DENSITY, N.STAGE_FORCE):
        b.ctx.run_stage(st)
    pa, va, ra = [t.cpu().numpy() for t in a.ctx.download()][:3]
    pb, vb, rb = [t.cpu().numpy() for t in b.ctx.download()][:3]
    assert np.max(np.abs(ra - rb) / rb) <= 1e-6
    assert np.max(np.abs(pa - pb) / np.maximum(np.abs(pb), 1e-3)) <= 1e-6
    sv = np.max(np.abs(vb))
    assert np.max(np.abs(va - vb)) <= 1e-5 * sv


def test_step_timed_is_step(gpu):
    """sph_step_timed runs sph_step's kernels: the same state bit for bit, and a positive
    time for each stage."""
    w = box_uniform(20000, seed=93)
    pos = f32(w.state.position)
    vel = f32(np.random.default_rng(93).normal(0.0, 0.5, size=pos.shape))

    def make():
        s = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, grid_voxel=2.0 ** -5)
        upload(s, pos, vel)
        return s

    a, b = make(), make()
    a.ctx.step(7)
    st = b.ctx.step_timed(7)
    print(f"stage ms: {st}")
    assert set(st) == {"sort", "density", "force"} and all(v > 0.0 for v in st.values())
    for ta, tb in zip(a.ctx.download(), b.ctx.download()):
        assert torch.equal(ta, tb)
    with pytest.raises(RuntimeError):
        b.ctx.step_timed(0)


def test_dense_cells_searched_whole(gpu):
    """300,000 particles at about 18 per cell: most cells hold more than the 6 particles below
    their last slab that the 3-bit offset fields keep, so the row walk searches them whole,
    taking the end of a row from the cell's count where the table entry keeps it (counts up to
    14 at this N) and from the next cell's start otherwise.  The fused step stays within the
    oracle's bars and agrees with the staged step (to 1e-5: sums of about 530 terms in two
    orders)."""
    w = box_uniform(300_000, seed=94, h=0.03)
    pos = f32(w.state.position)
    vel = f32(np.random.default_rng(94).normal(0.0, 0.5, size=pos.shape))
    edge = 2.0 ** -6

    def make():
        s = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, grid_voxel=edge)
        upload(s, pos, vel)
        return s

    a, b = make(), make()
    a.ctx.step(1)
    for st in (N.STAGE_SORT, N.STAGE_TABLE, N.STAGE_REORDER, N.STAGE_DENSITY, N.STAGE_FORCE):
        b.ctx.run_stage(st)
    pa, va, ra = [t.cpu().numpy() for t in a.ctx.download()][:3]
    pb, vb, rb = [t.cpu().numpy() for t in b.ctx.download()][:3]
    assert np.max(np.abs(ra - rb) / rb) <= 1e-5
    assert np.max(np.abs(va - vb)) <= 1e-5 * np.max(np.abs(vb))
    rp, rv, rrho, _, _ = oracle.step(pos, vel, consts=w.constants, ext=w.params.external_force, dt=1.0 / w.params.fps,
                                space=w.params.space_size, voxel=np.full(3, edge), use_grid=True)
    assert np.max(np.abs(ra - rrho) / rrho) <= 1e-5
    assert np.max(np.abs(pa - rp) / np.maximum(np.abs(rp), 1e-3)) <= 1e-5


def test_whole_cells_option(gpu):
    """sph_set_slabs(ctx, 0): the fused step searches whole cells and orders each cell by id,
    the reference's order (voxel_sph_strategy.py:84-88); the same pairs as with slabs, so the
    state agrees to FP32 rounding.  Switching with a state resident rewrites its keys: the
    next steps are those of a context made with that setting, bit for bit."""
    w = box_uniform(20000, seed=95)
    pos = f32(w.state.position)
    vel = f32(np.random.default_rng(95).normal(0.0, 0.5, size=pos.shape))
    edge = 2.0 ** -5

    def make(slabs):
        s = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, grid_voxel=edge, slabs=slabs)
        upload(s, pos, vel)
        return s

    whole, cut = make(False), make(True)
    whole.ctx.step(1)
    cut.ctx.step(1)
    dims = np.asarray(whole.ctx.grid_info()["dims"])
    c = np.clip(np.floor(pos / edge), 0, dims - 1).astype(np.int64)
    key = c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2])
    assert np.array_equal(read(whole, N.BUF_PERM), oracle.sort_map(key))
    pw, vw, rw = [t.cpu().numpy() for t in whole.ctx.download()][:3]
    pc, vc, rc = [t.cpu().numpy() for t in cut.ctx.download()][:3]
    assert np.max(np.abs(rw - rc) / rc) <= 1e-6
    assert np.max(np.abs(pw - pc) / np.maximum(np.abs(pc), 1e-3)) <= 1e-6
    # switch each context to the other setting and step on: bitwise the steps of a context
    # made with that setting from the same state
    refs = []
    for a, slabs in ((cut, False), (whole, True)):
        p, v = [t.cpu().numpy() for t in a.ctx.download()][:2]
        ref = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, grid_voxel=edge, slabs=slabs)
        upload(ref, p, v)
        refs.append(ref)
    cut.ctx.set_slabs(False)
    whole.ctx.set_slabs(True)
    with pytest.raises(RuntimeError):  # the last step's table is not readable after a switch
        read(cut, N.BUF_CELL_BEGIN)
    for a, b in zip((cut, whole), refs):
        a.ctx.step(3)
        b.ctx.step(3)
        for ta, tb in zip(a.ctx.download(), b.ctx.download()):
            assert torch.equal(ta, tb)


def test_set_slabs_refused_where_layout_is_fixed(gpu):
    from sph_amd.batched import BatchedSPH
    from sph_amd.strategy import VoxelSPHStrategy

    w = box_uniform(3000, seed=96)
    b = BatchedSPH([(w.params, w.constants)] * 2, device=gpu)
    with pytest.raises(RuntimeError):
        b.ctx.set_slabs(False)
    v = VoxelSPHStrategy(w.params, constants=w.constants, device=gpu)
    with pytest.raises(RuntimeError):
        v.ctx.set_slabs(False)
