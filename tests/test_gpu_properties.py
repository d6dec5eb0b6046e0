"""Size-independent properties of the GPU step, with no oracle in the loop.

- Renumbering the particles changes only the order in which a particle's sums are added (the
  cell order breaks ties by particle id): the step's cell table is bit-identical, densities and
  positions agree to FP32 rounding.  Run on the full 1M dam break.
- Shifting the whole state by whole cells shifts every key by one constant, so the cell order,
  the neighbour lists and their order are unchanged: with positions on a lattice of 2^-20 (so
  that the shifted coordinates and every coordinate difference are exact in FP32) the cell
  table moves by that constant and the density and the new velocity of every particle are
  bit-identical; the new positions differ by the shift up to the rounding of x + v dt.
"""
import numpy as np
import pytest
import torch

from sph_amd import _native as N
from sph_amd import workloads as W
from sph_amd.strategy import NaiveSPHStrategy

pytestmark = pytest.mark.gpu


def f32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def run(w, pos, vel, steps, device, grid=None):
    """A fresh context on (pos, vel) after `steps` steps in one call: positions, velocities,
    densities (particle order) and the cell table of the last step's sort."""
    strat = NaiveSPHStrategy(w.params, constants=w.constants, device=device, grid_voxel=grid)
    strat.ctx.upload(torch.from_numpy(np.ascontiguousarray(pos)).to(device),
                     torch.from_numpy(np.ascontiguousarray(vel)).to(device))
    strat.ctx.step(steps)
    p, v, rho = [t.cpu().numpy() for t in strat.ctx.download()][:3]
    begin = strat.ctx.read_buffer(N.BUF_CELL_BEGIN).cpu().numpy()
    strat.ctx.close()
    return p, v, rho, begin


def test_renumbered_particles_dam_break_1m(gpu):
    w = W.dam_break_1m()
    pos, vel = f32(w.state.position), f32(w.state.velocity)
    n = len(pos)
    perm = np.random.default_rng(71).permutation(n)
    p1, v1, rho1, begin1 = run(w, pos, vel, 1, gpu)
    q1, u1, sig1, begin2 = run(w, pos[perm], vel[perm], 1, gpu)
    # the sort of the step sees the same positions: the same cell table, bit for bit
    assert np.array_equal(begin1, begin2)
    # particle j of the renumbered run is particle perm[j] of the first
    assert np.max(np.abs(sig1 - rho1[perm]) / rho1[perm]) <= 1e-6
    assert np.max(np.abs(q1 - p1[perm]) / np.maximum(np.abs(p1[perm]), 1e-3)) <= 1e-6
    # and after three steps in one call (the force passes count the next sorts' keys)
    p3, _, rho3, _ = run(w, pos, vel, 3, gpu)
    q3, _, sig3, _ = run(w, pos[perm], vel[perm], 3, gpu)
    drho = np.max(np.abs(sig3 - rho3[perm]) / rho3[perm])
    dpos = np.max(np.abs(q3 - p3[perm]) / np.maximum(np.abs(p3[perm]), 1e-3))
    print(f"renumbered, 3 steps: rho {drho:.2e} pos {dpos:.2e}")
    assert drho <= 1e-5 and dpos <= 1e-5


@pytest.mark.parametrize("cells", [(8, 4, 16), (1, 0, 0), (0, 0, 5)])
def test_shift_by_whole_cells(gpu, cells):
    w = W.box_uniform(200_000, seed=72, h=0.012)
    g = 2.0 ** -7  # voxel_for(0.012): 128 cells per axis of the unit box
    lattice = 2.0 ** -20
    pos = np.round(w.state.position / lattice) * lattice
    vel = f32(np.random.default_rng(73).normal(0.0, 0.3, size=pos.shape))
    shift = np.asarray(cells, dtype=np.float64) * g
    assert pos.max() + shift.max() < 1.0 - 0.05  # away from the far walls during the step
    p, v, rho, begin = run(w, pos, vel, 1, gpu, grid=g)
    q, u, sig, begin_s = run(w, pos + shift, vel, 1, gpu, grid=g)
    dims = np.round(np.asarray(w.params.space_size) / g).astype(np.int64)
    off = int(cells[0] + dims[0] * (cells[1] + dims[1] * cells[2]))  # compute_1d_idx of the shift
    assert np.array_equal(begin_s[off:], begin[: len(begin) - off])
    assert np.all(begin_s[:off] == -1)
    assert np.array_equal(sig, rho)
    assert np.array_equal(u, v)
    assert np.max(np.abs((q - shift) - p)) <= 2.0 ** -23
