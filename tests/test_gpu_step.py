"""GPU parity of the plain step against the CPU FP64 oracle, through the C ABI.

Bars: voxel keys, sorted permutation and cell table bit-exact; density within 1e-5
relative; pressure / viscosity terms within 1e-5 of sum_j |term_ij| (the per-particle
magnitude the sum cancels from); positions within 1e-5 relative.  Inputs are rounded to
FP32 before both sides see them, so the comparison measures the kernels' arithmetic.
"""
import math

import numpy as np
import pytest
import torch

import oracle
from _brute import naive_terms, near_boundary
from sph_amd import _native as N
from sph_amd.constants import SPHConstants
from sph_amd.data_classes import SimulationParameters, SimulationState
from sph_amd.strategy import NaiveSPHStrategy, VoxelSPHStrategy
from sph_amd.workloads import box_uniform, pipe_flow

pytestmark = pytest.mark.gpu

TOL = 1e-5


def f32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def upload(strat, pos, vel):
    dev = strat.device
    strat.ctx.upload(torch.from_numpy(np.ascontiguousarray(pos, dtype=np.float64)).to(dev),
                     torch.from_numpy(np.ascontiguousarray(vel, dtype=np.float64)).to(dev))


def read(strat, which):
    return strat.ctx.read_buffer(which).cpu().numpy()


def oracle_step(w, pos, vel, formulation="NAIVE", rng=None, mode="BOX"):
    return oracle.step(pos, vel, consts=w.constants, ext=w.params.external_force, dt=1.0 / w.params.fps,
                       space=w.params.space_size, voxel=w.params.voxel_size, mode=mode, formulation=formulation,
                       pipe=w.params.pipe.to_numpy() if mode == "PIPE" else None, rng=rng)


def outside_points(pos, space):
    pos = pos.copy()
    L = np.asarray(space, dtype=np.float64)
    pts = np.array([[-0.3, 0.2, 0.2], [1.2, 0.5, 0.5], [0.5, -0.01, 0.9], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0],
                    [0.99999, 0.5, 0.5], [-2.0, -3.0, 5.0], [0.3, 1.7, -0.2]]) * L
    pos[: len(pts)] = f32(pts)
    return pos


@pytest.mark.parametrize("voxel", [0.0625, 0.0457])
def test_keys_sort_table_bit_exact(gpu, voxel):
    """voxel_kernels.py:88-105 keys, voxel_sph_strategy.py:84-88 order, :92-107 voxel_begin."""
    w = box_uniform(4096, seed=11)
    pos = outside_points(f32(w.state.position), w.params.space_size)
    params = SimulationParameters(4096, w.params.external_force, 1, 1000, w.params.pipe, w.params.space_size,
                                  np.array([voxel] * 3))
    strat = VoxelSPHStrategy(params, constants=w.constants, device=gpu)
    upload(strat, pos, np.zeros_like(pos))
    dims = oracle.space_dims(params.space_size, params.voxel_size, True)
    ref_keys = oracle.voxel_keys(pos, params.voxel_size, dims)
    assert np.array_equal(read(strat, N.BUF_KEYS), ref_keys)
    strat.ctx.run_stage(N.STAGE_SORT)
    ref_perm = oracle.sort_map(ref_keys)
    assert np.array_equal(ref_perm, np.argsort(ref_keys, kind="stable"))
    assert np.array_equal(read(strat, N.BUF_PERM), ref_perm)
    assert np.array_equal(read(strat, N.BUF_SORTED_KEYS), ref_keys[ref_perm])
    strat.ctx.run_stage(N.STAGE_TABLE)
    ref_begin = oracle.voxel_begin(ref_keys[ref_perm], int(np.prod(dims)))
    assert np.array_equal(read(strat, N.BUF_CELL_BEGIN), ref_begin)


def test_grid_keys_match_reference_inside_box(gpu):
    """The NAIVE path sorts on clamped keys; inside the box they equal the reference keys."""
    w = box_uniform(4096, seed=12)
    pos = f32(w.state.position)
    strat = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu)
    upload(strat, pos, np.zeros_like(pos))
    dims = oracle.space_dims(w.params.space_size, w.params.voxel_size, True)
    ref_keys = oracle.voxel_keys(pos, w.params.voxel_size, dims)
    assert np.array_equal(read(strat, N.BUF_KEYS), ref_keys)
    strat.ctx.run_stage(N.STAGE_SORT)
    assert np.array_equal(read(strat, N.BUF_PERM), np.argsort(ref_keys, kind="stable"))


def test_small_grid_sort_in_one_workgroup(gpu):
    """A problem of at most 8,192 cells and particles is counted, scanned and scattered by one
    workgroup with the counts in LDS: keys, order and cell table must still be the
    reference's (voxel_kernels.py:88-105, voxel_sph_strategy.py:84-107) bit for bit, on the
    grid the step searches (2^-4 here, 4,096 cells), and the step must match the oracle."""
    w = box_uniform(4096, seed=16)
    pos = f32(w.state.position)
    vel = f32(np.random.default_rng(16).normal(0.0, 0.5, size=pos.shape))
    grid = 2.0 ** -4
    strat = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, grid_voxel=grid)
    upload(strat, pos, vel)
    voxel = np.array([grid] * 3)
    dims = oracle.space_dims(w.params.space_size, voxel, True)
    assert int(np.prod(dims)) == 4096
    keys = oracle.voxel_keys(pos, voxel, dims)
    assert np.array_equal(read(strat, N.BUF_KEYS), keys)
    strat.ctx.run_stage(N.STAGE_SORT)
    perm = oracle.sort_map(keys)
    assert np.array_equal(read(strat, N.BUF_PERM), perm)
    assert np.array_equal(read(strat, N.BUF_SORTED_KEYS), keys[perm])
    strat.ctx.run_stage(N.STAGE_TABLE)
    assert np.array_equal(read(strat, N.BUF_CELL_BEGIN), oracle.voxel_begin(keys[perm], int(np.prod(dims))))
    out = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, grid_voxel=grid).compute_next_state(
        SimulationState(pos, vel, np.zeros(len(pos))))
    rp, rv, rrho, _, _ = oracle_step(w, pos, vel)
    assert np.max(np.abs(out.density - rrho) / rrho) <= TOL
    assert np.max(np.abs(out.position - rp) / np.maximum(np.abs(rp), 1e-3)) <= TOL


def _force_case(seed=13, n=4096):
    w = box_uniform(n, seed=seed)
    rng = np.random.default_rng(seed)
    pos = f32(w.state.position)
    vel = f32(rng.normal(0.0, 0.5, size=pos.shape))
    return w, pos, vel


def test_density_and_force_stages(gpu):
    """naive_kernels.py:7-80 stage by stage."""
    w, pos, vel = _force_case()
    c = w.constants
    strat = NaiveSPHStrategy(w.params, constants=c, device=gpu, keep_terms=True)
    upload(strat, pos, vel)
    for st in (N.STAGE_SORT, N.STAGE_TABLE, N.STAGE_REORDER, N.STAGE_DENSITY):
        strat.ctx.run_stage(st)
    rho = read(strat, N.BUF_DENSITY).astype(np.float64)
    rho_ref = oracle.naive_density(pos, c)
    rel = np.abs(rho - rho_ref) / rho_ref
    print(f"density max rel err {rel.max():.3e}")
    assert rel.max() <= TOL
    strat.ctx.run_stage(N.STAGE_FORCE)
    P = read(strat, N.BUF_PRESSURE).reshape(-1, 3).astype(np.float64)
    V = read(strat, N.BUF_VISCOSITY).reshape(-1, 3).astype(np.float64)
    # kernel error alone: the oracle's sums over the GPU's densities
    P_ref = oracle.naive_pressure(pos, rho, c)
    V_ref = oracle.naive_viscosity(pos, vel, rho, c)
    _, _, _, sP, sV = naive_terms(pos, vel, c, rho=rho)
    eP = np.linalg.norm(P - P_ref, axis=1) / np.maximum(sP, 1e-30)
    eV = np.linalg.norm(V - V_ref, axis=1) / np.maximum(sV, 1e-30)
    print(f"pressure err {eP.max():.3e} viscosity err {eV.max():.3e} (normalised)")
    assert eP.max() <= TOL and eV.max() <= TOL
    # end to end against the oracle's own densities (includes density rounding)
    P_ref2 = oracle.naive_pressure(pos, rho_ref, c)
    _, _, _, sP2, _ = naive_terms(pos, vel, c, rho=rho_ref)
    eP2 = np.linalg.norm(P - P_ref2, axis=1) / np.maximum(sP2, 1e-30)
    print(f"pressure err vs FP64 densities {eP2.max():.3e}")
    assert eP2.max() <= TOL


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("lanes", [1, 4])
def test_density_of_particles_outside_the_grid(gpu, lanes, fused):
    """naive_kernels.py:7-21 sums over all pairs, wherever the particles are.  The grid keeps
    particles beyond its faces in its edge cells and searches from their clamped positions;
    clusters of mutual neighbours beyond a face, straddling the far face and beyond a corner
    must get the all-pairs density, and so must the particles inside.  Through the staged
    stages (whole cells) and the fused step (x slabs, the edge cells' outer slabs extending
    beyond the grid)."""
    w = box_uniform(4096, seed=15)
    c = w.constants
    L = np.asarray(w.params.space_size, dtype=np.float64)
    pos = f32(w.state.position)
    rng = np.random.default_rng(15)
    centres = [np.array([-0.2, 0.3, 0.3]) * L, np.array([L[0], 0.45 * L[1], 0.4 * L[2]]),
               np.array([-0.1, -0.1, 1.1]) * L, np.array([0.5, 1.0 + c.h / L[1], 0.5]) * L]
    k = 0
    for ctr in centres:
        for _ in range(6):
            pos[k] = ctr + rng.uniform(-0.25, 0.25, size=3) * c.h  # mates within 0.87 h
            k += 1
    pos = f32(pos)
    strat = NaiveSPHStrategy(w.params, constants=c, device=gpu, lanes=lanes)
    upload(strat, pos, np.zeros_like(pos))
    if fused:
        strat.ctx.step(1)
    else:
        for st in (N.STAGE_SORT, N.STAGE_TABLE, N.STAGE_REORDER, N.STAGE_DENSITY):
            strat.ctx.run_stage(st)
    rho = read(strat, N.BUF_DENSITY).astype(np.float64)
    rho_ref = oracle.naive_density(pos, c)
    rel = np.abs(rho - rho_ref) / rho_ref
    print(f"lanes {lanes}: outside clusters rel err {rel[:k].max():.3e}, all {rel.max():.3e}")
    # every cluster particle has its five mates within h: (1 - r^2/h^2)^3 >= 0.25^3 each
    w_norm = c.mass * 315.0 / (64.0 * np.pi * c.h ** 3)
    assert np.all(rho_ref[:k] >= w_norm * (1.0 + 5 * 0.25 ** 3) * (1 - 1e-6))
    assert rel.max() <= TOL


def test_non_finite_particles_are_counted_not_summed(gpu):
    """naive_kernels.py:18 skips a pair only when norm_squared > h^2, so in the reference a
    NaN position makes every density NaN.  This build differs on purpose: a non-finite pair
    adds nothing (the pair test r^2 <= h^2 fails and (1 - r^2/h^2)^3 is clamped at zero), so
    a NaN or infinite particle has density 0, the finite particles get the all-pairs density
    of the finite particles alone, and the resident state counts the non-finite ones."""
    w, pos, vel = _force_case(seed=21)
    c = w.constants
    pos[5] = np.nan
    pos[6, 1] = np.inf
    pos[7, 2] = -np.inf
    pos[8, 0] = np.nan
    bad = ~np.isfinite(pos).all(axis=1)
    strat = NaiveSPHStrategy(w.params, constants=c, device=gpu)
    upload(strat, pos, vel)
    assert strat.diagnostics()["non_finite"] == int(bad.sum()) == 4
    for st in (N.STAGE_SORT, N.STAGE_TABLE, N.STAGE_REORDER, N.STAGE_DENSITY):
        strat.ctx.run_stage(st)
    rho = read(strat, N.BUF_DENSITY).astype(np.float64)
    assert np.all(rho[bad] == 0.0)
    rho_ref = oracle.naive_density(pos[~bad], c)
    rel = np.abs(rho[~bad] - rho_ref) / rho_ref
    print(f"finite particles' density rel err {rel.max():.3e}")
    assert rel.max() <= TOL


def test_full_step_matches_oracle(gpu):
    """compute_next_state (abstract_sph_strategy.py:31-46), NAIVE, BOX."""
    w, pos, vel = _force_case(seed=14)
    strat = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, keep_terms=True)
    out = strat.compute_next_state(SimulationState(pos, vel, np.zeros(len(pos))))
    rp, rv, rrho, rf, _ = oracle_step(w, pos, vel)
    drho = np.abs(out.density - rrho) / rrho
    dpos = np.abs(out.position - rp) / np.maximum(np.abs(rp), 1e-3)
    rho_, P_, V_, sP, sV = naive_terms(pos, vel, w.constants)
    scale_v = (sP + sV + np.linalg.norm(w.params.external_force)) / rrho / w.params.fps
    dvel = np.linalg.norm(out.velocity - rv, axis=1) / np.maximum(scale_v, 1e-30)
    dfor = np.linalg.norm(strat.result_force - rf, axis=1) / np.maximum(sP + sV, 1e-30)
    print(f"step: rho {drho.max():.2e} pos {dpos.max():.2e} vel(norm) {dvel.max():.2e} force(norm) {dfor.max():.2e}")
    assert drho.max() <= TOL and dpos.max() <= TOL
    assert dvel.max() <= TOL and dfor.max() <= TOL


def test_multi_step_trajectory(gpu):
    """20 steps of the box workload: positions stay within 1e-5 relative of the FP64 oracle."""
    w = box_uniform(4096, seed=15)
    pos, vel = f32(w.state.position), f32(w.state.velocity)
    strat = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu)
    out = strat.advance(SimulationState(pos, vel, np.zeros(len(pos))), 20)
    rp, rv = pos, vel
    for _ in range(20):
        rp, rv, rrho, _, _ = oracle_step(w, rp, rv)
    dpos = np.abs(out.position - rp) / np.maximum(np.abs(rp), 1e-3)
    drho = np.abs(out.density - rrho) / rrho
    print(f"20 steps: pos {dpos.max():.2e} rho {drho.max():.2e}")
    assert dpos.max() <= TOL
    assert drho.max() <= TOL


def test_box_collisions_all_walls(gpu):
    """base_kernels.py:75-98 on particles leaving through each face and a corner."""
    c = SPHConstants(mass=0.02, rho0=1000.0, h=0.0457, visc=3.5, k=3.0, damp=0.5)
    pos = f32([[0.0005, 0.5, 0.5], [0.9995, 0.5, 0.5], [0.5, 0.0005, 0.5], [0.5, 0.9995, 0.5],
               [0.5, 0.5, 0.0005], [0.5, 0.5, 0.9995], [0.0005, 0.0005, 0.9995], [0.3, 0.3, 0.3]])
    vel = f32([[-2, 0, 0], [2, 0, 0], [0, -2, 0], [0, 2, 0], [0, 0, -2], [0, 0, 2], [-3, -3, 3], [0.1, 0.2, 0.3]])
    params = SimulationParameters(len(pos), np.array([0.0, -9.81, 0.0]), 1, 1000, space_size=np.ones(3),
                                  voxel_size=np.full(3, 0.0625))
    strat = NaiveSPHStrategy(params, constants=c, device=gpu)
    out = strat.compute_next_state(SimulationState(pos, vel, np.zeros(len(pos))))
    rp, rv, _, _, _ = oracle.step(pos, vel, consts=c, ext=params.external_force, dt=1e-3, space=np.ones(3),
                                  voxel=params.voxel_size)
    assert np.allclose(out.position, rp, rtol=1e-6, atol=1e-7)
    assert np.allclose(out.velocity, rv, rtol=1e-5, atol=1e-6)
    # every bounced component flipped and damped
    assert out.velocity[0, 0] > 0 and out.velocity[1, 0] < 0 and abs(out.velocity[0, 0]) < 2


@pytest.mark.parametrize("grid", [None, 2.0 ** -4, 0.0391])
@pytest.mark.parametrize("n", [0, 1, 2, 63, 1000, 2049, 8192, 8193])
def test_small_and_ragged_sizes(gpu, n, grid):
    """Empty, single and ragged sizes; with the 2^-4 grid (4,096 cells) up to 8,192 particles
    go through the one-workgroup sort and 8,193 through the multi-kernel one.  A cell edge
    that is not a power of two (0.0391) forms the keys in FP64 and the x slabs of the cells
    from the FP32 quotient."""
    w = box_uniform(max(n, 1), seed=16)
    params = SimulationParameters(n, w.params.external_force, 1, 1000, w.params.pipe, w.params.space_size,
                                  w.params.voxel_size)
    pos = f32(w.state.position[:n])
    vel = np.zeros((n, 3))
    strat = NaiveSPHStrategy(params, constants=w.constants, device=gpu, grid_voxel=grid)
    out = strat.advance(SimulationState(pos, vel, np.zeros(n)), 3)
    rp, rv = pos, vel
    for _ in range(3):
        rp, rv, rrho, _, _ = oracle_step(w, rp, rv)
    assert out.position.shape == (n, 3)
    if n:
        assert np.allclose(out.position, rp, rtol=1e-6, atol=1e-7)
        assert np.allclose(out.density, rrho, rtol=TOL)
    if n == 1:
        # an isolated particle: rho = m W(0) = m 315 / (64 pi h^3)
        c = w.constants
        assert math.isclose(out.density[0], c.mass * 315 / (64 * math.pi * c.h ** 3), rel_tol=1e-6)


def test_all_particles_in_one_cell(gpu):
    c = SPHConstants(mass=0.02, rho0=1000.0, h=0.0457, visc=3.5, k=3.0, damp=0.5)
    rng = np.random.default_rng(17)
    pos = f32(0.5 + rng.uniform(0, 0.06, size=(700, 3)))
    vel = f32(rng.normal(0, 0.1, size=(700, 3)))
    params = SimulationParameters(700, np.array([0.0, -9.81, 0.0]), 1, 1000, space_size=np.ones(3),
                                  voxel_size=np.full(3, 0.0625))
    strat = NaiveSPHStrategy(params, constants=c, device=gpu)
    out = strat.compute_next_state(SimulationState(pos, vel, np.zeros(700)))
    rp, rv, rrho, _, _ = oracle.step(pos, vel, consts=c, ext=params.external_force, dt=1e-3, space=np.ones(3),
                                     voxel=params.voxel_size)
    assert np.allclose(out.density, rrho, rtol=TOL)
    assert np.allclose(out.position, rp, rtol=TOL)


def test_voxel_formulation(gpu):
    """voxel_kernels.py:29-211: capped neighbour lists, then rho / P / V of that formulation."""
    w = box_uniform(4096, seed=18)
    pos = f32(w.state.position)
    vel = f32(np.random.default_rng(18).normal(0, 0.3, size=pos.shape))
    c = w.constants
    strat = VoxelSPHStrategy(w.params, constants=c, device=gpu, keep_terms=True)
    upload(strat, pos, vel)
    for st in (N.STAGE_SORT, N.STAGE_TABLE, N.STAGE_REORDER, N.STAGE_DENSITY, N.STAGE_FORCE):
        strat.ctx.run_stage(st)
    lists = read(strat, N.BUF_NEIGHBOURS).reshape(-1, 33)
    dims = oracle.space_dims(w.params.space_size, w.params.voxel_size, True)
    tdims = oracle.space_dims(w.params.space_size, w.params.voxel_size, False)
    keys = oracle.voxel_keys(pos, w.params.voxel_size, dims)
    perm = oracle.sort_map(keys)
    begin = oracle.voxel_begin(keys[perm], int(np.prod(dims)))
    n = len(pos)
    nb = np.full((n, 32), -1, dtype=np.int32)
    nn = np.zeros(n, dtype=np.int32)
    for i in range(n):
        l = oracle.voxel_neighbours(pos, begin, perm, w.params.voxel_size, tdims, c.h, i)
        nn[i] = len(l)
        nb[i, : len(l)] = l
    # pairs within 1e-5 h of the boundary may fall either side in FP32; skip those particles
    close = near_boundary(pos, c.h, 1e-5)
    same = (lists[:, 0] == nn) & np.all((lists[:, 1:] == nb) | (nb == -1), axis=1)
    assert np.all(same | close), f"{np.sum(~(same | close))} neighbour lists differ"
    rho_ref, P_ref, V_ref = oracle.voxel_forces(pos, vel, nb, nn, c)
    ok = ~close
    rho = read(strat, N.BUF_DENSITY).astype(np.float64)
    assert np.max(np.abs(rho - rho_ref)[ok] / np.abs(rho_ref[ok])) <= TOL
    P = read(strat, N.BUF_PRESSURE).reshape(-1, 3)
    V = read(strat, N.BUF_VISCOSITY).reshape(-1, 3)
    # errors against the per-particle sum of the pair terms' magnitudes (the sums cancel)
    sP, sV = _voxel_term_scales(pos, vel, rho_ref, nb, nn, c)
    eP = np.linalg.norm(P - P_ref, axis=1) / np.maximum(sP, 1e-30)
    clipped = V_ref >= 1.0
    eV = np.linalg.norm(np.where(clipped, 0.0, V - V_ref), axis=1) / np.maximum(sV, 1e-30)
    print(f"voxel P err {eP[ok].max():.2e}, V err {eV[ok].max():.2e} (normalised)")
    assert eP[ok].max() <= TOL and eV[ok].max() <= TOL
    assert np.all(V[ok][clipped[ok]] == 1.0)


def _voxel_term_scales(pos, vel, rho, nb, nn, c):
    """Per particle, sum_j |(p_i/rho_i^2 + p_j/rho_j^2) grad W| and
    sum_j |v_j - v_i| / rho_j lap W m mu / rho_i over its VOXEL list (voxel_kernels.py:135-211)."""
    n = len(pos)
    p = c.k * (rho - c.rho0)
    sP, sV = np.zeros(n), np.zeros(n)
    gc, lc = 45.0 / (math.pi * c.h ** 6), 45.0 / (math.pi * c.h ** 6)
    for i in range(n):
        js = nb[i, :nn[i]]
        js = js[js != i]
        if len(js) == 0:
            continue
        r = np.linalg.norm(pos[i] - pos[js], axis=1)
        sP[i] = np.sum(np.abs(p[i] / rho[i] ** 2 + p[js] / rho[js] ** 2) * gc * (c.h - r) ** 2)
        sV[i] = np.sum(np.linalg.norm(vel[js] - vel[i], axis=1) / rho[js] * lc * (c.h - r)) * c.mass * c.visc / rho[i]
    return sP, sV


def test_pipe_mode_and_rng(gpu):
    """base_kernels.py:56-72, 101-127: wall reflection, mirror at x < 0, re-injection at the
    inlet with the particle's own xoroshiro128+ stream (seed 16435234, one stream per id)."""
    w = pipe_flow(n=4096, seed=19, steps=10)
    pos, vel = f32(w.state.position), f32(w.state.velocity)
    # particles about to leave through the outlet, the inlet and the wall
    pos[:40, 0] = 3.999
    vel[:40, 0] = 5.0
    pos[40:60, 0] = 0.0005
    vel[40:60, 0] = -3.0
    strat = NaiveSPHStrategy(w.params, constants=w.constants, mode="PIPE", device=gpu)
    rng0 = oracle.rng_init(16435234, len(pos))
    assert np.array_equal(read_rng_after_upload(strat, pos, vel), rng0)
    out = strat.advance(SimulationState(pos, vel, np.zeros(len(pos))), 3)
    rp, rv, rr = pos, vel, rng0
    for _ in range(3):
        rp, rv, _, _, rr = oracle_step(w, rp, rv, rng=rr, mode="PIPE")
    assert np.array_equal(read(strat, N.BUF_RNG).view(np.uint64), rr)
    moved = np.any(rr != rng0)
    assert moved
    err = np.abs(out.position - rp).max(axis=1)
    print("pipe max abs pos err", err.max(), "re-injected", int(np.sum(rr.reshape(-1, 2)[:, 0] != rng0.reshape(-1, 2)[:, 0])))
    assert err.max() <= TOL


def read_rng_after_upload(strat, pos, vel):
    upload(strat, pos, vel)
    return read(strat, N.BUF_RNG).view(np.uint64)


def test_diagnostics_fp64(gpu):
    w, pos, vel = _force_case(seed=20)
    strat = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu)
    out = strat.advance(SimulationState(pos, vel, np.zeros(len(pos))), 2)
    d = strat.diagnostics()
    m = w.constants.mass
    v = out.velocity
    assert math.isclose(d["mass"], m * len(pos))
    assert np.allclose(d["momentum"], m * v.sum(axis=0), rtol=1e-12, atol=1e-12)
    assert math.isclose(d["kinetic_energy"], 0.5 * m * np.sum(v * v), rel_tol=1e-12)
    assert d["non_finite"] == 0


@pytest.mark.parametrize("n", [3000, 20000])
@pytest.mark.parametrize("mode", ["BOX", "PIPE"])
def test_graph_replays_match_eager(gpu, mode, n):
    """A captured run of steps, replayed several times with eager work and synchronisation
    in between, advances the state exactly as eager steps do (PIPE: RNG streams included).
    3,000 particles sort in one workgroup; 20,000 take the multi-workgroup counting sort,
    whose counts and scan-tile sums the captured force passes make for the next step's sort."""
    w = box_uniform(n, seed=23) if mode == "BOX" else pipe_flow(n=n, seed=24)
    pos, vel = f32(w.state.position), f32(w.state.velocity)
    eager = NaiveSPHStrategy(w.params, constants=w.constants, mode=mode, device=gpu)
    graph = NaiveSPHStrategy(w.params, constants=w.constants, mode=mode, device=gpu)
    upload(eager, pos, vel)
    upload(graph, pos, vel)
    g = graph.capture(5)
    for r in range(3):
        eager.ctx.step(5)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        a = [t.cpu().numpy() for t in eager.ctx.download()]
        b = [t.cpu().numpy() for t in graph.ctx.download()]
        for x, y in zip(a, b):
            assert np.array_equal(x, y), r
        if mode == "PIPE":
            assert np.array_equal(read(eager, N.BUF_RNG), read(graph, N.BUF_RNG))


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("dense", [False, True])
def test_lanes_per_particle_match_oracle(gpu, lanes, dense):
    """Density and force with several lanes per particle: the step stays within the parity
    bars; dense = more neighbours than list slots (the rescan path of the force pass)."""
    if dense:
        w = box_uniform(3000, seed=25)
        pos = f32(np.random.default_rng(26).uniform([0.1, 0.1, 0.1], [0.22, 0.22, 0.22], size=(3000, 3)))
    else:
        w = box_uniform(4096, seed=27)
        pos = f32(w.state.position)
    vel = f32(np.random.default_rng(28).normal(0, 0.3, size=pos.shape))
    strat = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, keep_terms=True, launch=(64, 128))
    strat.ctx.set_lanes(lanes)
    out = strat.compute_next_state(SimulationState(pos, vel, np.zeros(len(pos))))
    rp, rv, rr, rf, _ = oracle_step(w, pos, vel)
    assert np.max(np.abs(out.density - rr) / rr) <= TOL
    # against the per-particle sum of the pair terms' magnitudes (the sums cancel)
    _, _, _, sP, sV = naive_terms(pos, vel, w.constants)
    ef = np.max(np.linalg.norm(strat.result_force - rf, axis=1) / np.maximum(sP + sV, 1e-30))
    print(f"lanes {lanes} dense {dense}: force err {ef:.2e}")
    assert ef <= TOL
    assert np.max(np.abs(out.position - rp) / np.maximum(np.abs(rp), 1.0)) <= TOL


def test_set_lanes_rejects_bad_values(gpu):
    w = box_uniform(100, seed=29)
    s = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu)
    for bad in (0, 3, 48, 128):
        with pytest.raises(RuntimeError, match="lanes per particle"):
            s.ctx.set_lanes(bad)
    v = VoxelSPHStrategy(w.params, constants=w.constants, device=gpu)
    with pytest.raises(RuntimeError, match="VOXEL"):
        v.ctx.set_lanes(4)


@pytest.mark.parametrize("lanes", [1, 4])
def test_far_neighbours_in_cell_order_are_rescanned(gpu, lanes):
    """List entries are int16 offsets in cell order.  In a thin slab (4 x 4 x 0.1 with four
    cell layers in z) a particle's neighbours in the next layer are ~75,000 slots away, so
    the density pass marks it for a rescan and the force pass walks its cells instead: the
    step still matches the FP64 oracle."""
    n = 300_000
    rng = np.random.default_rng(31)
    pos = f32(rng.uniform([0.0, 0.0, 0.0], [4.0, 4.0, 0.1], size=(n, 3)))
    vel = f32(rng.normal(0.0, 0.1, size=(n, 3)))
    c = SPHConstants(h=0.0457, mass=0.02, rho0=1000.0, k=3.0, visc=3.5, damp=0.5)
    vx = 0.03125
    params = SimulationParameters(n, np.array([0.0, -9.81, 0.0]), 1, 1000, box_uniform(16).params.pipe,
                                  np.array([4.0, 4.0, 0.1]), np.array([vx, vx, vx]))
    strat = NaiveSPHStrategy(params, constants=c, device=gpu, lanes=lanes)
    out = strat.compute_next_state(SimulationState(pos, vel, np.zeros(n)))
    rp, rv, rrho, _, _ = oracle.step(pos, vel, consts=c, ext=params.external_force, dt=1e-3, space=params.space_size,
                                     voxel=params.voxel_size, use_grid=True)
    drho = np.abs(out.density - rrho) / rrho
    dpos = np.abs(out.position - rp) / np.maximum(np.abs(rp), 1e-3)
    print(f"slab lanes {lanes}: rho {drho.max():.2e} pos {dpos.max():.2e}")
    assert drho.max() <= TOL and dpos.max() <= TOL


def _same_after(make, upload_state, steps):
    """Two contexts from make(): one runs `steps` steps in one sph_step call, the other in
    single-step calls; returns both contexts."""
    a, b = make(), make()
    upload_state(a)
    upload_state(b)
    a.ctx.step(steps)
    for _ in range(steps):
        b.ctx.step(1)
    return a, b


@pytest.mark.parametrize("mode,lanes,edge", [("BOX", 1, 2.0 ** -5), ("BOX", 4, 2.0 ** -5), ("PIPE", 1, None),
                                             ("BOX", 1, 2.0 ** -8)])
def test_steps_in_one_call_equal_single_calls(gpu, mode, lanes, edge):
    """One sph_step(k) call and k single-step calls give bitwise the same state, keys,
    permutation and cell table.  20,000 particles on 32,768 cells or more (the counting
    sort's multi-workgroup path, which keeps counts in device memory): with one lane per
    particle the k-step call's force passes count the keys and add up the scan's tile sums
    for the next sort, and the sorts after the first start at the scan; single calls count
    in every sort; with 4 lanes every sort counts.  Particles outside the grid (BOX) and
    re-injection (PIPE).  2^-8 cells: 256^3 cells in 4,096 scan tiles, more than a tile's
    workgroup sums itself, so the tile sums are scanned by a workgroup of their own first."""
    if mode == "BOX":
        w = box_uniform(20000, seed=31)
        pos = outside_points(f32(w.state.position), w.params.space_size)
        grid = edge
    else:
        w = pipe_flow(n=20000, seed=32)
        pos = f32(w.state.position)
        pos[:300, 0] = 3.9995
        grid = None
    vel = f32(np.random.default_rng(33).normal(0.0, 0.5, size=pos.shape))
    if mode == "PIPE":
        vel[:300, 0] = 5.0

    def make():
        return NaiveSPHStrategy(w.params, constants=w.constants, mode=mode, device=gpu, grid_voxel=grid, lanes=lanes)

    a, b = _same_after(make, lambda s: upload(s, pos, vel), 7)
    info = a.ctx.grid_info()
    print(f"{mode} lanes {lanes}: grid {info}")
    for ta, tb in zip(a.ctx.download(), b.ctx.download()):
        assert torch.equal(ta, tb)
    for which in (N.BUF_KEYS, N.BUF_PERM, N.BUF_SORTED_KEYS, N.BUF_CELL_BEGIN, N.BUF_DENSITY):
        assert torch.equal(a.ctx.read_buffer(which), b.ctx.read_buffer(which)), which
    if mode == "PIPE":
        assert torch.equal(a.ctx.read_buffer(N.BUF_RNG), b.ctx.read_buffer(N.BUF_RNG))
    # and a further single step after the k-step call
    a.ctx.step(1)
    b.ctx.step(1)
    for ta, tb in zip(a.ctx.download(), b.ctx.download()):
        assert torch.equal(ta, tb)


def test_multi_step_calls_across_lane_and_slab_switches(gpu):
    """Calls of several steps with the lanes and the slab setting changed between them: with
    one lane the force passes inside a call also add the next sort's tile sums (the scan
    then reads the array they went to and clears the other), with four lanes the sort adds
    them up itself, and a call's first sort always does.  The states equal single-step calls
    with the same switches, bit for bit, and a call that ends on an odd step leaves nothing
    behind for the next."""
    w = box_uniform(20000, seed=35)
    pos = f32(w.state.position)
    vel = f32(np.random.default_rng(35).normal(0.0, 0.5, size=pos.shape))
    plan = [(1, True, 3), (4, True, 4), (1, True, 5), (1, False, 2), (1, True, 3), (2, True, 1), (1, True, 4)]

    def make():
        s = NaiveSPHStrategy(w.params, constants=w.constants, device=gpu, grid_voxel=2.0 ** -6)
        upload(s, pos, vel)
        return s

    a, b = make(), make()
    for lanes, slabs, k in plan:
        for s in (a, b):
            s.ctx.set_lanes(lanes)
            s.ctx.set_slabs(slabs)
        a.ctx.step(k)
        for _ in range(k):
            b.ctx.step(1)
        for ta, tb in zip(a.ctx.download(), b.ctx.download()):
            assert torch.equal(ta, tb), (lanes, slabs, k)
    for which in (N.BUF_KEYS, N.BUF_PERM, N.BUF_SORTED_KEYS, N.BUF_CELL_BEGIN, N.BUF_DENSITY):
        assert torch.equal(a.ctx.read_buffer(which), b.ctx.read_buffer(which)), which


def test_batch_steps_in_one_call_equal_single_calls(gpu):
    """The batched form (per-instance cell offsets in one counting sort): one sph_step(6) call
    equals six single-step calls bitwise, instances with 1 and 4 lanes."""
    from sph_amd.batched import BatchedSPH

    ws = [box_uniform(n, seed=40 + k, h=h) for k, (n, h) in enumerate([(9000, 0.0457), (3000, 0.03), (12000, 0.05)])]

    def make():
        return BatchedSPH([(x.params, x.constants) for x in ws], device=gpu, cell_scale=0.5, lanes=[1, 4, 1])

    states = [SimulationState(f32(x.state.position), f32(np.random.default_rng(41).normal(0.0, 0.5, (len(x.state.position), 3))),
                              np.zeros(len(x.state.position))) for x in ws]
    a, b = _same_after(make, lambda s: s.upload(states), 6)
    for ta, tb in zip(a.ctx.download(), b.ctx.download()):
        assert torch.equal(ta, tb)
    assert torch.equal(a.ctx.read_buffer(N.BUF_CELL_BEGIN), b.ctx.read_buffer(N.BUF_CELL_BEGIN))
