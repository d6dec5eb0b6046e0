"""Pins the oracle's VOXEL formulation and its cone-wall collision with facts that the
mathematics fixes, independent of the oracle's own code.

VOXEL formulation (voxel_kernels.py:29-85, :108-211):
  - a three-particle chain worked by hand: density without the particle itself, the
    pressure factor p_i/rho_i^2 + p_j/rho_j^2 with no mass, the viscosity factor
    m mu / (rho_i rho_j) and the clip min(1, .) per component;
  - pair antisymmetry: with complete lists (fewer than 32 neighbours) and no clipped
    component, sum_i P_i = 0 and sum_i V_i = 0 (a symmetric factor times an
    antisymmetric kernel gradient / velocity difference);
  - mass scaling with rho0 = 0: rho ~ m, P ~ 1/m and V ~ 1/m (the NAIVE pressure ~ m);
  - an isolated particle has rho = 0, and the step leaves it non-finite;
  - the 32-neighbour cap and the enumeration order on a hand-built layout whose expected
    list is written out literally.

Cone collision (util_kernels.py:113-121, 152-204): the backtrack distance
d / (sin a + 0.001) with d from plane geometry, |v'| = |v|, v' + v parallel to the wall
line, and the re-advance by the backtracked distance along v'.
"""
import math

import numpy as np
import pytest

import oracle
from sph_amd.constants import SPHConstants

W9 = 315.0 / (64.0 * math.pi)  # poly6 constant for h = 1
G6 = -45.0 / math.pi           # spiky gradient constant for h = 1
L6 = 45.0 / math.pi            # viscosity Laplacian constant for h = 1


def _voxel_terms(pos, vel, consts, voxel, space):
    """rho, P, V of the VOXEL formulation from the oracle's own lists."""
    pos = np.asarray(pos, dtype=np.float64)
    dims = oracle.space_dims(space, voxel, True)
    tdims = oracle.space_dims(space, voxel, False)
    keys = oracle.voxel_keys(pos, voxel, dims)
    perm = oracle.sort_map(keys)
    begin = oracle.voxel_begin(keys[perm], int(np.prod(dims)))
    n = len(pos)
    nb = np.full((n, 32), -1, dtype=np.int32)
    nn = np.zeros(n, dtype=np.int32)
    for p in range(n):
        lst = oracle.voxel_neighbours(pos, begin, perm, voxel, tdims, consts.h, p)
        nn[p] = len(lst)
        nb[p, :len(lst)] = lst
    rho, P, V = oracle.voxel_forces(pos, vel, nb, nn, consts)
    return rho, P, V, nb, nn


def test_voxel_three_particle_chain_by_hand():
    """A - B - C on a line, |AB| = 0.5, |BC| = 0.8, |AC| = 1.3 > h = 1."""
    c = SPHConstants(mass=0.5, rho0=0.01, h=1.0, visc=3.0, k=2.0, damp=0.5)
    pos = np.array([[2.2, 2.5, 2.5], [2.7, 2.5, 2.5], [3.5, 2.5, 2.5]])
    vel = np.array([[0.0, 0.0, 0.0], [0.1, -0.1, 0.001], [0.02, 0.3, -0.2]])
    rho, P, V, nb, nn = _voxel_terms(pos, vel, c, np.ones(3), np.full(3, 8.0))
    assert nn.tolist() == [2, 3, 2]  # self included in every list
    # W(0.5) = W9 (1 - 0.25)^3, W(0.8) = W9 (1 - 0.64)^3; the particle itself is left out
    wab, wbc = W9 * 0.75 ** 3, W9 * 0.36 ** 3
    ra, rb, rc = 0.5 * wab, 0.5 * (wab + wbc), 0.5 * wbc
    assert np.allclose(rho, [ra, rb, rc], rtol=1e-14, atol=0)
    pa, pb, pc = 2.0 * (ra - 0.01), 2.0 * (rb - 0.01), 2.0 * (rc - 0.01)
    fa, fb, fc = pa / ra ** 2, pb / rb ** 2, pc / rc ** 2
    # grad W(d) = G6 (1 - r)^2 d / r: for d = x_A - x_B = (-0.5, 0, 0) it is -G6 0.25 along x
    g_ab = -G6 * 0.25   # x component of grad W(x_A - x_B)
    g_bc = -G6 * 0.04   # x component of grad W(x_B - x_C), d = (-0.8, 0, 0)
    expect_P = np.array([[(fa + fb) * g_ab, 0, 0],
                         [-(fb + fa) * g_ab + (fb + fc) * g_bc, 0, 0],
                         [-(fc + fb) * g_bc, 0, 0]])
    assert np.allclose(P, expect_P, rtol=1e-13, atol=1e-12)
    # V_i = min(1, sum_j (v_j - v_i) / rho_j lap W(r) m mu / rho_i), lap W = L6 (1 - r)
    mmu = 0.5 * 3.0
    raw_a = (vel[1] - vel[0]) / rb * L6 * 0.5 * mmu / ra
    raw_b = ((vel[0] - vel[1]) / ra * L6 * 0.5 + (vel[2] - vel[1]) / rc * L6 * 0.2) * mmu / rb
    raw_c = (vel[1] - vel[2]) / rb * L6 * 0.2 * mmu / rc
    raw = np.array([raw_a, raw_b, raw_c])
    assert raw[0, 0] > 1 and raw[0, 1] < -1 and raw[2, 0] > 1 and raw[2, 2] > 1  # clipped and unclipped cases
    assert np.allclose(V, np.minimum(1.0, raw), rtol=1e-13, atol=1e-13)
    assert V[0, 0] == 1.0 and V[2, 0] == 1.0 and V[2, 2] == 1.0


def _cluster(n, seed, spread):
    rng = np.random.default_rng(seed)
    pos = 2.0 + rng.uniform(-spread, spread, size=(n, 3))
    vel = rng.normal(0.0, 1e-4, size=(n, 3))
    return pos, vel


def test_voxel_pair_sums_are_antisymmetric():
    """With every list complete the pressure and viscosity pair terms cancel in the sum."""
    c = SPHConstants(mass=0.3, rho0=0.05, h=1.0, visc=2.0, k=4.0, damp=0.5)
    pos, vel = _cluster(24, 3, 0.9)
    rho, P, V, nb, nn = _voxel_terms(pos, vel, c, np.ones(3), np.full(3, 8.0))
    d = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    assert np.array_equal(nn, np.sum(d <= 1.0, axis=1))  # complete lists (all < 32)
    assert np.all(nn > 1) and np.any(nn < len(pos))       # every particle has a partner, not all pairs
    assert np.all(np.abs(V) < 1)                          # nothing clipped
    scale_p = np.sum(np.abs(P))
    scale_v = np.sum(np.abs(V))
    assert np.all(np.abs(P.sum(axis=0)) < 1e-13 * scale_p)
    assert np.all(np.abs(V.sum(axis=0)) < 1e-13 * scale_v)
    # and neither is trivially zero
    assert scale_p > 0 and scale_v > 0


def test_voxel_terms_scale_without_mass_in_pressure():
    """rho0 = 0: p = K rho.  VOXEL P = sum K (1/rho_i + 1/rho_j) grad W ~ 1/m and
    V ~ m / (rho rho) ~ 1/m; the NAIVE P = m sum K (rho_i + rho_j) / (2 rho_j) grad W ~ m."""
    pos, vel = _cluster(20, 5, 0.8)
    c1 = SPHConstants(mass=0.4, rho0=0.0, h=1.0, visc=1.5, k=3.0, damp=0.5)
    c2 = SPHConstants(mass=0.8, rho0=0.0, h=1.0, visc=1.5, k=3.0, damp=0.5)
    r1, P1, V1, _, _ = _voxel_terms(pos, vel, c1, np.ones(3), np.full(3, 8.0))
    r2, P2, V2, _, _ = _voxel_terms(pos, vel, c2, np.ones(3), np.full(3, 8.0))
    assert np.all(np.abs(V1) < 1) and np.all(np.abs(V2) < 1)
    assert np.allclose(r2, 2 * r1, rtol=1e-14)
    assert np.allclose(P2, P1 / 2, rtol=1e-12, atol=1e-12 * np.abs(P1).max())
    assert np.allclose(V2, V1 / 2, rtol=1e-12, atol=1e-12 * np.abs(V1).max())
    n1 = oracle.naive_pressure(pos, oracle.naive_density(pos, c1), c1)
    n2 = oracle.naive_pressure(pos, oracle.naive_density(pos, c2), c2)
    assert np.allclose(n2, 2 * n1, rtol=1e-12, atol=1e-12 * np.abs(n1).max())


def test_voxel_isolated_particle_goes_non_finite():
    """voxel_kernels.py:127-130 leaves the particle itself out of its density, so a particle
    with no neighbour has rho = 0 and the integration divides by it (base_kernels.py:47-52).
    The NAIVE density includes the particle itself and stays finite."""
    c = SPHConstants(mass=0.5, rho0=0.01, h=1.0, visc=3.0, k=2.0, damp=0.5)
    pos = np.array([[2.2, 2.5, 2.5], [2.7, 2.5, 2.5], [6.0, 6.0, 6.0]])
    vel = np.zeros((3, 3))
    kw = dict(consts=c, ext=[0.0, -9.81, 0.0], dt=1e-3, space=[8.0, 8.0, 8.0], voxel=[1.0, 1.0, 1.0])
    p, v, rho, f, _ = oracle.step(pos, vel, formulation="VOXEL", **kw)
    assert rho[2] == 0.0 and rho[0] > 0 and rho[1] > 0
    assert np.all(np.isfinite(p[:2])) and np.all(np.isfinite(v[:2]))
    assert not np.all(np.isfinite(p[2])) and not np.all(np.isfinite(v[2]))
    p, v, rho, f, _ = oracle.step(pos, vel, formulation="NAIVE", **kw)
    assert rho[2] > 0 and np.all(np.isfinite(p)) and np.all(np.isfinite(v))


# -- the 32-neighbour cap and its enumeration order (voxel_kernels.py:29-85) -------------
# Probe particle 20 at (1.5, 1.5, 1.5) in voxel (1, 1, 1); h = voxel = 1.  The 27 cells are
# visited in (dx, dy, dz) order with dz innermost, so the cells used come in this order:
#   (0, 1, 1) dx=-1: ids 33 5 17 40 2 28 within h, 11 36 beyond h
#   (1, 0, 1) dy=-1: ids 7 30 14 1 25 38 9 within h, 22 beyond h
#   (1, 1, 0) dz=-1: ids 3 19 27 35 12 41
#   (1, 1, 1) self : ids 20 (the probe) 6 31 15 39
#   (1, 1, 2) dz=+1: ids 26 0 34 10 18 4 37 23 29 13
#   (2, 1, 1) dx=+1: ids 8 16 21 24 32 (never reached)
# Within a cell the map is in id order (the stable sort), the probe is listed, and the cap
# cuts the (1, 1, 2) cell after its eighth accepted id.
CAP_GROUPS = [
    ((0.8, 1.5, 1.5), [33, 5, 17, 40, 2, 28], [((0.1, 1.1, 1.1), 11), ((0.2, 1.05, 1.9), 36)]),
    ((1.5, 0.8, 1.5), [7, 30, 14, 1, 25, 38, 9], [((1.5, 0.1, 1.5), 22)]),
    ((1.5, 1.5, 0.8), [3, 19, 27, 35, 12, 41], []),
    (None, [20, 6, 31, 15, 39], []),
    ((1.5, 1.5, 2.2), [26, 0, 34, 10, 18, 4, 37, 23, 29, 13], []),
    ((2.2, 1.5, 1.5), [8, 16, 21, 24, 32], []),
]
CAP_SELF = [(1.5, 1.5, 1.5), (1.4, 1.6, 1.5), (1.6, 1.4, 1.45), (1.3, 1.3, 1.7), (1.7, 1.7, 1.3)]
CAP_EXPECTED = [2, 5, 17, 28, 33, 40, 1, 7, 9, 14, 25, 30, 38, 3, 12, 19, 27, 35, 41, 6, 15, 20, 31, 39,
                0, 4, 10, 13, 18, 23, 26, 29]
CAP_PROBE = 20


def cap_layout():
    """Positions (42, 3) of the hand-built layout above, indexed by particle id."""
    pos = np.zeros((42, 3))
    for base, ids, far in CAP_GROUPS:
        for k, pid in enumerate(ids):
            pos[pid] = CAP_SELF[k] if base is None else np.add(base, np.multiply(k, (0.01, -0.01, 0.02)))
        for p, pid in far:
            pos[pid] = p
    return pos


def test_voxel_cap_order_literal():
    pos = cap_layout()
    d = np.linalg.norm(pos - pos[CAP_PROBE], axis=1)
    # the layout is what the comment says: 39 particles within h of the probe (the last cell's
    # five never reached), 42 in all
    assert np.sum(d <= 1.0) == 39 and len(set(map(tuple, pos))) == 42
    voxel, space = np.ones(3), np.full(3, 4.0)
    dims = oracle.space_dims(space, voxel, True)
    keys = oracle.voxel_keys(pos, voxel, dims)
    perm = oracle.sort_map(keys)
    begin = oracle.voxel_begin(keys[perm], int(np.prod(dims)))
    got = oracle.voxel_neighbours(pos, begin, perm, voxel, oracle.space_dims(space, voxel, False), 1.0, CAP_PROBE)
    assert got.tolist() == CAP_EXPECTED


# -- cone collision (util_kernels.py:113-121, 152-204) ----------------------------------
CONE = np.array([[0.0, 0.5, 0.5, 0.1, 1.0], [1.0, 0.5, 0.5, 0.2, 1.0]])  # radius 0.1 -> 0.2 over x in [0, 1]


def _cone_case(x, rad, theta, v):
    p = np.array([x, 0.5 + rad * math.cos(theta), 0.5 + rad * math.sin(theta)])
    return p, np.asarray(v, dtype=np.float64)


@pytest.mark.parametrize("x,rad,theta,v", [
    (0.4, 0.16, 0.7, "radial"),
    (0.25, 0.13, 2.9, "swirl"),
    (0.8, 0.185, -1.2, "back"),
])
def test_cone_collision_closed_form(x, rad, theta, v):
    c, s = math.cos(theta), math.sin(theta)
    vel = {"radial": (0.3, 0.5 * c, 0.5 * s),
           "swirl": (-0.2, 0.4 * c - 0.3 * s, 0.4 * s + 0.3 * c),
           "back": (-0.6, 0.05 * c, 0.05 * s)}[v]
    p, vel = _cone_case(x, rad, theta, vel)
    assert oracle.is_out_of_pipe(p, CONE, 0)
    # the wall line in the particle's half-plane: radius 0.1 + 0.1 x, slope 0.1; the distance
    # from the point (x, rad) to it in that plane
    d = (rad - (0.1 + 0.1 * x)) / math.sqrt(1.0 + 0.1 ** 2)
    edge = np.array([1.0, 0.1 * c, 0.1 * s])
    e_hat = edge / np.linalg.norm(edge)
    speed = np.linalg.norm(vel)
    sin_a = np.linalg.norm(np.cross(vel, e_hat)) / speed
    back = d / (sin_a + 0.001)  # distance travelled back along -v to the collision point
    cp = p - vel / speed * back
    p1, v1 = oracle.solve_collision(p, vel, CONE, 0)
    # reflection about the wall line: same speed, v' + v along the edge
    assert math.isclose(np.linalg.norm(v1), speed, rel_tol=1e-13)
    assert np.linalg.norm(np.cross(v1 + vel, e_hat)) < 1e-13 * speed
    assert np.allclose(v1, 2 * np.dot(e_hat, vel) * e_hat - vel, rtol=0, atol=1e-14)
    # re-advance from the collision point by the backtracked distance along v'
    assert np.allclose(p1, cp + v1 / speed * back, rtol=0, atol=1e-13)


def test_cone_collision_backtrack_constant():
    """The 0.001 in d / (sin a + 0.001) dominates for a particle moving nearly along the
    wall (sin a = 1e-4 here): the backtrack is d / 0.0011, not d / sin a."""
    theta = 0.3
    c, s = math.cos(theta), math.sin(theta)
    e_hat = np.array([1.0, 0.1 * c, 0.1 * s]) / math.sqrt(1.01)
    out = np.array([-0.1, 1.0 * c, 1.0 * s])  # in the half-plane, normal to the wall line
    out = out - np.dot(out, e_hat) * e_hat
    out /= np.linalg.norm(out)
    ang = 1e-4
    vel = 0.7 * (math.cos(ang) * e_hat + math.sin(ang) * out)
    p, _ = _cone_case(0.5, 0.1505, theta, (0, 0, 0))
    d = (0.1505 - 0.15) / math.sqrt(1.01)
    p1, v1 = oracle.solve_collision(p, vel, CONE, 0)
    back = d / (math.sin(ang) + 0.001)
    cp = p - vel / 0.7 * back
    assert np.allclose(v1, 2 * np.dot(e_hat, vel) * e_hat - vel, atol=1e-14)
    assert np.allclose(p1, cp + v1 / 0.7 * back, atol=1e-10)
