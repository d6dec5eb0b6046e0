"""Pins the oracle's FP64 adjoint (reverse mode of the NAIVE box step and of rollouts) with
finite differences of the oracle's own forward map, which the other oracle tests pin.

The forward map is piecewise smooth: the box collision (base_kernels.py:75-98) moves a
component that left the box to 1e-3 (or space - 1e-3) and scales its velocity by -damp,
so the map jumps where the bounce decision changes.  Finite differences that should see
the piece the adjoint differentiates replay the base point's bounce decisions
(branch_mode 2).  Checks are directional: <adjoint(g), delta> must equal
d/de <g, F(z + e delta)> for random delta, which pins every term of the pullback.
"""
from dataclasses import replace

import numpy as np
import pytest

import oracle
from sph_amd import workloads as W

EPS = 1e-6


def _setup(n, seed, steps, near_walls=False):
    w = W.box_uniform(n, seed=seed, steps=steps)
    p = w.params
    kw = dict(ext=p.external_force, dt=1.0 / p.fps, space=p.space_size)
    rng = np.random.default_rng(seed + 100)
    x = w.state.position.copy()
    v = rng.normal(0.0, 0.3, size=x.shape)
    if near_walls:
        # a slab against the floor and two side walls, moving into them: several particles
        # leave the box in the first step
        x = rng.uniform([0.0, 0.0, 0.0], [0.25, 0.03, 0.25], size=x.shape)
        v[:, 1] = -np.abs(v[:, 1]) * 3
        v[:, 0] -= 0.5
    return w.constants, kw, x, v


def _loss_and_dir(x, v, g_x, g_v, consts, kw, steps, branches):
    xT, vT = oracle.rollout(x, v, consts=consts, steps=steps, branches=branches, branch_mode=2, **kw)
    return float(np.sum(g_x * xT) + np.sum(g_v * vT))


@pytest.mark.parametrize("near_walls", [False, True])
def test_single_step_adjoint_directional(near_walls):
    c, kw, x, v = _setup(400, 5, 1, near_walls)
    rng = np.random.default_rng(9)
    gx1 = rng.normal(size=x.shape)
    gv1 = rng.normal(size=x.shape)
    gx0, gv0, gmu = oracle.adjoint_step(x, v, gx1, gv1, consts=c, **kw)
    br = np.zeros(3 * x.shape[0], np.uint8)
    oracle.rollout(x, v, consts=c, steps=1, branches=br, branch_mode=1, **kw)
    if near_walls:
        assert np.count_nonzero(br) >= 5, "the near-wall case must bounce"
    for trial in range(3):
        dx = rng.normal(size=x.shape) * (trial != 1)
        dv = rng.normal(size=x.shape) * (trial != 0)
        lp = _loss_and_dir(x + EPS * dx, v + EPS * dv, gx1, gv1, c, kw, 1, br)
        lm = _loss_and_dir(x - EPS * dx, v - EPS * dv, gx1, gv1, c, kw, 1, br)
        fd = (lp - lm) / (2 * EPS)
        ad = float(np.sum(gx0 * dx) + np.sum(gv0 * dv))
        assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1.0), (trial, ad, fd)
    lp = _loss_and_dir(x, v, gx1, gv1, replace(c, visc=c.visc + EPS), kw, 1, br)
    lm = _loss_and_dir(x, v, gx1, gv1, replace(c, visc=c.visc - EPS), kw, 1, br)
    fd = (lp - lm) / (2 * EPS)
    assert abs(gmu - fd) <= 1e-6 * max(abs(fd), 1e-3), (gmu, fd)


def test_bounced_component_has_no_position_cotangent():
    # a particle alone, leaving through the floor: x_y' is the constant 1e-3 and v_y' = -damp v_y
    c, kw, _, _ = _setup(1, 0, 1)
    x = np.array([[0.5, 0.0002, 0.5]])
    v = np.array([[0.0, -1.0, 0.0]])
    gx0, gv0, _ = oracle.adjoint_step(x, v, np.array([[0.0, 1.0, 0.0]]), np.zeros((1, 3)), consts=c, **kw)
    assert gx0[0, 1] == 0.0 and gv0[0, 1] == 0.0
    gx0, gv0, _ = oracle.adjoint_step(x, v, np.zeros((1, 3)), np.array([[0.0, 1.0, 0.0]]), consts=c, **kw)
    dt = kw["dt"]
    assert gx0[0, 1] == 0.0 and gv0[0, 1] == pytest.approx(-c.damp)
    # an unbounced component: x' = x + v' dt, v' = v + F / rho dt, single particle: F = f
    gx0, gv0, _ = oracle.adjoint_step(x, v, np.array([[1.0, 0.0, 0.0]]), np.zeros((1, 3)), consts=c, **kw)
    assert gx0[0, 0] == 1.0 and gv0[0, 0] == pytest.approx(dt)


@pytest.mark.parametrize("near_walls", [False, True])
def test_rollout_adjoint_matches_finite_differences(near_walls):
    steps, every = 12, 5  # segments of 5, 5, 2 steps
    c, kw, x, v = _setup(300, 7, steps, near_walls)
    rng = np.random.default_rng(3)
    target = x + rng.normal(0, 0.01, size=x.shape)
    loss, g, gmu = oracle.rollout_adjoint(x, v, consts=c, steps=steps, target=target, every=every, **kw)
    br = np.zeros(steps * 3 * x.shape[0], np.uint8)
    xT, _ = oracle.rollout(x, v, consts=c, steps=steps, branches=br, branch_mode=1, **kw)
    assert loss == pytest.approx(np.mean(np.sum((xT - target) ** 2, axis=1)), rel=1e-12)
    if near_walls:
        assert np.count_nonzero(br) >= 10

    def L(vv, cc):
        xt, _ = oracle.rollout(x, vv, consts=cc, steps=steps, branches=br, branch_mode=2, **kw)
        return np.mean(np.sum((xt - target) ** 2, axis=1))

    for k in rng.choice(x.size, 6, replace=False):
        e = np.zeros(x.size)
        e[k] = EPS
        fd = (L(v + e.reshape(x.shape), c) - L(v - e.reshape(x.shape), c)) / (2 * EPS)
        assert abs(g.ravel()[k] - fd) <= 1e-5 * abs(fd) + 1e-12, (k, g.ravel()[k], fd)
    fd = (L(v, replace(c, visc=c.visc + EPS)) - L(v, replace(c, visc=c.visc - EPS))) / (2 * EPS)
    assert abs(gmu - fd) <= 1e-5 * abs(fd) + 1e-14, (gmu, fd)


def test_checkpoint_interval_does_not_change_the_gradient():
    c, kw, x, v = _setup(200, 11, 9)
    target = x + 0.01
    ref = oracle.rollout_adjoint(x, v, consts=c, steps=9, target=target, every=1, **kw)
    for every in (2, 4, 9, 20):
        got = oracle.rollout_adjoint(x, v, consts=c, steps=9, target=target, every=every, **kw)
        assert got[0] == ref[0]
        np.testing.assert_array_equal(got[1], ref[1])
        assert got[2] == ref[2]
