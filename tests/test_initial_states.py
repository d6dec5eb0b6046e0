"""The reference's run configuration (config.py) as seeded functions: parameter values,
the pipe it builds, and the bounds and statistics of the two start-state generators."""
import math

import numpy as np
import pytest

from sph_amd import initial_states as S
from sph_amd.constants import SPHConstants


def test_box_preset_matches_config():
    params, consts, state, mode = S.preset("BOX", particle_count=2000, seed=1)
    # config.py:15-37 and 56-65
    assert mode == "BOX"
    assert params.particle_count == 2000 and params.duration == 15 and params.fps == 20
    assert np.array_equal(params.external_force, [0.0, -2.0, 0.0])
    assert np.array_equal(params.space_size, [40.0, 40.0, 40.0])
    assert np.array_equal(params.voxel_size, [2.0, 2.0, 2.0])
    assert consts == SPHConstants(mass=1.0, rho0=1.0, h=2.0, visc=0.5, k=10.0, damp=0.7)
    p, v = state.position, state.velocity
    assert p.shape == (2000, 3) and v.shape == (2000, 3) and np.all(state.density == 0)
    assert p[:, 0].min() >= 0 and p[:, 0].max() < 4.0          # 0.1 of the box along x
    assert p[:, 1].max() < 40.0 and p[:, 2].max() < 40.0 and p.min() >= 0
    base = np.array([1.5, -5.0, -5.0])
    assert np.all(np.abs(v - base) <= 0.5)
    assert np.allclose(v.mean(axis=0), base, atol=0.03)          # U[-0.5, 0.5) offsets
    assert abs(p[:, 1].mean() - 20.0) < 0.6


def test_pipe_preset_matches_config():
    params, consts, state, mode = S.preset("PIPE", particle_count=3000, seed=2)
    assert mode == "PIPE"
    assert np.array_equal(params.external_force, [2.0, 0.0, 0.0])
    assert np.array_equal(params.space_size, [40.0, 6.0, 6.0])
    pipe = params.pipe
    # config.py:68-76: five unit segments (radii 1, 1 -> 2.2, 2.2, 2.2 -> 1, 1) after the
    # default first one; transform (pipe_builder.py:116-154) spreads them over x in [0, 40]
    # with the axis at y = z = 3 and scales every radius so the largest (2.2) becomes 3
    segs = pipe.segments
    scale = 3.0 / 2.2
    assert len(segs) == 6
    assert math.isclose(pipe.get_length(), 40.0)
    assert [s.start_point[0] for s in segs] == pytest.approx([0, 40 / 6, 80 / 6, 120 / 6, 160 / 6, 200 / 6])
    assert all(s.start_point[1:] == (3.0, 3.0) for s in segs)
    assert [s.start_radius for s in segs] == pytest.approx([scale, scale, scale, 3.0, 3.0, scale])
    assert [s.end_radius for s in segs] == pytest.approx([scale, scale, 3.0, 3.0, scale, scale])
    p = state.position
    assert np.all(state.velocity == 0)
    assert p[:, 0].min() >= 0 and p[:, 0].max() < 40.0
    r = np.hypot(p[:, 1] - 3.0, p[:, 2] - 3.0)
    R = np.array([pipe.radius_at(x) for x in p[:, 0]])
    assert np.all(r <= 0.98 * R + 1e-12)
    # r = sqrt(u) R 0.98: (r / 0.98 R)^2 is uniform on [0, 1)
    u = (r / (0.98 * R)) ** 2
    assert abs(u.mean() - 0.5) < 0.03


def test_generators_are_seeded():
    a = S.start_state_box_wall(100, seed=5)
    b = S.start_state_box_wall(100, seed=5)
    c = S.start_state_box_wall(100, seed=6)
    assert np.array_equal(a.position, b.position) and np.array_equal(a.velocity, b.velocity)
    assert not np.array_equal(a.position, c.position)
    pipe = S.config_pipe()
    assert np.array_equal(S.start_state_inside_pipe(pipe, 50, seed=3).position,
                          S.start_state_inside_pipe(pipe, 50, seed=3).position)
