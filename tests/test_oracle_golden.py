"""The oracle against the reference's own known answers.

Values come from sim/tests/test_collisions.py:90-228 and common/tests/test_pipe.py:12-34
(the inputs and expected outputs those tests hold; the reference code is not run).  The
pipes are built with this package's PipeBuilder, which is pinned by
tests/test_pipe_builder.py against common/tests/test_pipe_builder.py.
"""
import math

import numpy as np
import pytest

import oracle
from sph_amd.data_classes import Pipe, Segment
from sph_amd.pipe_builder import PipeBuilder


def test_find_segment_known_answers():
    # test_collisions.py:90-106
    pos = np.array([[-0.1, 0, 2], [0.5, 3, 4], [1, 0, 0], [2.5, 0, -5], [5, 5, 5]], dtype=np.float64)
    pipe = PipeBuilder().add_roller_segment(1.0).add_increasing_segment(1.0, 3.0).get_result().to_numpy()
    got = [oracle.find_segment(p, pipe) for p in pos]
    assert got == [-1, 0, 1, 2, -1]


def test_vector_length_known_answers():
    # test_collisions.py:108-118
    got = [oracle.vector_length(v) for v in np.array([[0, 2, 2], [1, 1, 1]], dtype=np.float64)]
    assert got == [8 ** 0.5, 3 ** 0.5]


def test_distance_known_answers():
    # test_collisions.py:120-135
    a = np.array([[1, 1, 1], [4, 5, 6]], dtype=np.float64)
    b = np.array([[2, 3, 5], [-2, 5, 1]], dtype=np.float64)
    got = [oracle.distance_between_points(x, y) for x, y in zip(a, b)]
    assert got == [21 ** 0.5, 61 ** 0.5]


def test_x_at_segment_beginning_known_answers():
    # test_collisions.py:137-155
    pipe = PipeBuilder().add_increasing_segment(4, 1).add_lessening_segment(6, 1.5).get_result().to_numpy()
    assert [oracle.x_at_segment_beginning(pipe, s) for s in (0, 1, 2)] == [0, 1, 5]


def test_is_out_of_pipe_known_answers():
    # test_collisions.py:157-182
    pos = np.array([[0.5, 0.5, 0.5], [0.76, 0.9, 0.9], [0.01, 1, 0.5], [1.25, 0.8, 0.8], [1.5, 1.25, 1.40],
                    [1.99, 1.3, 1.6], [2.2, 1, 1], [2.6, 1, 1], [2.95, 0.2, 0.2]], dtype=np.float64)
    pipe = PipeBuilder().add_increasing_segment(1, 1).add_lessening_segment(1, 1.5).get_result().to_numpy()
    segs = [0, 0, 0, 1, 1, 1, 2, 2, 2]
    got = [oracle.is_out_of_pipe(p, pipe, s) for p, s in zip(pos, segs)]
    assert got == [False, True, True, False, True, True, False, True, False]


def test_collision_resolution_post_condition():
    # test_collisions.py:184-209: after solve_collision the particle is back inside
    pos = np.array([1.3, 1.5, 0.0])
    vel = np.array([-0.5, -0.0, 0.0])
    pipe = PipeBuilder().add_increasing_segment(1, 1).add_lessening_segment(2, 1.5).get_result().to_numpy()
    seg = oracle.find_segment(pos, pipe)
    assert oracle.is_out_of_pipe(pos, pipe, seg)
    p, v = oracle.solve_collision(pos, vel, pipe, seg)
    assert not oracle.is_out_of_pipe(p, pipe, oracle.find_segment(p, pipe))


def test_cylinder_reflection_closed_form():
    """Radial exit through a cylinder wall: the wall line is parallel to x, so
    cos(a) = 0, d = overshoot, back-track d / (1 + 0.001) (util_kernels.py:113-121), the
    radial velocity flips (util_kernels.py:191-193) and the particle re-advances the same
    distance: y' = R + d - 2 d / 1.001."""
    R, d = 1.0, 0.05
    pipe = Pipe([Segment((0, 0, 0), R, R, 2.0)]).to_numpy()
    pos = np.array([0.7, R + d, 0.0])
    vel = np.array([0.0, 0.8, 0.0])
    p, v = oracle.solve_collision(pos, vel, pipe, 0)
    assert np.allclose(v, [0.0, -0.8, 0.0], rtol=0, atol=1e-15)
    assert math.isclose(p[1], R + d - 2 * d / 1.001, rel_tol=1e-14)
    assert p[0] == 0.7 and p[2] == 0.0


def test_rand_position_inside_pipe_post_condition():
    # test_collisions.py:211-228: re-injected positions are inside the pipe
    pipe = PipeBuilder().add_increasing_segment(1, 1).add_lessening_segment(1, 1.5).get_result().to_numpy()
    n = 16
    rng_states = oracle.rng_init(43, n)
    pos = np.random.default_rng(0).random((n, 3)) * 3
    pos[:, 0] = 5.0  # past the outlet: collide_pipe re-injects at x = 0
    vel = np.zeros((n, 3))
    p, v, st = oracle.collide_pipe(pos, vel, pipe, rng_states)
    assert np.all(p[:, 0] == 0.0)
    assert not any(oracle.is_out_of_pipe(q, pipe, oracle.find_segment(q, pipe)) for q in p)
    assert not np.array_equal(st, rng_states)


def test_pipe_tables():
    # common/tests/test_pipe.py:12-34
    assert np.all(Segment(start_radius=4.5).to_numpy() == np.array([0, 0, 0, 4.5, 1]))
    assert np.all(Pipe([Segment()]).to_numpy() == np.array([[0, 0, 0, 1, 1], [1, 0, 0, 1, 1]]))
    pipe = PipeBuilder().with_starting_radius(3).add_roller_segment(1).add_increasing_segment(1, 2).get_result()
    assert np.all(pipe.to_numpy() == np.array([[0, 0, 0, 3, 1], [1, 0, 0, 1, 1], [2, 0, 0, 1, 1], [3, 0, 0, 3, 1]]))


def test_radius_table_semantics():
    """Between rows the wall runs from one start radius to the next
    (util_kernels.py:56-69): for a row of radius 1 followed by one of radius 3, the radius
    at the middle of the first segment is 2, and the cone formulas agree with linear
    interpolation in both directions."""
    up = Pipe([Segment((0, 0, 0), 1, 1, 2.0), Segment((2, 0, 0), 3, 3, 2.0)]).to_numpy()
    assert math.isclose(oracle.radius_in_position([1.0, 0, 0], up, 0), 2.0)
    down = Pipe([Segment((0, 0, 0), 3, 3, 2.0), Segment((2, 0, 0), 1, 1, 2.0)]).to_numpy()
    for x in (0.0, 0.5, 1.3, 2.0):
        assert math.isclose(oracle.radius_in_position([x, 0, 0], down, 0), 3.0 - x, rel_tol=1e-14, abs_tol=1e-14)
