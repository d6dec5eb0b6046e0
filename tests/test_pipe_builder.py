"""PipeBuilder against the expectations of common/tests/test_pipe_builder.py:10-133."""
import numpy as np
import pytest

from sph_amd.data_classes import Pipe, Segment
from sph_amd.pipe_builder import PipeBuilder


def test_basic_builder():
    assert PipeBuilder().get_result() == Pipe([Segment()])


def test_setting_first_segment():
    pipe = PipeBuilder().with_starting_position((1, 2, 3)).with_starting_length(20).with_starting_radius(4) \
        .with_ending_radius(5).get_result()
    assert pipe == Pipe([Segment(start_point=(1, 2, 3), start_radius=4, end_radius=5, length=20)])


@pytest.mark.parametrize("call,msg", [
    (lambda: PipeBuilder().with_starting_length(-10), PipeBuilder._NEGATIVE_LENGTH_MESSAGE),
    (lambda: PipeBuilder().with_starting_radius(-1), PipeBuilder._NEGATIVE_RADIUS_MESSAGE),
    (lambda: PipeBuilder().with_ending_radius(-1), PipeBuilder._NEGATIVE_RADIUS_MESSAGE),
    (lambda: PipeBuilder().add_roller_segment(20).with_starting_radius(1), PipeBuilder._FIRST_SEGMENT_MESSAGE),
    (lambda: PipeBuilder().add_lessening_segment(1, -10), PipeBuilder._NEGATIVE_CHANGE_MESSAGE),
])
def test_assertion_messages(call, msg):
    with pytest.raises(AssertionError) as err:
        call()
    assert str(err.value) == msg


def test_negative_end_after_lessening():
    with pytest.raises(AssertionError):
        PipeBuilder().add_lessening_segment(1, -1)
    with pytest.raises(AssertionError):
        PipeBuilder().add_lessening_segment(1, 2)  # end radius would be -1


def test_segments():
    assert PipeBuilder().with_ending_radius(4).add_roller_segment(30).get_result() == \
        Pipe([Segment(end_radius=4), Segment((1, 0, 0), 4, 4, 30)])
    assert PipeBuilder().with_ending_radius(7).add_lessening_segment(length=1, change=6).get_result() == \
        Pipe([Segment(end_radius=7), Segment((1, 0, 0), 7, 1, 1)])
    assert PipeBuilder().with_ending_radius(5).add_increasing_segment(length=1, change=5).get_result() == \
        Pipe([Segment(end_radius=5), Segment((1, 0, 0), 5, 10, 1)])
    assert PipeBuilder().with_starting_radius(10).with_ending_radius(5).add_roller_segment(length=1) \
        .add_increasing_segment(1, 5).get_result() == \
        Pipe([Segment(start_radius=10, end_radius=5), Segment((1, 0, 0), 5, 5, 1), Segment((2, 0, 0), 5, 10, 1)])


def test_pipe_transform():
    # common/tests/test_pipe_builder.py:112-133
    pipe = PipeBuilder().add_increasing_segment(2., 3.).add_roller_segment(1.).add_increasing_segment(2., 3.) \
        .transform(600, 600, 70).get_result()
    expected = PipeBuilder().with_starting_position((0., 300., 300.)).with_starting_radius(10.) \
        .with_ending_radius(10.).with_starting_length(100.).add_increasing_segment(200., 30.) \
        .add_roller_segment(100.).add_increasing_segment(200., 30.).get_result().to_numpy()
    assert np.all(pipe.to_numpy() == expected)
