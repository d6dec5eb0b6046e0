"""Fluent construction of a Pipe, with the interface of the reference's
common/pipe_builder.py:6-161 (same method names, argument meaning, assertion messages and
resulting segments)."""
from __future__ import annotations

from typing import Optional, Tuple

from .data_classes import Pipe, Segment


class PipeBuilder:
    """Starts from one default Segment; configure it with the with_* methods, then append
    segments that continue from the previous one (pipe_builder.py:6-20)."""

    _FIRST_SEGMENT_MESSAGE = "First segment can't be change after adding new segments"
    _NEGATIVE_RADIUS_MESSAGE = "Radius must be positive"
    _NEGATIVE_CHANGE_MESSAGE = "Change must be positive"
    _NEGATIVE_LENGTH_MESSAGE = "Length must be positive"

    def __init__(self) -> None:
        self._editing_first = True
        self._pipe = Pipe(segments=[Segment()])

    @property
    def _first(self) -> Segment:
        return self._pipe.segments[0]

    # -- first segment (pipe_builder.py:22-67) --------------------------------------------
    def with_starting_position(self, position: Tuple[float, float, float]) -> "PipeBuilder":
        assert self._editing_first, self._FIRST_SEGMENT_MESSAGE
        self._first.start_point = position
        return self

    def with_starting_radius(self, radius: float) -> "PipeBuilder":
        assert self._editing_first, self._FIRST_SEGMENT_MESSAGE
        assert radius > 0, self._NEGATIVE_RADIUS_MESSAGE
        self._first.start_radius = radius
        return self

    def with_ending_radius(self, radius: float) -> "PipeBuilder":
        assert self._editing_first, self._FIRST_SEGMENT_MESSAGE
        assert radius > 0, self._NEGATIVE_RADIUS_MESSAGE
        self._first.end_radius = radius
        return self

    def with_starting_length(self, length: float) -> "PipeBuilder":
        assert self._editing_first, self._FIRST_SEGMENT_MESSAGE
        assert length > 0, self._NEGATIVE_LENGTH_MESSAGE
        self._first.length = length
        return self

    # -- appended segments (pipe_builder.py:69-118) ---------------------------------------
    def add_roller_segment(self, length) -> "PipeBuilder":
        """A cylinder continuing at the previous segment's end radius."""
        self._editing_first = False
        assert length > 0, self._NEGATIVE_LENGTH_MESSAGE
        prev = self._pipe.segments[-1]
        x, y, z = prev.start_point
        self._pipe.segments.append(Segment(start_point=(x + prev.length, y, z), start_radius=prev.end_radius,
                                           end_radius=prev.end_radius, length=length))
        return self

    def add_lessening_segment(self, length, change) -> "PipeBuilder":
        """A truncated cone whose end radius is smaller by change."""
        self.add_roller_segment(length)
        assert change > 0, self._NEGATIVE_CHANGE_MESSAGE
        last = self._pipe.segments[-1]
        assert change < last.end_radius, "After change radius must be positive"
        last.end_radius = last.end_radius - change
        return self

    def add_increasing_segment(self, length, change) -> "PipeBuilder":
        """A truncated cone whose end radius is larger by change."""
        self.add_roller_segment(length)
        assert change > 0, self._NEGATIVE_CHANGE_MESSAGE
        last = self._pipe.segments[-1]
        last.end_radius = last.end_radius + change
        return self

    # -- placement (pipe_builder.py:120-155) ----------------------------------------------
    def transform(self, space_size_x: float, space_size_yz: float, max_radius: Optional[float] = None) -> "PipeBuilder":
        """Scales lengths so the pipe spans [0, space_size_x] along x, puts its axis at
        y = z = space_size_yz / 2 and scales radii so the largest is max_radius
        (space_size_yz / 2 when not given)."""
        segs = self._pipe.segments
        scale_x = space_size_x / sum(s.length for s in segs)
        x = 0.0
        for s in segs:
            s.start_point = (x, space_size_yz / 2.0, space_size_yz / 2.0)
            s.length = s.length * scale_x
            x += s.length
        if not max_radius:
            max_radius = space_size_yz / 2.0
        widest = 0.0
        for s in segs:
            widest = max(widest, s.start_radius, s.end_radius)
        scale_r = max_radius / widest
        for s in segs:
            s.start_radius = s.start_radius * scale_r
            s.end_radius = s.end_radius * scale_r
        return self

    def get_result(self) -> Pipe:
        return self._pipe
