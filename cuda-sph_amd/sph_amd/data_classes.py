"""State and parameter types of the SPH step.

Same class names, fields, defaults and pipe-table format as the reference's
common/data_classes.py:7-85, so code written against the reference's types works
unchanged.  Mutable defaults use default_factory (the reference shares one default
instance between all objects).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np


@dataclass
class Segment:
    """One pipe segment along +x (data_classes.py:7-27)."""

    start_point: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    start_radius: float = 1.0
    end_radius: float = 1.0
    length: float = 1.0

    def to_numpy(self) -> np.ndarray:
        """Row [x, y, z, start_radius, length] (data_classes.py:13-16)."""
        x, y, z = self.start_point
        return np.array([x, y, z, self.start_radius, self.length], dtype=np.float64)

    def radius_at(self, t: float) -> float:
        """Radius at axial coordinate t, linear from start to end (data_classes.py:18-27)."""
        if self.start_radius == self.end_radius:
            return self.start_radius
        frac = (t - self.start_point[0]) / self.length
        return self.start_radius + (self.end_radius - self.start_radius) * frac


@dataclass(frozen=True)
class Pipe:
    """Chain of segments (data_classes.py:30-67)."""

    segments: List[Segment] = field(default_factory=list)

    def to_numpy(self) -> np.ndarray:
        """Collision table: one row per segment plus a closing row at the far end that
        carries the last segment's end radius (data_classes.py:34-44).  Only the last
        segment's end radius is used; between rows the wall runs from one start radius to
        the next (util_kernels.py:56-69)."""
        if not self.segments:
            return np.asarray([])
        rows = [s.to_numpy() for s in self.segments]
        closing = self.segments[-1].to_numpy()
        closing[0] += closing[4]
        closing[3] = self.segments[-1].end_radius
        rows.append(closing)
        return np.stack(rows)

    def get_length(self) -> float:
        total = 0.0
        for s in self.segments:
            total += s.length
        return total

    def find_segment(self, t: float) -> int:
        """Index of the first segment with start <= t <= end, -1 if none (data_classes.py:48-57)."""
        for i, s in enumerate(self.segments):
            if s.start_point[0] <= t <= s.start_point[0] + s.length:
                return i
        return -1

    def radius_at(self, t: float) -> float:
        return self.segments[self.find_segment(t)].radius_at(t)


def _vec(values) -> np.ndarray:
    return np.asarray(values)


@dataclass(frozen=True)
class SimulationParameters:
    """Run parameters (data_classes.py:70-78).  dt of the step is 1 / fps."""

    particle_count: int = 100
    external_force: np.ndarray = field(default_factory=lambda: _vec([0, 0, 0]))
    duration: int = 10
    fps: int = 20
    pipe: Pipe = field(default_factory=lambda: Pipe([Segment()]))
    space_size: np.ndarray = field(default_factory=lambda: _vec([1, 1, 1]))
    voxel_size: np.ndarray = field(default_factory=lambda: _vec([1, 1, 1]))


@dataclass(frozen=True)
class SimulationState:
    """Particle state (data_classes.py:81-85): position (n, 3), velocity (n, 3), density (n)."""

    position: np.ndarray = field(default_factory=lambda: np.asarray([[1, 1, 1]]))
    velocity: np.ndarray = field(default_factory=lambda: np.asarray([[1, 1, 1]]))
    density: np.ndarray = field(default_factory=lambda: np.asarray([1]))
