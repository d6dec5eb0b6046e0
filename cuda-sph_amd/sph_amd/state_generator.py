"""Frame iterator with the protocol of sim/src/state_generator.py:14-56."""
from __future__ import annotations

import logging
from timeit import default_timer as timer

from .data_classes import SimulationParameters, SimulationState
from .strategy import make_strategy

logger = logging.getLogger(__name__)


class StateGenerator:
    """`for state in StateGenerator(start_state, params): ...` yields duration * fps states
    (state_generator.py:15-37).  The strategy name plays the role of config.SIM_STRATEGY."""

    def __init__(self, start_state: SimulationState, params: SimulationParameters, strategy: str = "NAIVE",
                 **strategy_kw) -> None:
        self.current_state = start_state
        self.current_frame_idx = 0
        self.n_frames = params.duration * params.fps
        self.sph_strategy = make_strategy(strategy, params, **strategy_kw)
        logger.info("Simulation parameters: %s", params)

    def __next__(self) -> SimulationState:
        if self.current_frame_idx >= self.n_frames:
            raise StopIteration
        start = timer()
        self.current_state = self.sph_strategy.compute_next_state(self.current_state)
        self.current_frame_idx += 1
        logger.info("frame %d / %d computed in %.4f s", self.current_frame_idx, self.n_frames, timer() - start)
        return self.current_state

    def __iter__(self) -> "StateGenerator":
        return self
