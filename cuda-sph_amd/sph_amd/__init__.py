"""MI355X-native SPH step with the API of iwoplaza/cuda-sph's simulator.

Importing the package does not touch the GPU; the strategies open libsph_amd.so and the
device when they are constructed.
"""
from .constants import MAX_NEIGHBOURS, RNG_SEED, SPHConstants
from .data_classes import Pipe, Segment, SimulationParameters, SimulationState

__all__ = [
    "SPHConstants", "MAX_NEIGHBOURS", "RNG_SEED", "Pipe", "Segment", "SimulationParameters", "SimulationState",
    "SPHStrategy", "NaiveSPHStrategy", "VoxelSPHStrategy", "make_strategy", "StateGenerator", "BatchedSPH",
    "DifferentiableSPH", "rollout",
]


def __getattr__(name):
    if name in ("SPHStrategy", "NaiveSPHStrategy", "VoxelSPHStrategy", "make_strategy"):
        from . import strategy
        return getattr(strategy, name)
    if name == "StateGenerator":
        from .state_generator import StateGenerator
        return StateGenerator
    if name == "BatchedSPH":
        from .batched import BatchedSPH
        return BatchedSPH
    if name in ("DifferentiableSPH", "rollout"):
        from . import differentiable
        return getattr(differentiable, name)
    raise AttributeError(name)
