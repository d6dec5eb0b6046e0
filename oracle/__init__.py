"""CPU FP64 oracle of the reference SPH step -- test infrastructure only.

Used by tests/, __graft_entry__.smoke() (as the checker) and bench.py's cpu_baseline leg.
The product package never imports it.
"""
from .oracle import *  # noqa: F401,F403
