"""Newmark-beta implicit dynamics (reference: solver/dynamic_implicit.py:15-533).

Present for API compatibility; the consistent-mass Newmark path is not part of
the accelerated hot path yet, so solve() raises.
"""
from .basesolver import BaseSolver


class DynamicImplicitSolver(BaseSolver):
    def __init__(self, maximum_iteration: int = 10000, deltaT: float = 1e-2, time_end: float = 1.0,
                 tol_error: float = 1e-5) -> None:
        super().__init__()
        self.maximum_iteration = maximum_iteration
        self._deltaT = deltaT
        self._time_end = time_end
        self.tol_error = tol_error
        self._gamma = 0.5
        self._beta = 0.25

    def set_deltaT(self, deltaT: float):
        self._deltaT = deltaT

    def solve(self, GC0=None, GV0=None, *args, **kwargs):
        raise NotImplementedError('Newmark-beta dynamics is not on the accelerated path')
