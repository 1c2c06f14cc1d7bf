"""Solver base class (reference: solver/basesolver.py:10-50)."""


class BaseSolver:
    def __init__(self) -> None:
        self.assembly = None
        self.GC = None

    def initialize(self, assembly, *args, **kwargs):
        self.assembly = assembly

    def solve(self, GC0=None, *args, **kwargs):
        pass
