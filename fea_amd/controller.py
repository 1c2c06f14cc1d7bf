"""FEAController (reference: FEA/controller.py:13-56)."""


class FEAController:

    def __init__(self, maximum_iteration: int = 10000) -> None:
        self.assembly = None
        self.solver = None

    def initialize(self, *args, **kwargs):
        """controller.py:29-40."""
        self.assembly.initialize(*args, **kwargs)
        self.solver.initialize(assembly=self.assembly)

    def solve(self, GC0=None, if_initialize: bool = True, *args, **kwargs):
        """controller.py:42-56."""
        if if_initialize:
            self.initialize()
        return self.solver.solve(GC0=GC0, *args, **kwargs)
