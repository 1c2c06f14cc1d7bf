"""Zero Dirichlet condition on reference-point DOFs (UX, UY, UZ, RX, RY, RZ)
(reference: boundarys/boundary_condition_rp.py:5-31)."""
from .base import BaseBoundary


class Boundary_Condition_RP(BaseBoundary):

    def __init__(self, rp_name: str, indexDoF=(0, 1, 2, 3, 4, 5)) -> None:
        super().__init__()
        self.rp_name = rp_name
        self.indexDoF = list(indexDoF)
        self._constraint_index = None

    def initialize(self, assembly):
        super().initialize(assembly)
        self._constraint_index = assembly.get_reference_point(self.rp_name)._RGC_index

    def modify_RGC(self, RGC):
        for i in self.indexDoF:
            RGC[self._constraint_index][i] = 0.0
        return RGC

    def set_required_DoFs(self, RGC_remain_index):
        for i in self.indexDoF:
            RGC_remain_index[self._constraint_index][i] = False
        return RGC_remain_index
