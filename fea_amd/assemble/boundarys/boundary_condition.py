"""Zero Dirichlet condition on a node set (reference: boundary_condition.py:5-41)."""
import numpy as np

from .base import BaseBoundary


class Boundary_Condition(BaseBoundary):

    def __init__(self, instance_name: str, set_nodes_name: str, indexDoF=(0, 1, 2)) -> None:
        super().__init__()
        self.set_nodes_name = set_nodes_name
        self.instance_name = instance_name
        self.indexDoF = list(indexDoF)
        self.index_nodes = None
        self._constraint_index = None

    def initialize(self, assembly):
        super().initialize(assembly)
        inst = self._assembly.get_instance(self.instance_name)
        self.index_nodes = np.asarray(inst.set_nodes[self.set_nodes_name], dtype=np.int64)
        self._constraint_index = inst._RGC_index

    def modify_RGC(self, RGC):
        """boundary_condition.py:27-33: zero the constrained components."""
        for i in self.indexDoF:
            RGC[self._constraint_index][self.index_nodes, i] = 0.0
        return RGC

    def set_required_DoFs(self, RGC_remain_index):
        """boundary_condition.py:35-41: constrained DOFs leave the GC."""
        for i in self.indexDoF:
            RGC_remain_index[self._constraint_index][self.index_nodes, i] = False
        return RGC_remain_index
