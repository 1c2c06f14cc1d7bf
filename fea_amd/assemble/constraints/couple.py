"""Kinematic coupling of an instance node set to a reference point
(reference: constraints/couple.py:7-292).

The node set follows the point rigidly: u_b = v + R(z) y_b - y_b, y_b = X_b - X_rp
(couple.py:36-43).  The coupled node DOFs leave the generalized coordinates and the
point's 6 DOFs enter them (couple.py:57-64).  The residual / tangent condensation and the
kinematics run in the HIP kernels of fea_amd/csrc/couple.hip through
``fea_amd.couple.CoupledPlan``; this object only holds the definition.
"""
import numpy as np

from .base import BaseConstraint


class Couple(BaseConstraint):

    def __init__(self, instance_name: str, set_nodes_name: str, rp_name: str) -> None:
        super().__init__()
        self.instance_name = instance_name
        self.set_nodes_name = set_nodes_name
        self.rp_name = rp_name
        self._indexNodes = None
        self._rp_index = None
        self._couple_index = None
        self._ref_location = None

    def initialize(self, assembly):
        """couple.py:22-34."""
        super().initialize(assembly)
        inst = assembly.get_instance(self.instance_name)
        self._indexNodes = np.asarray(inst.set_nodes[self.set_nodes_name], dtype=np.int64)
        rp = assembly.get_reference_point(self.rp_name)
        self._rp_index = rp._RGC_index
        self._couple_index = inst._RGC_index
        X = np.asarray(inst.nodes.detach().cpu().numpy(), dtype=np.float64)
        self._ref_location = X[self._indexNodes] - np.asarray(rp.node.detach().cpu().numpy(), dtype=np.float64)

    def set_required_DoFs(self, RGC_remain_index):
        """couple.py:57-64: the coupled node DOFs leave, the point's 6 DOFs enter."""
        RGC_remain_index[self._couple_index][self._indexNodes] = False
        RGC_remain_index[self._rp_index][:] = True
        return RGC_remain_index
