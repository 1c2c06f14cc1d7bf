"""Dead-load uniform traction (new; the reference has no counterpart).

f_a = integral over the undeformed facets of N_a t dA0, with the reference-area
Gauss rule of the reference's surfaces (basesurface.py:144-150).  It has no
stiffness; its potential is f . u.
"""
import numpy as np
import torch

from ... import groupops as _nr
from ...device import Group
from .base import BaseLoad


class Traction(BaseLoad):

    def __init__(self, instance_name: str, surface_set: str, traction) -> None:
        super().__init__()
        self.instance_name = instance_name
        self.surface_name = surface_set
        self.traction = np.asarray(traction, dtype=np.float64).reshape(3)
        self.surface_element = []
        self._load_index = None

    def initialize(self, assembly):
        super().initialize(assembly)
        inst = assembly.get_instance(self.instance_name)
        self._load_index = inst._RGC_index
        self.surface_element = [s for s in inst.surfaces.get_elements(self.surface_name)
                                if s._elems.shape[0] > 0]

    def device_groups(self, assembly):
        off = int(assembly.RGC_list_indexStart[self._load_index]) // 3
        return [Group('traction', s.num_nodes_per_elem, s._elems.cpu().numpy() + off, load=self,
                      in_k=False, name=self.surface_name) for s in self.surface_element]

    def _X(self):
        inst = self._assembly.get_instance(self.instance_name)
        return torch.as_tensor(inst.nodes).to('cuda', torch.float64).contiguous().reshape(-1)

    def get_stiffness(self, RGC, if_onlyforce=False, *args, **kwargs):
        X = self._X()
        off = int(self._assembly.RGC_list_indexStart[self._load_index])
        idx, val = [], []
        for s in self.surface_element:
            e = torch.as_tensor(s._elems).cuda()
            idx.append((e.unsqueeze(1) * 3 + torch.arange(3, device='cuda').view(1, 3, 1)).reshape(-1) + off)
            val.append(_nr.surface_traction_force(s.num_nodes_per_elem, s._elems, X, self.traction))
        F_idx, F_val = torch.cat(idx), torch.cat(val)
        if if_onlyforce:
            return F_idx, F_val
        empty_i = torch.zeros([2, 0], dtype=torch.int64, device='cuda')
        return F_idx, F_val, empty_i, torch.zeros([0], dtype=torch.float64, device='cuda')

    def get_potential_energy(self, RGC):
        F_idx, F_val = self.get_stiffness(RGC, if_onlyforce=True)
        u = torch.cat([_nr.as_u(r) for r in RGC]) if isinstance(RGC, (list, tuple)) else _nr.as_u(RGC)
        return (F_val * u[F_idx]).sum()

    def set_required_DoFs(self, RGC_remain_index):
        return RGC_remain_index
