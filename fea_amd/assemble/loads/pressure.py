"""Follower pressure (reference: loads/pressure.py:5-169).

Pi_p = -p V with V = 1/3 sum_g w_g det[x, x_xi, x_eta] over the surface facets;
the force is -p dV/dx and the stiffness -p d2V/dx2 (returned with that sign,
as the reference does, and subtracted by the assembly).
"""
import numpy as np
import torch

from ... import groupops as _nr
from ...device import Group
from .base import BaseLoad


class Pressure(BaseLoad):

    def __init__(self, instance_name: str, surface_set: str, pressure: float) -> None:
        super().__init__()
        self.surface_name = surface_set
        self.instance_name = instance_name
        self.pressure = float(pressure)
        self.surface_element = []
        self._load_index = None

    def initialize(self, assembly):
        super().initialize(assembly)
        inst = assembly.get_instance(self.instance_name)
        self._load_index = inst._RGC_index
        self.surface_element = [s for s in inst.surfaces.get_elements(self.surface_name)
                                if s._elems.shape[0] > 0]

    def _offset(self):
        return int(self._assembly.RGC_list_indexStart[self._load_index]) // 3

    def device_groups(self, assembly):
        off = self._offset()
        return [Group('pressure', s.num_nodes_per_elem, s._elems.cpu().numpy() + off, load=self,
                      name=self.surface_name) for s in self.surface_element]

    @property
    def _Vdot_indices(self):
        off = self._index_start_nodes()
        return torch.cat([(torch.as_tensor(s._elems).cuda().unsqueeze(1) * 3 +
                           torch.arange(3, device='cuda').view(1, 3, 1)).reshape(-1)
                          for s in self.surface_element]) + off

    @property
    def _Vdot_2_indices(self):
        out = []
        for s in self.surface_element:
            e = torch.as_tensor(s._elems).cuda()
            E, n = e.shape
            A = e.view(E, 1, n, 1, 1)
            B = e.view(E, 1, 1, 1, n)
            i = torch.arange(3, device='cuda').view(1, 3, 1, 1, 1)
            j = torch.arange(3, device='cuda').view(1, 1, 1, 3, 1)
            rows = (A * 3 + i).expand(E, 3, n, 3, n).reshape(-1)
            cols = (B * 3 + j).expand(E, 3, n, 3, n).reshape(-1)
            out.append(torch.stack([rows, cols]))
        return torch.cat(out, dim=1) + self._index_start_nodes()

    def _index_start_nodes(self):
        return int(self._assembly.RGC_list_indexStart[self._load_index])

    def _pos(self, RGC):
        inst = self._assembly.get_instance(self.instance_name)
        X = torch.as_tensor(inst.nodes).to('cuda', torch.float64).contiguous().reshape(-1)
        U = _nr.as_u(RGC[self._load_index])
        return X, U

    def get_stiffness(self, RGC, if_onlyforce=False, *args, **kwargs):
        """pressure.py:80-132: (F idx, -p dV, K idx, -p d2V)."""
        X, U = self._pos(RGC)
        Fs, Ks = [], []
        for s in self.surface_element:
            Fv, Kv = _nr.surface_pressure_per_facet(s.num_nodes_per_elem, s._elems, X, U, self.pressure,
                                                    if_onlyforce)
            Fs.append(Fv)
            Ks.append(Kv)
        if if_onlyforce:
            return self._Vdot_indices, torch.cat(Fs)
        return self._Vdot_indices, torch.cat(Fs), self._Vdot_2_indices, torch.cat(Ks)

    def get_potential_energy(self, RGC):
        """pressure.py:134-159: -p V."""
        X, U = self._pos(RGC)
        V = sum(_nr.surface_volume(s.num_nodes_per_elem, s._elems, X, U) for s in self.surface_element)
        return -self.pressure * V

    def set_required_DoFs(self, RGC_remain_index):
        """pressure.py:161-169: facet nodes are required DOFs."""
        for s in self.surface_element:
            RGC_remain_index[self._load_index][np.unique(s._elems.cpu().numpy())] = True
        return RGC_remain_index
