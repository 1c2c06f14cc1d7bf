"""Part / Instance / surface sets (reference: FEA/assemble/part.py)."""
from __future__ import annotations

import numpy as np
import torch

from .elements import BaseElement, surfaces
from .obj_base import BaseObj


class _Surfaces:
    """Named surface sets: name -> [(element indices, local face id)] (part.py:14-93)."""

    def __init__(self):
        self._surface_dict = {}
        self._surface_elements = {}
        self._initialized = False

    def initialize(self, part):
        self._surface_elements.clear()
        for name in self._surface_dict:
            els = part.extract_surfaces(name)
            for se in els:
                se.initialize(part)
            self._surface_elements[name] = els
            self._initialized = True

    def get_elements(self, name):
        return self._surface_elements.get(name, [])

    def __getitem__(self, key):
        return self._surface_dict[key]

    def __setitem__(self, key, value):
        self._surface_dict[key] = value

    def __delitem__(self, key):
        del self._surface_dict[key]

    def __contains__(self, key):
        return key in self._surface_dict

    def keys(self):
        return list(self._surface_dict.keys())


class Part:
    def __init__(self, nodes) -> None:
        self.nodes = nodes
        self.elems: dict[str, BaseElement] = {}
        self.set_nodes: dict[str, np.ndarray] = {}
        self.surfaces = _Surfaces()

    def initialize(self, *args, **kwargs):
        for e in self.elems.values():
            e.initialize()
        self.surfaces.initialize(self)

    def add_element(self, element: BaseElement, name: str = None):
        if name is None:
            number = len(self.elems)
            while ('element-%d' % number) in self.elems:
                number += 1
            name = 'element-%d' % number
        self.elems[name] = element
        return name

    def delete_element(self, name: str):
        if name in self.elems:
            del self.elems[name]
        else:
            raise ValueError(f"Element '{name}' not found in the model.")

    def add_surface_set(self, name: str, elements):
        self.surfaces[name] = elements
        return name

    def delete_surface_set(self, name: str):
        if name in self.surfaces:
            del self.surfaces[name]
        else:
            raise KeyError(f"Surface set '{name}' not found in the model.")

    def extract_surfaces(self, name: str):
        """part.py:291-313: facets of every element group for each (elems, face) entry."""
        out = []
        surf_ind = None
        for elem_ind, surf_ind in self.surfaces[name]:
            for e in self.elems.values():
                out += e.extract_surface(surf_ind, elem_ind)
        if len(out) == 0:
            raise ValueError(f"Surface {surf_ind} not found in the model.")
        return surfaces.merge_surfaces(out)

    def refine_RGC(self, RGC):
        return RGC


class Instance(BaseObj):
    def __init__(self, part: Part) -> None:
        super().__init__()
        self.part = part
        self._RGC_requirements = tuple(torch.as_tensor(part.nodes).shape)
        self._translation = torch.zeros(3, dtype=torch.float64)
        self._rotation = torch.zeros(3, dtype=torch.float64)

    @property
    def rotation_matrix(self):
        if float(torch.norm(torch.as_tensor(self._rotation, dtype=torch.float64))) != 0.0:
            raise NotImplementedError('instance rotations are not supported on the accelerated path')
        return torch.eye(3, dtype=torch.float64)

    @property
    def elems(self):
        return self.part.elems

    @property
    def nodes(self):
        self.rotation_matrix  # validates
        X = torch.as_tensor(self.part.nodes, dtype=torch.float64)
        return X + torch.as_tensor(self._translation, dtype=torch.float64).to(X.device).view(1, 3)

    @property
    def surfaces(self):
        return self.part.surfaces

    @property
    def set_nodes(self):
        return self.part.set_nodes

    def set_required_DoFs(self, RGC_remain_index):
        for e in self.part.elems.values():
            RGC_remain_index[self._RGC_index] = e.set_required_DoFs(RGC_remain_index[self._RGC_index])
        return RGC_remain_index

    def initialize(self, assembly):
        super().initialize(assembly=assembly)

    def refine_RGC(self, RGC):
        return RGC

    def potential_energy(self, RGC):
        return sum(e.potential_Energy(RGC[self._RGC_index]) for e in self.part.elems.values())

    def structural_stiffness(self, RGC, if_onlyforce=False, *args, **kwargs):
        """part.py:424-431: concatenated per-group COO lists, offset into the global RGC."""
        Ri, Rv, Ki, Kv = [], [], [], []
        for e in self.part.elems.values():
            out = e.structural_Force(RGC=RGC[self._RGC_index], if_onlyforce=if_onlyforce)
            Ri.append(out[0])
            Rv.append(out[1])
            if not if_onlyforce:
                Ki.append(out[2])
                Kv.append(out[3])
        if if_onlyforce:
            return torch.cat(Ri) + self._index_start, torch.cat(Rv)
        return torch.cat(Ri) + self._index_start, torch.cat(Rv), torch.cat(Ki, dim=1) + self._index_start, torch.cat(Kv)
