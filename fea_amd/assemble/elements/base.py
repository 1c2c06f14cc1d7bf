"""Element plugin registry (reference: elements/base.py:10-148)."""
from __future__ import annotations

import numpy as np
import torch


class BaseElement:
    _subclasses: dict = {}

    def __init_subclass__(cls):
        """Register subclasses by class name for initialize_element (base.py:13-15)."""
        cls._subclasses[cls.__name__] = cls

    def __init__(self, elems_index, elems, part) -> None:
        self._elems_index = elems_index
        self._elems = elems
        self.part = part
        self.materials = None
        self.density = torch.tensor([1.0], dtype=torch.float64)
        self.num_nodes_per_elem: int = 0
        self._num_gaussian: int = 0

    def initialize(self, *args, **kwargs):
        pass

    def set_materials(self, materials):
        self.materials = materials

    def set_density(self, density):
        if isinstance(density, (int, float)):
            density = torch.tensor([float(density)], dtype=torch.float64)
        self.density = torch.as_tensor(density)

    def set_required_DoFs(self, RGC_remain_index: np.ndarray) -> np.ndarray:
        return RGC_remain_index

    def extract_surface(self, surface_ind: int, elems_ind):
        return []

    def refine_RGC(self, RGC, nodes):
        return RGC
