"""Reference point: 6 generalized coordinates (translation v, rotation vector z)
(reference: assemble/reference_points.py:5-24)."""
import numpy as np
import torch

from .obj_base import BaseObj


class ReferencePoint(BaseObj):

    def __init__(self, node=None) -> None:
        super().__init__()
        if isinstance(node, (list, tuple, np.ndarray)):
            node = torch.tensor(np.asarray(node, dtype=np.float64).tolist(), dtype=torch.float64)
        self.node = node
        self._RGC_requirements = 6
