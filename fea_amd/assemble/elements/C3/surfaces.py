"""Surface facets T3 / T6 / Q4 (reference: elements/C3/surfaces).

A facet object only carries connectivity and the facet's type; the Gauss
rules live in the HIP surface kernels (triangle.py:22-46 T3, :57-93 T6,
quadrilateral.py:13-57 Q4).
"""
import numpy as np
import torch


class BaseSurface:
    _subclasses = {}
    num_nodes_per_elem = 0
    _num_gaussian = 0

    def __init_subclass__(cls):
        cls._subclasses[cls.__name__] = cls

    def __init__(self, elems) -> None:
        self._elems = torch.as_tensor(elems, dtype=torch.int64).reshape(-1, self.num_nodes_per_elem)
        self._part = None

    def initialize(self, part) -> None:
        self._part = part

    @property
    def surf_elems_circ(self):
        return self._elems


class T3(BaseSurface):
    num_nodes_per_elem = 3
    _num_gaussian = 3


class T6(BaseSurface):
    num_nodes_per_elem = 6
    _num_gaussian = 3

    @property
    def surf_elems_circ(self):
        return self._elems[:, [0, 3, 1, 4, 2, 5]]


class Q4(BaseSurface):
    num_nodes_per_elem = 4
    _num_gaussian = 4


def initialize_surfaces(surface_elems) -> BaseSurface:
    """Facet factory by node count (surfaces/__init__.py:7-32)."""
    surface_elems = torch.as_tensor(surface_elems)
    n = surface_elems.shape[1]
    if n == 3:
        return T3(surface_elems)
    if n == 6:
        return T6(surface_elems)
    if n == 4:
        return Q4(surface_elems)
    raise ValueError(f"Unsupported surface element type: {n}.")


def merge_surfaces(surfaces):
    """Merge facets of equal type (surfaces/__init__.py:34-85)."""
    if len(surfaces) <= 1:
        return surfaces
    by = {'T3': [], 'T6': [], 'Q4': []}
    for s in surfaces:
        by[s.__class__.__name__].append(s._elems)
    out = []
    for name, cls in (('T3', T3), ('T6', T6), ('Q4', Q4)):
        if by[name]:
            e = torch.cat(by[name], dim=0)
            if e.numel() > 0:
                out.append(cls(e))
    return out
