from .C3base import Element_3D
from .tetrahedral import C3D4, C3D10
from .brick import C3D8
from . import surfaces
from .surfaces import initialize_surfaces, merge_surfaces, BaseSurface, T3, T6, Q4
