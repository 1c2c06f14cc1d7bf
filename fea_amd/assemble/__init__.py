from .part import Part, Instance
from .assembly import Assembly
from .reference_points import ReferencePoint
from . import elements, loads, boundarys, constraints
from .elements import materials
from .elements.C3 import surfaces
