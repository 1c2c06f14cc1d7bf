from .part import Part, Instance
from .assembly import Assembly
from . import elements, loads, boundarys
from .elements import materials
from .elements.C3 import surfaces
