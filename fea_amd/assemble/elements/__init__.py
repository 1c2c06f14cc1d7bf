"""Element registry and factory (reference: elements/__init__.py:1-35)."""
from .base import BaseElement
from .C3 import C3D4, C3D8, C3D10, Element_3D, surfaces
from . import materials
from .C3.surfaces import initialize_surfaces, merge_surfaces, T3, T6, Q4, BaseSurface


def initialize_element(element_type: str, elems_index, elems, part, *args, **kwargs) -> BaseElement:
    """Instantiate an element group by class name (elements/__init__.py:12-35)."""
    element_class = BaseElement._subclasses.get(element_type)
    if element_class is None:
        raise ValueError(f"Element type '{element_type}' is not recognized.")
    return element_class(elems_index=elems_index, elems=elems, part=part)
