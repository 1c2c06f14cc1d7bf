from .base import BaseBoundary
from .boundary_condition import Boundary_Condition

__all__ = ['BaseBoundary', 'Boundary_Condition']
