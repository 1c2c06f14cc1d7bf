from .base import BaseBoundary
from .boundary_condition import Boundary_Condition
from .boundary_condition_rp import Boundary_Condition_RP

__all__ = ['BaseBoundary', 'Boundary_Condition', 'Boundary_Condition_RP']
