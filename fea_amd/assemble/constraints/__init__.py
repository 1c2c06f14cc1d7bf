from .base import BaseConstraint
from .couple import Couple

__all__ = ['BaseConstraint', 'Couple']
