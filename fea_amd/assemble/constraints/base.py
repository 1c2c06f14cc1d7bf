"""Constraint base class (reference: constraints/base.py:5-31)."""
from ..obj_base import BaseObj


class BaseConstraint(BaseObj):
    def set_required_DoFs(self, RGC_remain_index):
        return RGC_remain_index
