"""Boundary base class (reference: boundarys/base.py:6-39)."""
from ..obj_base import BaseObj


class BaseBoundary(BaseObj):
    def set_required_DoFs(self, RGC_remain_index):
        return RGC_remain_index

    def modify_RGC(self, RGC):
        return RGC
