"""RGC bookkeeping base class (obj_base.py:11-59 of the reference)."""
from __future__ import annotations


class BaseObj:
    """Object owning a block of redundant generalized coordinates (RGC)."""

    def __init__(self) -> None:
        self._RGC_requirements = [0]
        self._RGC_index = None
        self._index_start = None
        self._assembly = None

    def set_RGC_index(self, index: int) -> None:
        self._RGC_index = index

    def set_required_DoFs(self, RGC_remain_index):
        return RGC_remain_index

    def modify_RGC(self, RGC):
        return RGC

    def initialize(self, assembly):
        self._assembly = assembly
        self._index_start = assembly.RGC_list_indexStart[self._RGC_index]

    def initialize_dynamic(self):
        pass

    def reinitialize(self, RGC):
        pass
