"""Load base class (reference: loads/base.py:16-61)."""
from ..obj_base import BaseObj


class BaseLoad(BaseObj):
    """A load contributes F (and K) in global RGC indices; the assembly subtracts
    them from the internal force (assembly.py:337-341)."""

    def get_stiffness(self, RGC, if_onlyforce: bool = False, *args, **kwargs):
        raise NotImplementedError

    def get_potential_energy(self, RGC):
        raise NotImplementedError("get_potential_energy method not implemented")

    def device_groups(self, assembly):
        """Groups this load contributes to the device plan (facets with global node ids)."""
        return []
