"""Force and moment on a reference point (reference: loads/concentrate_force.py:5-49,
loads/moment.py:5-48): constant generalized forces on the point's translation v and
rotation vector z, with potentials F.v and M.z.  On the device path they enter the
reduced residual of ``fea_couple_condense`` (R -= load_scale * (F, M))."""
import numpy as np
import torch

from .base import BaseLoad


class _PointLoad(BaseLoad):
    _SLICE = slice(0, 3)

    def __init__(self, rp_name: str, value) -> None:
        super().__init__()
        self.rp_name = rp_name
        self.rp_index = None
        self._value = np.asarray(value, dtype=np.float64).reshape(3)
        self._indices_force = None

    def initialize(self, assembly):
        super().initialize(assembly)
        self.rp_index = assembly.get_reference_point(self.rp_name)._RGC_index
        s = assembly.RGC_list_indexStart[self.rp_index] + self._SLICE.start
        self._indices_force = torch.arange(s, s + 3)

    def generalized_load(self):
        """(6,) force / moment vector of this load on its point."""
        out = np.zeros(6)
        out[self._SLICE] = self._value
        return out

    def get_stiffness(self, RGC, if_onlyforce=False, *args, **kwargs):
        f = torch.as_tensor(self._value)
        if if_onlyforce:
            return self._indices_force, f
        return self._indices_force, f, torch.zeros([2, 0], dtype=torch.int64), torch.zeros([0])

    def get_potential_energy(self, RGC):
        return (torch.as_tensor(self._value, dtype=torch.float64) *
                torch.as_tensor(RGC[self.rp_index][self._SLICE]).cpu().to(torch.float64)).sum()

    def set_required_DoFs(self, RGC_remain_index):
        RGC_remain_index[self.rp_index][self._SLICE] = True
        return RGC_remain_index


class Concentrate_Force(_PointLoad):
    _SLICE = slice(0, 3)

    def __init__(self, rp_name: str, force) -> None:
        super().__init__(rp_name, force)

    @property
    def force(self):
        return self._value

    @force.setter
    def force(self, value):
        self._value = np.asarray(value, dtype=np.float64).reshape(3)


class Moment(_PointLoad):
    _SLICE = slice(3, 6)

    def __init__(self, rp_name: str, moment) -> None:
        super().__init__(rp_name, moment)

    @property
    def moment(self):
        return self._value

    @moment.setter
    def moment(self, value):
        self._value = np.asarray(value, dtype=np.float64).reshape(3)
