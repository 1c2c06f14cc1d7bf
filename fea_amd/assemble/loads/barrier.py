"""The reference's contact loads: ContactSelf (one surface) and Contact (two surfaces),
a Gauss-point-pair barrier between T3 facets (reference: loads/contact.py:10-1218).

These objects hold the definition; the pair search, energy, gradient and Hessian run in
the HIP kernels of fea_amd/csrc/barrier.hip through ``fea_amd.barrier.BarrierSet``, and the
assembly's ``LivePlan`` keeps the CSR sparsity covering the contact pairs.  Parameters and
defaults are the reference's (ContactBase.__init__, contact.py:11-50; ContactSelf,
contact.py:250-285; the candidate radius is 2.0 x penalty_threshold_h in both).
"""
import numpy as np

from .base import BaseLoad


class _BarrierLoad(BaseLoad):

    def __init__(self, penalty_distance_g: float = 1e-5, penalty_factor_g: float = 20.0, penalty_degree: int = 9,
                 penalty_threshold_h: float = 1.5, penalty_ratio_h: float = 0.5, penalty_start_f: float = -0.0,
                 penalty_end_f: float = -0.8, search_ratio: float = 2.0) -> None:
        super().__init__()
        self._penalty_distance_g = float(penalty_distance_g)
        self._penalty_factor_g = float(penalty_factor_g)
        self._penalty_degree = int(penalty_degree)          # the reference's polynomial law is unused
        self._penalty_threshold_h = float(penalty_threshold_h)
        self._penalty_ratio_h = float(penalty_ratio_h)
        self._penalty_start_f = float(penalty_start_f)
        self._penalty_end_f = float(penalty_end_f)
        self._search_ratio = float(search_ratio)
        self._ignore_min_normal, self._ignore_max_normal = 0.8, 1.2

    def params(self):
        return {'c_g': self._penalty_distance_g, 'k_g': self._penalty_factor_g, 'h_t': self._penalty_threshold_h,
                'r_h': self._penalty_ratio_h, 's_f': self._penalty_start_f, 'e_f': self._penalty_end_f,
                'ign_min': self._ignore_min_normal, 'ign_max': self._ignore_max_normal,
                'search_ratio': self._search_ratio}

    @staticmethod
    def _facets(assembly, instance_name, surface_name):
        """Global-node T3 facets of the first surface group of a set (contact.py:296, :756-757)."""
        inst = assembly.get_instance(instance_name)
        groups = inst.surfaces.get_elements(surface_name)
        if not groups:
            raise ValueError('surface %s not found on instance %s' % (surface_name, instance_name))
        s = groups[0]
        if s.num_nodes_per_elem != 3:
            raise NotImplementedError('the contact barrier is built for T3 facets')
        off = int(assembly.RGC_list_indexStart[inst._RGC_index]) // 3
        return s._elems.cpu().numpy().astype(np.int64) + off

    def device_groups(self, assembly):
        return []


class ContactSelf(_BarrierLoad):

    def __init__(self, instance_name: str, surface_name: str, ignore_min_normal: float = 0.8,
                 ignore_max_normal: float = 1.2, initial_detact_ratio: float = 2., **kwargs) -> None:
        super().__init__(search_ratio=initial_detact_ratio, **kwargs)
        self.instance_name = self.instance_name1 = self.instance_name2 = instance_name
        self.surface_name = self.surface_name1 = self.surface_name2 = surface_name
        self._ignore_min_normal = float(ignore_min_normal)
        self._ignore_max_normal = float(ignore_max_normal)
        self.is_self_contact = True

    def facets(self, assembly):
        return self._facets(assembly, self.instance_name, self.surface_name), None

    def set_required_DoFs(self, RGC_remain_index):
        """contact.py:707-712: the surface's nodes are generalized coordinates."""
        inst = self._assembly.get_instance(self.instance_name)
        nodes = np.unique(inst.surfaces.get_elements(self.surface_name)[0]._elems.cpu().numpy())
        RGC_remain_index[inst._RGC_index][nodes] = True
        return RGC_remain_index


class Contact(_BarrierLoad):

    def __init__(self, instance_name1: str, instance_name2: str, surface_name1: str, surface_name2: str,
                 **kwargs) -> None:
        super().__init__(**kwargs)
        self.instance_name1, self.instance_name2 = instance_name1, instance_name2
        self.surface_name1, self.surface_name2 = surface_name1, surface_name2
        self.is_self_contact = False

    def facets(self, assembly):
        return (self._facets(assembly, self.instance_name1, self.surface_name1),
                self._facets(assembly, self.instance_name2, self.surface_name2))
