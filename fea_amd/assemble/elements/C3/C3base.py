"""Element_3D: the drop-in for the reference's solid element base class
(FEA/assemble/elements/C3/C3base.py:16-697), backed by the HIP element kernels.

The hot path (Assembly + solvers) never goes through the per-group methods
below; they exist so that code written against the reference's element API
(structural_Force, potential_Energy, the Gauss-point attributes and the
group COO pattern) keeps working, each computed by the same kernels.
"""
from __future__ import annotations

import numpy as np
import torch

from .... import groupops as _nr
from ..base import BaseElement
from .surfaces import initialize_surfaces


class Element_3D(BaseElement):
    etype = 0                 # nodes per element (= C-ABI element type id)
    ngp = 0
    face_nodes: dict = {}     # surface index -> local node order of the facet

    def __init__(self, elems=None, elems_index=None, part=None):
        elems = torch.as_tensor(elems, dtype=torch.int64)
        if elems_index is None:
            elems_index = torch.arange(elems.shape[0], dtype=torch.int64)
        super().__init__(torch.as_tensor(elems_index), elems, part)
        self.num_nodes_per_elem = self.etype
        self._num_gaussian = self.ngp
        self.num_surfaces = len(self.face_nodes)
        self._cache = {}

    # ------------------------------------------------------------------ setup
    def initialize(self, *args, **kwargs) -> None:
        if self._elems.dim() != 2 or self._elems.shape[1] != self.etype:
            raise ValueError('%s expects connectivity [E, %d], got %s'
                             % (type(self).__name__, self.etype, tuple(self._elems.shape)))
        self._cache = {}

    def conn_np(self):
        return self._elems.detach().cpu().numpy().astype(np.int64)

    def kernel_params(self):
        if self.materials is None:
            raise ValueError('element group has no material (set_materials)')
        return self.materials.kernel_params()

    def set_required_DoFs(self, RGC_remain_index: np.ndarray) -> np.ndarray:
        """C3base.py:499-504 (full-order elements: every referenced node is required)."""
        RGC_remain_index[np.unique(self.conn_np())] = True
        return RGC_remain_index

    def extract_surface(self, surface_ind: int, elems_ind):
        """Facets of local face `surface_ind` of the elements listed in elems_ind
        (tetrahedral.py:45-64 / :122-156, brick.py:166-201)."""
        if surface_ind not in self.face_nodes:
            raise ValueError(f"Invalid surface index: {surface_ind}")
        index_now = np.where(np.isin(np.asarray(self._elems_index.cpu()), np.asarray(elems_ind)))[0]
        local = self.face_nodes[surface_ind]
        facets = self._elems[torch.as_tensor(index_now, dtype=torch.int64)][:, local]
        if facets.shape[0] == 0 and self.empty_surface_is_dropped:
            return []
        return [initialize_surfaces(facets)]

    empty_surface_is_dropped = False

    # ------------------------------------------------------------------ lazy Gauss data
    def _nodes(self):
        return torch.as_tensor(self.part.nodes).to('cuda', torch.float64).contiguous()

    def _prepared(self):
        if 'prep' not in self._cache:
            self._cache['prep'] = _nr.elem_prepare(self.etype, self.ngp, self._elems, self._nodes())
        return self._cache['prep']

    @property
    def shape_function_d1_gaussian(self):
        """dN/dX [g, E, 3, nen] (C3base.py:22-31, :208)."""
        return self._prepared()[0]

    @property
    def gaussian_weight(self):
        """Signed w*detJ [g, E] (C3base.py:214)."""
        return self._prepared()[1]

    @property
    def shape_function_d0_gaussian(self):
        """N [g, E, nen] (C3base.py:33-34, :211)."""
        return self._prepared()[2]

    def _group_plan(self):
        if 'plan' not in self._cache:
            self._cache['plan'] = _nr.group_plan(self)
        return self._cache['plan']

    @property
    def _indices_force(self):
        """DOF of each R_e entry in [a, j, e] order (C3base.py:148-153)."""
        if 'ifor' not in self._cache:
            c = self._elems.to('cuda')
            idx = c.t().unsqueeze(1) * 3 + torch.arange(3, device='cuda').view(1, 3, 1)
            self._cache['ifor'] = idx.reshape(-1)
        return self._cache['ifor']

    @property
    def _indices_matrix(self):
        """Coalesced group pattern [2, nnz_grp] in sorted (row, col) order (C3base.py:142-145)."""
        return self._group_plan().csr_indices_coo()

    @property
    def _index_matrix_coalesce(self):
        """Position of each [a, j, b, k, e] entry in the group pattern (C3base.py:129-141)."""
        if 'coal' not in self._cache:
            self._cache['coal'] = _nr.coalesce_map(self, self._group_plan())
        return self._cache['coal']

    # ------------------------------------------------------------------ reference API
    def structural_Force(self, RGC: torch.Tensor, rotation_matrix=None, if_onlyforce: bool = False,
                         *args, **kwargs):
        """C3base.py:413-452: (_indices_force, R_e [a,j,e] flat, _indices_matrix, K values)."""
        _nr.require_identity(rotation_matrix)
        U = _nr.as_u(RGC)
        Re = _nr.elem_residual_per_node(self, U)
        if if_onlyforce:
            return self._indices_force, Re
        plan = self._group_plan()
        _, val = plan.tangent(U, 1.0, _nr.F64, with_residual=False)
        return self._indices_force, Re, self._indices_matrix, val.clone()

    def potential_Energy(self, RGC: torch.Tensor):
        """C3base.py:387-411: sum_g W w detJ over the group (float64 0-dim tensor)."""
        return _nr.elem_energy(self, _nr.as_u(RGC))

    def get_volumn(self, U=None):
        """C3base.py:482-497."""
        if U is None:
            return self.gaussian_weight.sum()
        raise NotImplementedError('deformed element volume is not part of the accelerated path')

    def get_mass_matrix(self, rotation_matrix=None):
        raise NotImplementedError('consistent mass matrices are not part of the accelerated path; '
                                  'the explicit solver uses the lumped-mass kernel')
