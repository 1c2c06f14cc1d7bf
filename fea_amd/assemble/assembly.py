"""Assembly: registries, DOF numbering and the global residual / tangent
(reference: FEA/assemble/assembly.py:10-767).

``initialize`` reproduces the reference's RGC/GC bookkeeping (name-sorted
registries, RGC blocks, required-DOF masks, assembly.py:76-185) and then builds
one ``DevicePlan`` for the whole model: every element group of every instance
and every surface load becomes a kernel group writing into one precomputed
free-DOF CSR.  A model with reference points and Couple constraints gets a
``CoupledPlan`` (fea_amd/couple.py) with the same interface over the state
[nodal displacements | 6 per reference point].  ``assemble_Stiffness_Matrix`` / ``assemble_force`` /
``_total_Potential_Energy`` run entirely in HIP kernels.
"""
from __future__ import annotations

import numpy as np
import torch

from ..device import DevicePlan, Group
from .._native import F64
from . import boundarys, constraints, loads
from .part import Instance, Part
from .reference_points import ReferencePoint


class Assembly:
    def __init__(self):
        self.device = torch.device('cuda')
        self._parts: dict[str, Part] = {}
        self._instances: dict[str, Instance] = {}
        self._reference_points = {}
        self._loads: dict = {}
        self._constraints: dict = {}
        self._boundarys: dict = {}
        self.RGC = []
        self._RGC_size = []
        self.RGC_remain_index = []
        self.RGC_list_indexStart = []
        self._GC_list_indexStart = []
        self.GC = None
        self.plan: DevicePlan | None = None

    # ------------------------------------------------------------------ initialization
    def _allocate_RGC(self, size):
        """assembly.py:448-468."""
        index_now = len(self.RGC)
        size = (int(size),) if np.ndim(size) == 0 else tuple(int(s) for s in size)
        self.RGC.append(torch.zeros(size, dtype=torch.float64, device=self.device))
        self.RGC_remain_index.append(np.zeros(size, dtype=bool))
        self._RGC_size.append(size)
        self.RGC_list_indexStart.append(self.RGC_list_indexStart[-1] + int(np.prod(size)))
        return index_now

    def initialize(self, *args, **kwargs):
        """assembly.py:76-185 + construction of the device plan."""
        self._parts = dict(sorted(self._parts.items()))
        self._instances = dict(sorted(self._instances.items()))
        self._loads = dict(sorted(self._loads.items()))
        self._constraints = dict(sorted(self._constraints.items()))
        self._reference_points = dict(sorted(self._reference_points.items()))
        self._boundarys = dict(sorted(self._boundarys.items()))
        self.RGC, self.RGC_remain_index, self._RGC_size = [], [], []
        self.RGC_list_indexStart = [0]
        RGC_index = None
        for ins in self._instances.values():
            RGC_index = self._allocate_RGC(ins._RGC_requirements)
            ins.set_RGC_index(RGC_index)
        for rp in self._reference_points.values():
            RGC_index = self._allocate_RGC(rp._RGC_requirements)
            rp.set_RGC_index(RGC_index)
        if RGC_index is not None:
            self.RGC[RGC_index][-1] = 1e-5            # initial guess quirk, assembly.py:113
        for f in self._loads.values():
            f.set_RGC_index(self._allocate_RGC(f._RGC_requirements))
        for c in self._constraints.values():
            c.set_RGC_index(self._allocate_RGC(c._RGC_requirements))
        for b in self._boundarys.values():
            b.set_RGC_index(self._allocate_RGC(b._RGC_requirements))

        for part in self._parts.values():
            part.initialize(self)
        for ins in self._instances.values():
            ins.initialize(self)
        for l in self._loads.values():
            l.initialize(self)
        for c in self._constraints.values():
            c.initialize(self)
        for b in self._boundarys.values():
            b.initialize(self)

        for ins in self._instances.values():
            self.RGC_remain_index = ins.set_required_DoFs(self.RGC_remain_index)
        for f in self._loads.values():
            self.RGC_remain_index = f.set_required_DoFs(self.RGC_remain_index)
        for c in self._constraints.values():
            self.RGC_remain_index = c.set_required_DoFs(self.RGC_remain_index)
        for b in self._boundarys.values():
            self.RGC_remain_index = b.set_required_DoFs(self.RGC_remain_index)
        self.RGC_remain_index_flatten = torch.tensor(
            np.concatenate([m.reshape(-1) for m in self.RGC_remain_index]), dtype=torch.bool)

        self._build_plan()
        self.GC = self._RGC2GC(self.RGC)
        self._GC_list_indexStart = np.cumsum([0] + [int(m.sum()) for m in self.RGC_remain_index]).tolist()

    def _build_plan(self):
        X, masks, groups, offsets = [], [], [], {}
        node_off = 0
        for name, ins in self._instances.items():
            if self.RGC_list_indexStart[ins._RGC_index] != 3 * node_off:
                raise RuntimeError('instance RGC blocks must be contiguous')
            Xi = ins.nodes.detach().cpu().numpy().astype(np.float64)
            X.append(Xi)
            masks.append(self.RGC_remain_index[ins._RGC_index].reshape(-1))
            offsets[name] = node_off
            for ename, e in ins.part.elems.items():
                groups.append(Group('elem', e.etype, e.conn_np() + node_off, params=e.kernel_params(),
                                    rho=float(torch.as_tensor(e.density).reshape(-1)[0]),
                                    name='%s/%s' % (name, ename)))
            node_off += Xi.shape[0]
        self._n_nodes = node_off
        for f in self._loads.values():
            groups += f.device_groups(self)
        X = np.concatenate(X, axis=0) if X else np.zeros((0, 3))
        free = np.concatenate(masks) if masks else np.zeros(0, dtype=bool)
        self._coupled = bool(self._reference_points or self._constraints)
        barrier_loads = [l for l in self._loads.values() if isinstance(l, loads.barrier._BarrierLoad)]
        if barrier_loads:
            from ..barrier import BarrierSet, LivePlan
            base = self._base_plan_builder(X, free, groups, offsets)
            Xd = torch.as_tensor(X, dtype=torch.float64, device=self.device).reshape(-1).contiguous()
            sets = []
            for name, l in self._loads.items():
                if l in barrier_loads:
                    c1, c2 = l.facets(self)
                    sets.append(BarrierSet(c1, c2, Xd, l.params(), self.device, name=name))
            self.plan = LivePlan(lambda extra: base(extra), sets, node_off)
            self._K_indices = None
            return
        self.plan = self._base_plan_builder(X, free, groups, offsets)([])
        self._K_indices = None

    def _base_plan_builder(self, X, free, groups, offsets):
        """A function of extra groups building the DevicePlan (or CoupledPlan) of the model."""
        if self._coupled:
            from ..couple import CoupledPlan
            rp_names = list(self._reference_points.keys())
            if not rp_names:
                raise ValueError('a Couple needs a reference point')
            if 3 * X.shape[0] != self.RGC_list_indexStart[self._reference_points[rp_names[0]]._RGC_index]:
                raise RuntimeError('reference point RGC blocks must follow the instances')
            rps = []
            for rn in rp_names:
                rp = self._reference_points[rn]
                load = np.zeros(6)
                for l in self._loads.values():
                    if isinstance(l, (loads.Concentrate_Force, loads.Moment)) and l.rp_name == rn:
                        load += l.generalized_load()
                rps.append({'node': np.asarray(rp.node.detach().cpu().numpy(), dtype=np.float64),
                            'free': self.RGC_remain_index[rp._RGC_index].copy(), 'load': load})
            couples = []
            for c in self._constraints.values():
                if not isinstance(c, constraints.Couple):
                    raise NotImplementedError('constraint %s' % type(c).__name__)
                couples.append((c._indexNodes + offsets[c.instance_name], rp_names.index(c.rp_name)))
            return lambda extra: CoupledPlan(X, free, groups + extra, couples, rps, device=self.device)
        return lambda extra: DevicePlan(X, free, groups + extra, device=self.device)

    def initialize_dynamic(self):
        for ins in self._instances.values():
            ins.initialize_dynamic()
        for l in self._loads.values():
            l.initialize_dynamic()

    def reinitialize(self, RGC):
        self.RGC = RGC
        self.GC = self._RGC2GC(self.RGC)

    # ------------------------------------------------------------------ GC <-> RGC
    def _u_from(self, RGC=None, GC=None):
        """Solver state from RGC or GC: the nodal float64 field [3N] of the instances, followed
        by the 6 coordinates of every reference point when the model has them."""
        if RGC is None:
            if GC is None:
                raise ValueError("Either RGC or GC must be provided.")
            return self.plan.gc_to_u(torch.as_tensor(GC, dtype=torch.float64, device=self.device).contiguous())
        blocks = [torch.as_tensor(RGC[ins._RGC_index]).to(self.device, torch.float64).reshape(-1)
                  for ins in self._instances.values()]
        blocks += [torch.as_tensor(RGC[rp._RGC_index]).to(self.device, torch.float64).reshape(-1)
                   for rp in self._reference_points.values()]
        U = torch.cat(blocks).contiguous()
        if self._coupled:
            self.plan.sync(U)                         # Couple.modify_RGC (couple.py:36-43)
        return U

    def _GC2RGC(self, GC):
        """assembly.py:470-492: scatter GC into RGC blocks (constrained entries zero, coupled
        nodes from their reference point)."""
        U = self._u_from(GC=GC)
        RGC = []
        for i, size in enumerate(self._RGC_size):
            RGC.append(torch.zeros(size, dtype=torch.float64, device=self.device))
        Un = U[:3 * self._n_nodes].view(-1, 3)
        for ins in self._instances.values():
            s = self.RGC_list_indexStart[ins._RGC_index] // 3
            RGC[ins._RGC_index] = Un[s:s + ins._RGC_requirements[0]].clone()
        for k, rp in enumerate(self._reference_points.values()):
            RGC[rp._RGC_index] = U[3 * self._n_nodes + 6 * k:3 * self._n_nodes + 6 * k + 6].clone()
        for b in self._boundarys.values():
            RGC = b.modify_RGC(RGC)
        return RGC

    def _RGC2GC(self, RGC):
        """assembly.py:494-499: free entries of every block, concatenated."""
        return self.plan.u_to_gc(self._u_from(RGC=RGC))

    def refine_RGC(self, RGC):
        return [r.clone().detach() for r in RGC]

    # ------------------------------------------------------------------ assembly
    def K_indices(self):
        v = getattr(self.plan, 'version', 0)
        if self._K_indices is None or self._K_version != v:
            self._K_indices = self.plan.csr_indices_coo()
            self._K_version = v
        return self._K_indices

    def assemble_Stiffness_Matrix(self, RGC=None, GC=None):
        """assembly.py:283-309: (R [n_free], K indices [2, nnz], K values [nnz]).

        K is returned already coalesced, in CSR order (the reference returns the
        same matrix as an uncoalesced COO list that _linear_solver.py:18-19
        coalesces into exactly this CSR)."""
        U = self._u_from(RGC, GC)
        R, val = self.plan.tangent(U, 1.0, F64)
        return R.clone(), self.K_indices(), val.clone()

    def _contact_loads(self):
        return [l for l in self._loads.values() if isinstance(l, loads.NodeToSurfaceContact)]

    def assemble_force(self, RGC=None, GC=None):
        """assembly.py:230-281 (contact loads: force only, subtracted as external forces)."""
        U = self._u_from(RGC, GC)
        R = self.plan.residual(U, 1.0, F64).clone()
        for c in self._contact_loads():
            R -= self.plan.u_to_gc(c.state(self.plan).forces(U, search=True))
        return R

    def _total_Potential_Energy(self, RGC=None, GC=None):
        """assembly.py:389-415: internal energy minus load potentials (plus contact energy)."""
        U = self._u_from(RGC, GC)
        E = self.plan.energy(U, 1.0)
        for c in self._contact_loads():
            st = c.state(self.plan)
            st.forces(U, search=True)
            E += st.energy()
        return torch.tensor(E, dtype=torch.float64)

    # ------------------------------------------------------------------ registries
    @staticmethod
    def _auto_name(store, base):
        number = len(store)
        while ('%s-%d' % (base, number)) in store:
            number += 1
        return '%s-%d' % (base, number)

    def add_part(self, part: Part, name: str = None) -> None:
        if name is None:
            name = self._auto_name(self._parts, part.__class__.__name__)
        if name in self._parts:
            raise ValueError(f"Part with name {name} already exists in the assembly.")
        self._parts[name] = part

    def get_part(self, name: str) -> Part:
        if name not in self._parts:
            raise ValueError(f"Part with name {name} does not exist in the assembly.")
        return self._parts[name]

    def delete_part(self, name: str) -> None:
        if name not in self._parts:
            raise ValueError(f"Part with name {name} does not exist in the assembly.")
        del self._parts[name]

    def add_instance(self, instance: Instance, name: str = None) -> None:
        if name is None:
            name = self._auto_name(self._instances, instance.__class__.__name__)
        if name in self._instances:
            raise ValueError(f"Instance with name {name} already exists in the assembly.")
        self._instances[name] = instance

    def get_instance(self, name: str) -> Instance:
        if name not in self._instances:
            raise ValueError(f"Instance with name {name} does not exist in the assembly.")
        return self._instances[name]

    def delete_instance(self, name: str) -> None:
        if name not in self._instances:
            raise ValueError(f"Instance with name {name} does not exist in the assembly.")
        del self._instances[name]

    def add_reference_point(self, rp: ReferencePoint, name: str = None):
        """assembly.py:557-576: default names rp-0, rp-1, ..."""
        if name is None:
            number = len(self._reference_points)
            while ('rp-%d' % number) in self._reference_points:
                number += 1
            name = 'rp-%d' % number
        self._reference_points[name] = rp
        return name

    def get_reference_point(self, name: str) -> ReferencePoint:
        if name in self._reference_points:
            return self._reference_points[name]
        raise ValueError(f"Reference point '{name}' not found in the model.")

    def delete_reference_point(self, name: str):
        if name not in self._reference_points:
            raise ValueError(f"Reference point '{name}' not found in the model.")
        del self._reference_points[name]

    def add_load(self, load: loads.BaseLoad, name: str = None):
        if name is None:
            name = self._auto_name(self._loads, load.__class__.__name__)
        self._loads[name] = load
        return name

    def add_loads(self, loads_dict):
        for name, load in loads_dict.items():
            self.add_load(load, name)

    def get_load(self, name: str):
        if name in self._loads:
            return self._loads[name]
        raise ValueError(f"Load '{name}' not found in the model.")

    def delete_load(self, name: str):
        if name not in self._loads:
            raise ValueError(f"Load '{name}' not found in the model.")
        del self._loads[name]

    def delete_all_loads(self):
        self._loads.clear()

    def add_constraint(self, constraint: constraints.BaseConstraint, name: str = None):
        if name is None:
            name = self._auto_name(self._constraints, constraint.__class__.__name__)
        self._constraints[name] = constraint
        return name

    def get_constraint(self, name: str):
        if name in self._constraints:
            return self._constraints[name]
        raise ValueError(f"Constraint '{name}' not found in the model.")

    def delete_constraint(self, name: str):
        if name not in self._constraints:
            raise ValueError(f"Constraint '{name}' not found in the model.")
        del self._constraints[name]

    def add_boundary(self, boundary: boundarys.BaseBoundary, name: str = None):
        if name is None:
            name = self._auto_name(self._boundarys, boundary.__class__.__name__)
        self._boundarys[name] = boundary
        return name

    def get_boundary(self, name: str):
        if name in self._boundarys:
            return self._boundarys[name]
        raise ValueError(f"Boundary '{name}' not found in the model.")

    def delete_boundary(self, name: str):
        if name not in self._boundarys:
            raise ValueError(f"Boundary '{name}' not found in the model.")
        del self._boundarys[name]
