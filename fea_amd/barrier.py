"""The reference's Gauss-point-pair contact barrier on the device (reference:
loads/contact.py:10-1218, ContactSelf / Contact) and a plan whose sparsity follows it.

``BarrierSet`` holds one contact load's T3 facets (global node ids) and runs, on the GPU,
the candidate search (facet centroids within ``radius`` at the current configuration,
hashed grid of ``fea_contact_grid_build`` + ``fea_barrier_search``), the energy
(``fea_barrier_energy``), the coverage test of a staged pair set (``fea_barrier_missing``)
and the choice of that set: the pairs within ``radius + skin`` whose self-contact ratio is
not identically zero (``fea_barrier_active``).

``LivePlan`` wraps a ``DevicePlan`` / ``CoupledPlan`` whose groups include one 6-node
"barrier" group per contact load (pair = facet p's nodes then facet q's): the element
pattern builder turns the staged pairs into CSR couplings and staging slots like any
element, and ``fea_barrier_assemble`` writes every staged pair's rows and blocks (zero for
pairs that are not candidates at the configuration).  Before a tangent or residual, the
exact candidate pairs at the state are checked against the staged set; a pair outside it
rebuilds the plan around the pairs within ``radius + skin`` of that state.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat
from ._native import F32, F64, call


class BarrierParams(ctypes.Structure):
    """Mirror of ``fea_barrier_params`` (include/fea_amd.h)."""
    _fields_ = [(f, ctypes.c_double) for f in ('c_g', 'k_g', 'h_t', 'r_h', 's_f', 'e_f', 'ign_min', 'ign_max',
                                               'radius')] + [('self_contact', ctypes.c_int32),
                                                             ('reserved', ctypes.c_int32)]


def _i32(a, device):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)


class BarrierSet:
    def __init__(self, conn1, conn2, X, params, device, name=''):
        self.device = device
        self.name = name
        self.self_contact = conn2 is None
        c1 = np.ascontiguousarray(conn1, dtype=np.int32).reshape(-1, 3)
        c2 = c1 if conn2 is None else np.ascontiguousarray(conn2, dtype=np.int32).reshape(-1, 3)
        self.c1_np, self.c2_np = c1, c2
        self.F1, self.F2 = c1.shape[0], c2.shape[0]
        self.c1 = _i32(c1 if self.F1 else np.zeros((1, 3)), device)
        self.c2 = self.c1 if conn2 is None else _i32(c2 if self.F2 else np.zeros((1, 3)), device)
        self.X = X                                                    # device float64 [3N]
        p = BarrierParams()
        for k in ('c_g', 'k_g', 'h_t', 'r_h', 's_f', 'e_f', 'ign_min', 'ign_max'):
            setattr(p, k, float(params[k]))
        self.radius = float(params['search_ratio']) * float(params['h_t'])
        p.radius = self.radius
        p.self_contact = 1 if self.self_contact else 0
        self.P = p
        self.bits = max(10, int(np.ceil(np.log2(max(2, 2 * self.F2)))))
        nb = 1 << self.bits
        z = lambda n, dt=torch.int32: torch.zeros(max(1, n), dtype=dt, device=device)
        self.cent = z(3 * self.F2, torch.float64)
        self.key, self.bstart, self.bitems = z(self.F2), z(nb + 1), z(self.F2)
        self.work = z(max(nb, self.F2))
        self.count, self.offsets = z(self.F1), z(self.F1 + 1)
        self.pi, self.pj = z(1024), z(1024)
        self.flag = z(1024)
        self.missing = z(1)
        self.epart = z(1, torch.float64)
        self.eout = z(1, torch.float64)
        self.n_pairs = 0
        # the staged pair set (pattern), ascending keys i * F2 + j
        self.spi, self.spj = np.zeros(0, dtype=np.int32), np.zeros(0, dtype=np.int32)
        self.spi_d, self.spj_d = z(1), z(1)
        self.keys = torch.zeros(1, dtype=torch.int64, device=device)
        self.origin = np.zeros(3)

    def _mp(self):
        return ctypes.addressof(self.P)

    @staticmethod
    def _s():
        return nat.stream_ptr()

    # ------------------------------------------------------------------ search
    def search(self, U, radius):
        """Candidate pairs at the configuration X + U within ``radius`` into (pi, pj)[:n]; returns n."""
        s = self._s()
        if self.F1 == 0 or self.F2 == 0:
            self.n_pairs = 0
            return 0
        call('fea_contact_grid_build', self.F2, self.c2, self.X, U, self.origin, float(radius), self.bits,
             self.cent, self.key, self.bstart, self.bitems, self.work, s)
        call('fea_barrier_search', self.F1, self.c1, self.X, U, self.cent, self.bstart, self.bitems, self.origin,
             float(radius), self.bits, float(radius), 1 if self.self_contact else 0, None, self.count, None, None, s)
        call('fea_scan_i32', self.F1, self.count, self.offsets, s)
        n = int(self.offsets[self.F1].item())
        if n > self.pi.numel():
            cap = int(n * 1.5) + 1024
            self.pi = torch.zeros(cap, dtype=torch.int32, device=self.device)
            self.pj = torch.zeros(cap, dtype=torch.int32, device=self.device)
            self.flag = torch.zeros(cap, dtype=torch.int32, device=self.device)
        if n:
            call('fea_barrier_search', self.F1, self.c1, self.X, U, self.cent, self.bstart, self.bitems,
                 self.origin, float(radius), self.bits, float(radius), 1 if self.self_contact else 0, self.offsets,
                 None, self.pi, self.pj, s)
        self.n_pairs = n
        return n

    def energy(self, U):
        """Barrier energy at X + U (device float64 scalar tensor [1])."""
        n = self.search(U, self.radius)
        if n == 0:
            self.eout.zero_()
            return self.eout
        nb = nat.lib().fea_barrier_energy_blocks(n)
        if self.epart.numel() < nb:
            self.epart = torch.zeros(nb, dtype=torch.float64, device=self.device)
        s = self._s()
        call('fea_barrier_energy', self._mp(), n, self.pi, self.pj, self.c1, self.c2, self.X, U, self.epart, s)
        call('fea_reduce_sum', nb, self.epart, self.eout, s)
        return self.eout

    def covered(self, U):
        """True when every candidate pair at X + U with a nonzero ratio is staged."""
        n = self.search(U, self.radius)
        if n == 0:
            return True
        self.missing.zero_()
        call('fea_barrier_missing', self._mp(), n, self.pi, self.pj, self.c1, self.c2, self.X, self.keys,
             int(self.spi.shape[0]), int(max(1, self.F2)), self.missing, self._s())
        return int(self.missing.item()) == 0

    def restage(self, U, skin):
        """Stage the pairs within radius + skin at X + U whose ratio is not identically zero."""
        n = self.search(U, self.radius + skin)
        if n:
            call('fea_barrier_active', self._mp(), n, self.pi, self.pj, self.c1, self.c2, self.X, self.flag,
                 self._s())
            pi = self.pi[:n].cpu().numpy()
            pj = self.pj[:n].cpu().numpy()
            keep = self.flag[:n].cpu().numpy() != 0
            pi, pj = pi[keep], pj[keep]
        else:
            pi = pj = np.zeros(0, dtype=np.int32)
        keys = pi.astype(np.int64) * max(1, self.F2) + pj
        o = np.argsort(keys, kind='stable')
        self.spi, self.spj = pi[o].astype(np.int32), pj[o].astype(np.int32)
        self.spi_d = _i32(self.spi if self.spi.size else np.zeros(1), self.device)
        self.spj_d = _i32(self.spj if self.spj.size else np.zeros(1), self.device)
        self.keys = torch.as_tensor(keys[o] if keys.size else np.zeros(1, dtype=np.int64), device=self.device)

    def group_conn(self):
        """Pseudo-element connectivity [P, 6] of the staged pairs (facet p's nodes, facet q's)."""
        return np.concatenate([self.c1_np[self.spi], self.c2_np[self.spj]], axis=1) if self.spi.size else \
            np.zeros((0, 6), dtype=np.int32)

    # ------------------------------------------------------------------ assembly
    def assemble(self, U, dst_k, kstage, kprec, dst_r, rstage):
        n = int(self.spi.shape[0])
        if n == 0:
            return
        call('fea_barrier_assemble', self._mp(), kprec, n, self.spi_d, self.spj_d, self.c1, self.c2, self.X, U,
             dst_k, kstage, dst_r, rstage, self._s())


class LivePlan:
    """A plan whose sparsity follows the contact pairs (see the module docstring).
    ``build(extra_groups)`` constructs the underlying plan; attributes not defined here
    are the underlying plan's, so a rebuild is invisible to callers except through
    ``version`` (solvers keyed on the CSR re-create their linear solvers when it changes)."""

    def __init__(self, build, barriers, n_nodes, skin_ratio=0.5):
        self.__dict__['_build'] = build
        self.__dict__['barriers'] = barriers
        self.__dict__['n_nodes_'] = int(n_nodes)
        self.__dict__['skin'] = [skin_ratio * b.radius for b in barriers]
        self.__dict__['version'] = 0
        self.__dict__['rebuilds'] = 0
        U0 = torch.zeros(3 * n_nodes, dtype=torch.float64, device=barriers[0].device)
        for b, sk in zip(barriers, self.skin):
            b.restage(U0, sk)
        self.__dict__['_plan'] = build(self._groups())

    def _groups(self):
        from .device import Group
        return [Group('barrier', 6, b.group_conn(), load=b, name='barrier:%s' % b.name) for b in self.barriers]

    def __getattr__(self, name):
        return getattr(self.__dict__['_plan'], name)

    def _nodal(self, U):
        return U[:3 * self.n_nodes_]

    def ensure(self, U):
        """Rebuild the plan when a candidate pair at state U is not staged."""
        Un = self._nodal(U)
        if all(b.covered(Un) for b in self.barriers):
            return False
        for b, sk in zip(self.barriers, self.skin):
            b.restage(Un, sk)
        self.__dict__['_plan'] = self._build(self._groups())
        self.__dict__['version'] += 1
        self.__dict__['rebuilds'] += 1
        return True

    def tangent(self, U, *args, **kwargs):
        self.ensure(U)
        return self._plan.tangent(U, *args, **kwargs)

    def residual(self, U, *args, **kwargs):
        self.ensure(U)
        return self._plan.residual(U, *args, **kwargs)

    def energy(self, U, load_scale=1.0):
        E = self._plan.energy(U, load_scale)
        Un = self._nodal(U)
        for b in self.barriers:
            E += float(b.energy(Un).item())
        return E
