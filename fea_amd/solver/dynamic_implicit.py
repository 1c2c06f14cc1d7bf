"""Newmark-beta implicit dynamics (reference: solver/dynamic_implicit.py:15-533).

Each time step minimises the incremental functional

    Pi_inc(GC) = Pi(GC) + (GC - GC_pre) . M (GC - GC_pre) / (2 beta dt^2)

(get_incremental_energy / get_incremental_stiffness_matrix, :100-144) from the
predictor GC_pre = GC_n + dt GV_n + (1/2 - beta) dt^2 GA_n (:181 writes the
literal 0.25, which is this value at the default beta = 1/4; any other beta
needs 1/2 - beta for the converged step to satisfy M GA + R = 0 with the
acceleration update of :261-281) with the Newton / Armijo iteration of the static solver, the effective tangent
K + M / (beta dt^2) in the tangent's own CSR (the consistent mass M is
assembled once into the same pattern, C3base.py:348-384 / assembly.py:421-443)
and the reference's stopping rule max|dGC| < tol and max|R| < tol (:468); then
the Newmark update of velocity and acceleration (:261-281).

Deviation: the initial acceleration solves M GA0 = -R(GC0), the equation of
motion (the reference solves M GA0 = +R, :266-269, the opposite sign; the two
agree when the initial state is in equilibrium).  ``rel_tol`` / ``cg_tol`` /
``cg_max_iter`` are the static solver's options; the reference's Pardiso branch
is replaced by PCG as there.
"""
from __future__ import annotations

import time

import torch

from .. import _native as nat
from .._native import F64, call
from ._linear_solver import PCG
from .static_implicit import StaticImplicitSolver


class DynamicImplicitSolver(StaticImplicitSolver):

    def __init__(self, maximum_iteration: int = 10000, deltaT: float = 1e-2, time_end: float = 1.0,
                 tol_error: float = 1e-5, *, rel_tol: float | None = None, cg_tol: float = 1e-5,
                 cg_max_iter: int = 1500, armijo_slack: float = 0.0, verbose: bool = False,
                 beta: float = 0.25, gamma: float = 0.5, store: bool = True) -> None:
        super().__init__(maximum_iteration, tol_error, precision='fp64', rel_tol=rel_tol, load_steps=1,
                         cg_tol=cg_tol, cg_max_iter=cg_max_iter, armijo_slack=armijo_slack, predictor=False,
                         verbose=verbose)
        self._deltaT = float(deltaT)
        self._time_end = float(time_end)
        self._beta, self._gamma = float(beta), float(gamma)
        self.store = bool(store)
        self._GC_list, self._GV_list, self._GA_list, self._time_list = [], [], [], []
        self._GC_pre = None

    def initialize(self, assembly, *args, **kwargs):
        super().initialize(assembly)
        self.assembly.initialize_dynamic()
        self._GC_list, self._GV_list, self._GA_list, self._time_list = [], [], [], []

    def set_deltaT(self, deltaT: float):
        self._deltaT = float(deltaT)

    # ------------------------------------------------------------------ incremental functional
    def _c(self):
        return 1.0 / (self._beta * self._deltaT ** 2)

    def _inertia(self, U):
        """(d, M d) with d = GC(U) - GC_pre (free-DOF float64 device vectors)."""
        p = self.assembly.plan
        d = p.u_to_gc(U) - self._GC_pre
        Md = torch.empty_like(d)
        p.spmv(p.mass_matrix(), d, Md, F64)
        return d, Md

    def _energy(self, U, scale):
        """get_incremental_energy (dynamic_implicit.py:100-111)."""
        d, Md = self._inertia(U)
        return self.assembly.plan.energy(U, scale) + 0.5 * self._c() * float(torch.dot(d, Md))

    def _tangent(self, U, scale):
        """get_incremental_stiffness_matrix (dynamic_implicit.py:115-144): R + M d c, K + M c."""
        p = self.assembly.plan
        R, K = p.tangent(U, scale, F64)
        _, Md = self._inertia(U)
        c = self._c()
        Re = R + c * Md
        if getattr(self, '_Keff', None) is None or self._Keff.numel() != K.numel():
            self._Keff = torch.empty_like(K)
        self._Keff.copy_(K)
        self._Keff.add_(p.mass_matrix(), alpha=c)
        return Re, self._Keff

    def _converged_abs(self, dmax, rmax, tol_error):
        """dynamic_implicit.py:468 (no separate small-residual exit)."""
        return dmax < tol_error and rmax < tol_error

    # ------------------------------------------------------------------ Newmark
    def get_next_velocity(self, GC_now, GC0, GV0, GA0, deltaT):
        """dynamic_implicit.py:261-281."""
        b, g = self._beta, self._gamma
        GV = g / (b * deltaT) * (GC_now - GC0) + (1 - g / b) * GV0 + deltaT * (1 - g / (2 * b)) * GA0
        GA = (GC_now - GC0) / (b * deltaT ** 2) - GV0 / (b * deltaT) - (1 / (2 * b) - 1) * GA0
        return GV, GA

    def get_current_acceleration(self, GC0, GV0=None):
        """M GA0 = -R(GC0) by Jacobi-PCG on the consistent mass (the reference solves
        M GA0 = +R with Pardiso, dynamic_implicit.py:266-269)."""
        p = self.assembly.plan
        U = p.gc_to_u(torch.as_tensor(GC0, dtype=torch.float64, device=p.device).contiguous())
        R = p.residual(U, 1.0, F64).clone()
        pcg = PCG(p.n_free, p.nnz, p.rowptr, p.col, p.diag_slot, F64, p.device, grp=p.grp,
                  n_groups=p.spmv_groups, spmv_col=p.spmv_col)
        pcg.set_matrix(p.mass_matrix())
        return -pcg.solve(R, None, tol=1e-28, max_iter=20000).clone()

    def solve(self, GC0=None, GV0=None, *args, **kwargs) -> bool:
        """dynamic_implicit.py:146-235; returns True, or False if a step fails to converge."""
        t_start = time.time()
        p = self.assembly.plan
        GC = (self.assembly.GC if GC0 is None else torch.as_tensor(GC0)).to(p.device, torch.float64).clone()
        GV = torch.zeros_like(GC) if GV0 is None else torch.as_tensor(GV0).to(p.device, torch.float64).clone()
        GA = self.get_current_acceleration(GC)
        t, dt = 0.0, self._deltaT
        self._GC_list, self._GV_list, self._GA_list, self._time_list = [GC], [GV], [GA], [0.0]
        self.stats = {'cg_iters': 0, 'energy_evals': 0, 'residual_history': [], 'increments': 0, 'steps': 0}
        self._r_ref = None
        self._increment = 1
        while t < self._time_end - 1e-12 * max(1.0, self._time_end):
            self._GC_pre = GC + dt * GV + (0.5 - self._beta) * dt ** 2 * GA
            U = p.gc_to_u(GC.contiguous())
            self._iter_now = 0
            if not self._newton(U, 1.0, self.tol_error):
                self.assembly.GC = GC
                return False
            GC_new = p.u_to_gc(U)
            GV, GA = self.get_next_velocity(GC_new, GC, GV, GA, dt)
            GC = GC_new
            t += dt
            self.stats['steps'] += 1
            if self.store:
                self._GC_list.append(GC)
                self._GV_list.append(GV)
                self._GA_list.append(GA)
                self._time_list.append(t)
        self.GC = GC
        self.assembly.GC = GC
        self.assembly.RGC = self.assembly.refine_RGC(self.assembly._GC2RGC(GC))
        self.stats['time'] = time.time() - t_start
        return True
