"""Static implicit Newton solver (reference: solver/static_implicit.py:12-353).

The Newton / Armijo line-search / Jacobi-CG structure of the reference is kept
(same constants: c1 = 0.3, alpha halving to 1e-12, CG tolerance on the
squared scaled residual, 1500 / 6000 CG iterations, warm start from the
previous direction, the low-alpha bookkeeping).  All heavy work runs in HIP
kernels on the device plan; the host only takes the branch decisions.

Additions (keyword-only, defaults reproduce the reference's rules except for the
round-off step below):

* In every mode the line search takes the full Newton step when the predicted
  energy decrease |dGC . R| is at most 1e-13 |E| (below the energy's float64
  resolution, so the Armijo test cannot resolve it); the reference backtracks
  alpha to 0 there and decides from max|R|.

* ``precision``: 'fp64' (everything FP64); 'mixed' (FP64 element arithmetic,
  tangent stored once-rounded in FP32, PCG on the FP32 matrix with FP64
  vectors and recurrences); 'mixed-f32' (as 'mixed' but the element stiffness
  contraction and its staging in FP32); 'mixed-ir' (FP32 element contraction, FP32
  inner PCG inside an FP64 iterative-refinement loop of ``ir_steps`` passes).
  In every mode the residual, energy, convergence test and J-terms are FP64.
* ``rel_tol``: converge on ||R||_2 <= rel_tol ||R_ref||_2 with R_ref the
  residual of the full load at u = 0 (the external load norm), or when a full
  Newton step changes the displacement by at most ``step_tol`` relative to
  max|u| (the correction has reached float64 round-off, below which the
  residual cannot fall); the reference tests |R|_inf < tol_error absolutely
  (static_implicit.py:290).
* ``load_steps``: ramp all loads in equal increments (the reference has no
  increments; its harnesses loop the load by hand, test_multicase.py:64-72).
* ``armijo_slack``: relative energy round-off allowance in the Armijo test, so
  the line search does not reject full steps whose energy change is below
  float64 resolution near convergence (0 reproduces the reference).
* ``predictor``: with load increments, start increment k from the linear
  extrapolation 2 u_{k-1} - u_{k-2} of the converged states (True / 'linear'), the
  quadratic one 3 u_{k-1} - 3 u_{k-2} + u_{k-3} ('quadratic', from the third
  increment), the cubic one ('cubic') or any order m <= 6 (an int: the degree-m
  polynomial through the last m + 1 converged states, lower while fewer exist), the
  order chosen per increment ('auto': the order m <= 6 whose prediction of u_{k-1}
  from u_{k-2}, u_{k-3}, ... misses it least in the 2-norm, raised by one when that is
  the highest order testable, and replaced by u_{k-1} when the prediction's potential
  energy at the new load exceeds u_{k-1}'s), or from u_{k-1} (False: for paths with large rotations,
  where the extrapolation can start far from the next equilibrium).
* ``forcing='ew'``: inexact-Newton forcing term of Eisenstat and Walker
  (choice 2, gamma = 0.9, alpha = 2 with the safeguard gamma eta_{k-1}^2)
  for the PCG relative tolerance eta_k, clamped to [sqrt(cg_tol), eta_max];
  the default 'fixed' keeps the reference's constant cg_tol.
* ``preconditioner='block'``: node-block (3x3) Jacobi instead of the reference's point
  Jacobi scaling (other CG iterates, the same Newton solution); ``'amg'``: a
  smoothed-aggregation multigrid V-cycle (solver/amg.py) built once per model from the
  tangent at u = 0, its finest level smoothed with the current tangent
  (``amg_degree``: Chebyshev degree of the smoother).
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

from .. import _native as nat
from .._native import F32, F64, call, ptr
from ._linear_solver import PCG
from .basesolver import BaseSolver


_PREDICTORS = {False: 0, True: 1, 'none': 0, 'linear': 1, 'quadratic': 2, 'cubic': 3}
AUTO_MAX_ORDER = 6


def predictor_order(predictor):
    """Extrapolation order of a ``predictor`` argument: False / 'none' 0, True / 'linear' 1,
    'quadratic' 2, 'cubic' 3, or an int in [0, 6]; 'auto' gives the largest order it may pick."""
    if predictor == 'auto':
        return AUTO_MAX_ORDER
    if isinstance(predictor, (bool, str)) and predictor in _PREDICTORS:
        return _PREDICTORS[predictor]
    if isinstance(predictor, int) and 0 <= predictor <= 6:
        return int(predictor)
    raise ValueError("predictor must be False, True, 'linear', 'quadratic', 'cubic' or an order 0..6")


class StaticImplicitSolver(BaseSolver):

    def __init__(self, maximum_iteration: int = 10000, tol_error: float = 1e-5, *,
                 precision: str = 'fp64', rel_tol: float | None = None, load_steps: int = 1,
                 cg_tol: float = 1e-5, cg_max_iter: int = 1500, ir_steps: int = 1,
                 armijo_slack: float = 0.0, predictor=True, verbose: bool = False,
                 pcg_chunk: int = 32, progress=None, step_tol: float = 1e-12, forcing: str = 'fixed',
                 eta_max: float = 1e-2, preconditioner: str = 'jacobi', amg_degree: int = 2) -> None:
        super().__init__()
        if preconditioner not in ('jacobi', 'block', 'amg'):
            raise ValueError("preconditioner must be 'jacobi', 'block' or 'amg'")
        if preconditioner == 'amg' and precision == 'mixed-ir':
            raise ValueError("preconditioner='amg' needs float64 PCG vectors (not 'mixed-ir')")
        self.preconditioner = preconditioner
        self.amg_degree = int(amg_degree)
        if precision not in ('fp64', 'mixed', 'mixed-f32', 'mixed-ir'):
            raise ValueError("precision must be 'fp64', 'mixed', 'mixed-f32' or 'mixed-ir'")
        self.maximum_iteration = maximum_iteration
        self.tol_error = tol_error
        self._iter_now = 0
        self._maximum_step_length = 1e10
        self._low_alpha_count = 0
        self.precision = precision
        self.rel_tol = rel_tol
        self.step_tol = float(step_tol)
        if forcing not in ('fixed', 'ew'):
            raise ValueError("forcing must be 'fixed' or 'ew'")
        self.forcing = forcing
        self.eta_max = float(eta_max)
        self.load_steps = int(load_steps)
        self.cg_tol = float(cg_tol)
        self.cg_max_iter = int(cg_max_iter)
        self.ir_steps = int(ir_steps)
        self.armijo_slack = float(armijo_slack)
        self.predictor = predictor_order(predictor)
        self.predictor_auto = predictor == 'auto'
        self.progress = progress      # callable(increment, stats) after each Newton iteration
        self.verbose = verbose
        self.pcg_chunk = int(pcg_chunk)
        self.stats = {}

    def initialize(self, assembly, *args, **kwargs):
        super().initialize(assembly)
        from ..assemble.loads.contact import NodeToSurfaceContact
        if any(isinstance(l, NodeToSurfaceContact) for l in assembly._loads.values()):
            raise NotImplementedError('node-to-surface contact has no tangent; use DynamicExplicitSolver')
        self._pcg = {}
        if self.preconditioner == 'amg':
            if getattr(assembly.plan, 'n_rp', 0):
                # the hierarchy aggregates node DOF triples with their rigid-body modes; the
                # reduced system of a Couple model also has 6 generalized DOFs per point
                raise NotImplementedError("preconditioner='amg' is not available for models with reference "
                                          "points (Couple constraints): use 'block' or 'jacobi'")
            self._build_amg()

    def _build_amg(self):
        """The multigrid hierarchy of this model, from its float64 tangent at u = 0 under the
        full load (built once per model, like the sparsity pattern)."""
        from .amg import AMG
        p = self.assembly.plan
        U = torch.zeros(3 * p.n_nodes, dtype=torch.float64, device=p.device)
        _, K = p.tangent(U, 1.0, F64)
        self.amg = AMG(p, K, degree=self.amg_degree)
        self.amg.version = getattr(p, 'version', 0)
        self.amg_setup_s = self.amg.setup_time

    # ------------------------------------------------------------------ API
    def solve(self, GC0=None, *args, **kwargs):
        """static_implicit.py:44-76; returns the converged GC or False."""
        t0 = time.time()
        if GC0 is None:
            GC0 = self.assembly.GC
        result = self._solve_iteration(GC=GC0, tol_error=self.tol_error)
        if isinstance(result, bool):
            return result
        self.GC = result
        self.assembly.GC = self.GC
        self.assembly.RGC = self.assembly.refine_RGC(self.assembly._GC2RGC(self.GC))
        self.stats['time'] = time.time() - t0
        if self.verbose:
            print('total_iter:%d, total_time:%.2f' % (self._iter_now, self.stats['time']))
        return self.GC

    def get_total_energy(self, GC_now):
        return self.assembly._total_Potential_Energy(GC=GC_now)

    def get_stiffness_matrix(self, GC_now):
        return self.assembly.assemble_Stiffness_Matrix(GC=GC_now)

    # ------------------------------------------------------------------ internals
    def _pcg_for(self, mat_dtype, vec_dtype):
        # keyed on the plan's CSR version: a contact plan re-creates its pattern when the
        # contact pairs leave the staged set
        key = (mat_dtype, vec_dtype, getattr(self.assembly.plan, 'version', 0))
        if key not in self._pcg:
            self._pcg = {k: v for k, v in self._pcg.items() if k[2] == key[2]}
            p = self.assembly.plan
            amg = None
            if self.preconditioner == 'amg':
                if getattr(self, 'amg', None) is None or self.amg.version != key[2]:
                    self._build_amg()                # the pattern changed (contact pairs re-staged)
                amg = self.amg
            self._pcg[key] = PCG(p.n_free, p.nnz, p.rowptr, p.col, p.diag_slot, mat_dtype, p.device,
                                 vec_dtype=vec_dtype, grp=p.grp, n_groups=p.spmv_groups, chunk=self.pcg_chunk,
                                 precond=self.preconditioner, spmv_col=p.spmv_col, amg=amg)
        return self._pcg[key]

    def _vec_stats(self, x):
        """(sum x^2, max|x|) of a float64 device vector, one host sync."""
        p = self.assembly.plan
        n = x.numel()
        nb = nat.lib().fea_vec_blocks(n)
        part = p._scratch('stats', 2 * nb)
        call('fea_vec_stats', n, x, part, nat.stream_ptr())
        h = part.view(nb, 2).cpu().numpy()
        return float(h[:, 0].sum()), float(h[:, 1].max())

    def _tangent(self, U, scale):
        p = self.assembly.plan
        if self.precision == 'fp64':
            return p.tangent(U, scale, F64)
        if self.precision == 'mixed':
            return p.tangent(U, scale, F64, val_dtype=F32)
        # FP32 contraction and staging; the residual of the same pass is float64
        return p.tangent(U, scale, F32, val_dtype=F32)

    def _solve_linear(self, K, R, dGC0, iter_now, alpha0, tol=None):
        """static_implicit.py:297-350 (CG branch; the Pardiso branch is replaced by PCG)."""
        p = self.assembly.plan
        if alpha0 < 1e-1:
            self._low_alpha_count += 1
        else:
            self._low_alpha_count = 0
        max_iter = 4 * self.cg_max_iter if (iter_now % 20 == 0 or self._low_alpha_count > 0) else self.cg_max_iter
        s = nat.stream_ptr()
        n = p.n_free
        b = p._scratch('rhs', n)
        b.zero_()
        call('fea_axpy', n, -1.0, F64, R, b, s)              # b = -R
        out = p._scratch('dgc', n)
        if self.precision in ('fp64', 'mixed', 'mixed-f32'):
            pcg = self._pcg_for(F64 if self.precision == 'fp64' else F32, F64)
            pcg.set_matrix(K)
            x = pcg.solve(b, dGC0, tol=self.cg_tol if tol is None else tol, max_iter=max_iter)
            out.copy_(x)
            self.stats['cg_iters'] = self.stats.get('cg_iters', 0) + pcg.iterations
        else:
            pcg = self._pcg_for(F32, F32)
            pcg.set_matrix(K)
            out.copy_(dGC0)
            r = p._scratch('ir_r', n)
            nb = nat.lib().fea_vec_blocks(n)
            part = p._scratch('ir_part', nb)
            for _ in range(max(1, self.ir_steps)):
                call('fea_spmv_residual_mixed', n, p.rowptr, p.col, K, out, b,
                     r, part, s)
                d = pcg.solve(r, None, tol=self.cg_tol if tol is None else tol, max_iter=max_iter)
                call('fea_axpy', n, 1.0, F32, d, out, s)
                self.stats['cg_iters'] = self.stats.get('cg_iters', 0) + pcg.iterations
        return out

    def _line_search(self, U, dGC, R, energy0, scale, Unew):
        """static_implicit.py:98-177: Armijo backtracking on the total energy."""
        p = self.assembly.plan
        alpha, c1 = 1.0, 0.3
        dd = p.dot(dGC, R)
        deltaE = float(dd.item())
        sign = 1.0
        if deltaE > 0:
            sign, deltaE = -1.0, -deltaE
        if not math.isfinite(deltaE):
            raise ValueError('dGC has nan or inf')
        if abs(deltaE) <= 1e-13 * abs(energy0):
            # the predicted decrease is below the energy's rounding: the Armijo test
            # cannot resolve it, so take the full Newton step (only near convergence).
            # A deviation in every mode: the reference backtracks alpha to 0 here and
            # then decides convergence from max|R| (static_implicit.py:98-177, :290).
            p.dof_axpy(U, sign, dGC, Unew)
            self.stats['energy_evals'] = self.stats.get('energy_evals', 0) + 1
            return 1.0, self._energy(Unew, scale)
        slack = self.armijo_slack * abs(energy0)
        while True:
            p.dof_axpy(U, sign * alpha, dGC, Unew)
            e_new = self._energy(Unew, scale)
            self.stats['energy_evals'] = self.stats.get('energy_evals', 0) + 1
            if (not math.isfinite(e_new)) or e_new > energy0 + c1 * deltaE * alpha + slack:
                alpha *= 0.5
                if alpha < 1e-12:
                    return 0.0, energy0
            else:
                return alpha, e_new

    def _energy(self, U, scale):
        """The functional the Newton iteration minimises (total potential energy)."""
        return self.assembly.plan.energy(U, scale)

    def _converged_abs(self, dmax, rmax, tol_error):
        """Absolute convergence rule without ``rel_tol`` (static_implicit.py:290)."""
        return (dmax < tol_error and rmax < tol_error) or rmax < 1e-6

    def _newton(self, U, scale, tol_error):
        p = self.assembly.plan
        energy = self._energy(U, scale)
        dGC = torch.zeros(p.n_free, dtype=torch.float64, device=p.device)
        Unew = torch.empty_like(U)
        low_alpha, alpha, it = 0, 0.0, 0
        r0 = self._r_ref
        r_prev, eta_prev = None, self.eta_max
        while True:
            if self._iter_now >= self.maximum_iteration:
                return False
            R, K = self._tangent(U, scale)
            it += 1
            self._iter_now += 1
            rss, rmax = self._vec_stats(R)
            self.stats['residual_history'].append(math.sqrt(rss))
            if self.rel_tol is not None:
                if r0 is None:
                    r0 = math.sqrt(rss)
                if math.sqrt(rss) <= self.rel_tol * r0:
                    return True
            tol = None
            if self.forcing == 'ew':
                rn = math.sqrt(rss)
                eta_min = math.sqrt(self.cg_tol)
                if r_prev is None:
                    eta = self.eta_max
                else:
                    eta = 0.9 * (rn / r_prev) ** 2
                    if 0.9 * eta_prev ** 2 > 0.1:
                        eta = max(eta, 0.9 * eta_prev ** 2)
                eta = min(max(eta, eta_min), self.eta_max)
                r_prev, eta_prev = rn, eta
                tol = eta * eta
            d = self._solve_linear(K, R, dGC, it, alpha, tol)
            dGC.copy_(d)
            alpha, e_new = self._line_search(U, dGC, R, energy, scale, Unew)
            if alpha == 0:
                if self.rel_tol is not None or rmax > tol_error:
                    return False
                return True
            if alpha < 0.01:
                low_alpha += 1
            else:
                low_alpha = max(0, low_alpha - 5)
            if low_alpha > 10:
                return False
            U.copy_(Unew)
            energy = e_new
            if self.rel_tol is not None and alpha == 1.0:
                _, dmax = self._vec_stats(dGC)
                _, umax = self._vec_stats(U)
                if dmax <= self.step_tol * umax:
                    return True
            if self.progress is not None:
                self.progress(self._increment, dict(self.stats, newton_iters=self._iter_now, alpha=alpha,
                                                    residual=math.sqrt(rss)))
            if self.verbose:
                print('iter %3d  alpha %.3g  energy %.10e  |R|inf %.3e' % (self._iter_now, alpha, energy, rmax))
            if self.rel_tol is None:
                _, dmax = self._vec_stats(dGC)
                if self._converged_abs(dmax, rmax, tol_error):
                    return True

    @staticmethod
    def _extrapolate(hist, m, out):
        """out <- equal-step Lagrange extrapolation of degree m through hist[0..m] (newest
        first) to the next level: sum_j (-1)^j C(m+1, j+1) hist[j]."""
        s = nat.stream_ptr()
        out.zero_()
        for j in range(m + 1):
            call('fea_axpy', out.numel(), (-1.0) ** j * math.comb(m + 1, j + 1), F64, hist[j], out, s)

    def _auto_order(self, p, hist, max_order):
        """The order whose extrapolation from hist[1:] misses hist[0] least (2-norm); one
        higher when the best is the highest order hist can test (the miss still falls)."""
        top = min(max_order, len(hist) - 2)
        if not hasattr(self, '_pred_tmp') or self._pred_tmp.numel() != hist[0].numel():
            self._pred_tmp = torch.empty_like(hist[0])
        t = self._pred_tmp
        miss = []
        for m in range(top + 1):
            self._extrapolate(hist[1:], m, t)
            call('fea_axpy', t.numel(), -1.0, F64, hist[0], t, nat.stream_ptr())
            miss.append(float(p.dot(t, t).item()))
        best = int(np.argmin(miss))
        if best == top and top + 1 <= min(max_order, len(hist) - 1):
            best = top + 1
        return best

    def _solve_iteration(self, GC, tol_error):
        """static_implicit.py:179-295, once per load increment."""
        p = self.assembly.plan
        self._iter_now = 0
        self.stats = {'cg_iters': 0, 'energy_evals': 0, 'residual_history': [], 'increments': 0}
        U = p.gc_to_u(torch.as_tensor(GC, dtype=torch.float64, device=p.device).contiguous())
        self._r_ref = None
        if self.rel_tol is not None:
            zero = torch.zeros_like(U)
            rss, _ = self._vec_stats(p.residual(zero, 1.0, F64))
            self._r_ref = math.sqrt(rss) if rss > 0 else None
        # converged states of the previous load levels, newest first (u_0 = the start)
        order = self.predictor
        keep = order + 2 if self.predictor_auto else order + 1
        hist = [U.clone()] if order and self.load_steps > 1 else []
        self.stats['predictor_orders'] = []
        for k in range(1, self.load_steps + 1):
            if hist and k >= 2:
                m = self._auto_order(p, hist, order) if self.predictor_auto else min(order, len(hist) - 1)
                self._extrapolate(hist, m, U)
                if hasattr(p, 'sync'):
                    p.sync(U)                      # coupled nodes back onto their reference points
                if self.predictor_auto and m > 0:
                    # keep the extrapolation only where its potential energy at the new load
                    # is below the last converged state's
                    scale = k / self.load_steps
                    if not self._energy(U, scale) <= self._energy(hist[0], scale):
                        m = 0
                        U.copy_(hist[0])
                self.stats['predictor_orders'].append(m)
            self._increment = k
            ok = self._newton(U, k / self.load_steps, tol_error)
            self.stats['increments'] = k
            if not ok:
                if self.verbose:
                    print('load increment %d failed to converge' % k)
                return False
            if hist:
                hist.insert(0, U.clone())
                del hist[keep:]
        self.stats['newton_iters'] = self._iter_now
        return p.u_to_gc(U)
