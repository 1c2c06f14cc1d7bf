"""Sensitivities through hand-written adjoint kernels behind torch.autograd.Function.

The reference obtains shape and material sensitivities by running torch
autograd through its assembly (C3base.py:387-452 and pressure.py:80-132 are
plain torch ops on the coordinates, so ``requires_grad`` nodes and a
``backward()`` give gradients at the cost of storing every intermediate).
Here the backward passes are explicit kernels that recompute Gauss-point
quantities from (X, U, lambda):

* ``InternalForce``     f_int(X, U; theta) of the element groups, nodal [3N];
                        backward: fea_elem_adjoint -> d/dX, d/dU (= K lambda)
                        and d/dtheta, theta = per-group (w1, w2, kv).
* ``PressureForce``     the follower-pressure external force -p dV/dx, nodal;
                        backward: -p Hess(V) g for both X and U (fea_surface_pressure_hvp).
* ``StaticEquilibrium`` U*(X, theta) from the static Newton solver; backward:
                        the adjoint solve K(U*) lambda = g (FP64 PCG: the forward solver's
                        multigrid hierarchy when it has one, node-block Jacobi otherwise)
                        then -lambda . dR/dX and -lambda . dR/dtheta with
                        R = f_int - f_ext (pressure).
* ``compliance_gradient``  J = f_ext(U*, X) . U* and dJ/dX, dJ/dtheta.

theta of an element group is its kernel parameter vector (w1, w2, kv) of
W = w1 (I1bar - 3) + w2 (I2bar - 3) + kv/2 (J - 1)^2; ``material_gradient``
maps d/dtheta to the reference's material parameters.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from ._native import F64, call


def _elem_groups(plan):
    return [(gi, g) for gi, g in enumerate(plan.groups) if g.kind == 'elem']


def _check_loads(plan):
    for g in plan.groups:
        if g.kind == 'traction':
            raise NotImplementedError('shape sensitivities of traction loads (dead loads on the '
                                      'reference area) are not implemented')


def _theta_of(plan):
    return torch.tensor(np.stack([g.params for _, g in _elem_groups(plan)]), dtype=torch.float64)


def _set_theta(plan, theta):
    th = theta.detach().cpu().numpy().reshape(-1, 3)
    for k, (_, g) in enumerate(_elem_groups(plan)):
        g.params[:] = th[k]


def elem_vjp(plan, U, lam, want_x=True, want_u=False, want_theta=False):
    """d(lam . f_int)/dX, /dU (nodal [3N]) and /dtheta ([groups, 3]) at (plan.X, U)."""
    seg, dsts, (sx, su) = plan.nodal_staging()
    s = nat.stream_ptr()
    gx = gu = gth = None
    if want_x:
        sx.zero_()
    if want_u:
        su.zero_()
    parts = []
    for gi, g in _elem_groups(plan):
        part = None
        if want_theta:
            nb = nat.lib().fea_elem_adjoint_blocks(g.n)
            part = torch.zeros(max(1, 3 * nb), dtype=torch.float64, device=plan.device)
            parts.append((part, nb))
        if g.n == 0:
            continue
        call('fea_elem_adjoint', g.ntype, g.n, g.conn, plan.X, U, g.params, lam, dsts[gi],
             sx if want_x else None, su if want_u else None, part, s)
    if want_x:
        gx = torch.empty(3 * plan.n_nodes, dtype=torch.float64, device=plan.device)
        call('fea_gather_nodal', plan.n_nodes, seg, sx, gx, s)
    if want_u:
        gu = torch.empty(3 * plan.n_nodes, dtype=torch.float64, device=plan.device)
        call('fea_gather_nodal', plan.n_nodes, seg, su, gu, s)
    if want_theta:
        out = torch.zeros(len(parts), 3, dtype=torch.float64, device=plan.device)
        for k, (part, nb) in enumerate(parts):
            for j in range(3):
                if nb:
                    call('fea_reduce_sum', nb, part[j * nb:(j + 1) * nb], out[k, j:j + 1], s)
        gth = out
    return gx, gu, gth


def pressure_hvp(plan, U, lam, load_scale=1.0):
    """sum over pressure groups of p Hess(V)(X + U) lam, nodal [3N]."""
    seg, dsts, (sx, _) = plan.nodal_staging()
    s = nat.stream_ptr()
    sx.zero_()
    for gi, g in enumerate(plan.groups):
        if g.kind == 'pressure' and g.n:
            call('fea_surface_pressure_hvp', g.ntype, g.n, g.conn, plan.X, U,
                 float(g.load.pressure) * load_scale, lam, dsts[gi], sx, s)
    out = torch.empty(3 * plan.n_nodes, dtype=torch.float64, device=plan.device)
    call('fea_gather_nodal', plan.n_nodes, seg, sx, out, s)
    return out


def _nodal_residual(plan, U, kinds, load_scale=1.0):
    """Nodal [3N] residual rows of the given group kinds (every node, constrained too)."""
    seg, dsts, (sx, _) = plan.nodal_staging()
    s = nat.stream_ptr()
    sx.zero_()
    for gi, g in enumerate(plan.groups):
        if g.kind not in kinds or g.n == 0:
            continue
        if g.kind == 'elem':
            call('fea_elem_residual', g.ntype, F64, g.n, g.conn, plan.X, U, g.params, dsts[gi], sx, s)
        elif g.kind == 'pressure':
            call('fea_surface_pressure', g.ntype, F64, g.n, g.conn, plan.X, U,
                 float(g.load.pressure) * load_scale, None, None, dsts[gi], sx, s)
    out = torch.empty(3 * plan.n_nodes, dtype=torch.float64, device=plan.device)
    call('fea_gather_nodal', plan.n_nodes, seg, sx, out, s)
    return out


def _vec(t, plan):
    t = torch.as_tensor(t, dtype=torch.float64, device=plan.device).reshape(-1).contiguous()
    if t.numel() != 3 * plan.n_nodes:
        raise ValueError('nodal vectors have %d entries' % (3 * plan.n_nodes))
    return t


class InternalForce(torch.autograd.Function):
    """f_int(X, U; theta), nodal [3N], with the hand-written backward."""

    @staticmethod
    def forward(ctx, X, U, theta, plan):
        plan.set_coordinates(X.detach())
        _set_theta(plan, theta)
        Ud = _vec(U.detach(), plan)
        ctx.plan = plan
        ctx.save_for_backward(X.detach().clone(), Ud.clone(), theta.detach().clone())
        return _nodal_residual(plan, Ud, ('elem',))

    @staticmethod
    def backward(ctx, g):
        X, U, theta = ctx.saved_tensors
        plan = ctx.plan
        plan.set_coordinates(X)
        _set_theta(plan, theta)
        need = ctx.needs_input_grad
        gx, gu, gth = elem_vjp(plan, U, _vec(g, plan), want_x=need[0], want_u=need[1], want_theta=need[2])
        if gth is not None:
            gth = gth.to(theta.device)
        return gx, gu, gth, None


class PressureForce(torch.autograd.Function):
    """External follower-pressure force -p dV/dx (pressure.py:80-132), nodal [3N]."""

    @staticmethod
    def forward(ctx, X, U, plan, load_scale=1.0):
        plan.set_coordinates(X.detach())
        Ud = _vec(U.detach(), plan)
        ctx.plan, ctx.scale = plan, load_scale
        ctx.save_for_backward(X.detach().clone(), Ud.clone())
        return -_nodal_residual(plan, Ud, ('pressure',), load_scale)

    @staticmethod
    def backward(ctx, g):
        X, U = ctx.saved_tensors
        plan = ctx.plan
        plan.set_coordinates(X)
        h = -pressure_hvp(plan, U, _vec(g, plan), ctx.scale)
        return h, h.clone(), None, None


class StaticEquilibrium(torch.autograd.Function):
    """U*(X, theta): the static Newton equilibrium of the controller's model.

    Backward: K(U*) lambda = g on the free DOFs (the tangent is symmetric: the
    pressure part is the Hessian of p V), then
    dX = -(lambda . df_int/dX + p Hess(V) lambda), dtheta = -lambda . df_int/dtheta."""

    @staticmethod
    def forward(ctx, X, theta, fe, GC0=None, adjoint_tol=1e-24, adjoint_max_iter=20000):
        asm = fe.assembly
        plan = asm.plan
        _check_loads(plan)
        plan.set_coordinates(X.detach())
        _set_theta(plan, theta)
        if GC0 is None:
            GC = fe.solver.solve(torch.zeros_like(asm.GC))
        else:
            # warm start from a nearby equilibrium (a perturbed mesh or material): one Newton
            # solve at the full load instead of the solver's load increments from rest
            steps = fe.solver.load_steps
            fe.solver.load_steps = 1
            try:
                GC = fe.solver.solve(GC0.detach().clone())
            finally:
                fe.solver.load_steps = steps
        if isinstance(GC, bool):
            raise RuntimeError('static solve did not converge')
        U = plan.gc_to_u(GC).clone()
        ctx.fe, ctx.tol, ctx.max_iter = fe, adjoint_tol, adjoint_max_iter
        ctx.save_for_backward(X.detach().clone(), theta.detach().clone(), U)
        return U

    @staticmethod
    def backward(ctx, g):
        from .solver import PCG
        X, theta, U = ctx.saved_tensors
        fe = ctx.fe
        plan = fe.assembly.plan
        plan.set_coordinates(X)
        _set_theta(plan, theta)
        _, K = plan.tangent(U, 1.0, F64)
        # the forward solver's multigrid hierarchy preconditions the adjoint solve too when it
        # has one (its finest level smooths with this K); node-block Jacobi otherwise
        amg = getattr(fe.solver, 'amg', None) if getattr(fe.solver, 'preconditioner', None) == 'amg' else None
        pcg = PCG(plan.n_free, plan.nnz, plan.rowptr, plan.col, plan.diag_slot, F64, plan.device,
                  grp=plan.grp, n_groups=plan.spmv_groups, spmv_col=plan.spmv_col,
                  precond='amg' if amg is not None else 'block', amg=amg)
        pcg.set_matrix(K)
        rhs = plan.u_to_gc(_vec(g, plan)).contiguous()
        lam_free = pcg.solve(rhs, None, tol=ctx.tol, max_iter=ctx.max_iter)
        ctx.adjoint_iters = pcg.iterations
        lam = plan.gc_to_u(lam_free.contiguous()).clone()
        need = ctx.needs_input_grad
        gx = gth = None
        ex, _, eth = elem_vjp(plan, U, lam, want_x=need[0], want_u=False, want_theta=need[1])
        if need[0]:
            gx = -(ex + pressure_hvp(plan, U, lam))
        if need[1]:
            gth = -eth.to(theta.device)
        return gx, gth, None, None, None, None


def compliance(fe, X=None, theta=None, GC0=None):
    """J = f_ext(U*, X) . U* (pressure loads) as a differentiable torch scalar.  ``GC0``: start
    the equilibrium solve from this state at the full load (a nearby equilibrium, e.g. for
    finite differences) instead of ramping the load from rest."""
    plan = fe.assembly.plan
    if X is None:
        X = plan.X.reshape(-1).clone().requires_grad_(True)
    if theta is None:
        theta = _theta_of(plan).requires_grad_(True)
    U = StaticEquilibrium.apply(X, theta, fe, GC0)
    f = PressureForce.apply(X, U, plan)
    return torch.sum(f * U), X, theta


def compliance_gradient(fe):
    """(J, dJ/dX nodal [3N], dJ/dtheta [groups, 3]) at the controller's current mesh."""
    J, X, theta = compliance(fe)
    J.backward()
    return float(J.detach()), X.grad, theta.grad


def material_gradient(material, g_theta):
    """Map d/d(w1, w2, kv) of one group to its material's parameters:
    NeoHookean (mu, kappa): w1 = mu/2, kv = kappa;
    MooneyRivlin (C10, C01, D1): w1 = C10, w2 = C01, kv = 2/D1."""
    g = [float(v) for v in g_theta]
    name = type(material).__name__
    if name == 'NeoHookean':
        return {'mu': 0.5 * g[0], 'kappa': g[2]}
    if name == 'MooneyRivlin':
        D1 = float(torch.as_tensor(material.D1).reshape(-1)[0])
        return {'C10': g[0], 'C01': g[1], 'D1': -2.0 / D1 ** 2 * g[2]}
    raise NotImplementedError(name)
