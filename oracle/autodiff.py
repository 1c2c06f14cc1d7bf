"""Autograd gradients of the element and pressure forces (test infrastructure only).

The reference obtains sensitivities by running torch autograd through its
assembly (FEA/assemble/elements/C3/C3base.py:387-452 are plain torch ops on
the nodal coordinates and displacements).  This module restates that path in
torch float64: the strain energy of a group
    Pi(X, U) = sum_g sum_e  w_g detJ0 W(F),   F = I + U^T dN/dX,
(shape tables and Gauss rules of oracle.elements; W the hyperelastic energy
of oracle.materials, which its tests pin against the reference's stress by
complex-step differentiation), the internal force f_int = dPi/dU by autograd,
and vector-Jacobian products of f_int by a second autograd pass.  Follower
pressure uses the enclosed-volume functional V(x) = sum_g w_g/3 det[x, x_xi,
x_eta] of pressure.py:95-105 (oracle.surfaces) with f = p dV/dx.
"""
import numpy as np
import torch

from . import elements as el
from . import surfaces as osurf


def _t(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float64))


def strain_energy(etype, conn, X, U, w1, w2, kv):
    """Pi for one element group; X, U [N,3] tensors; w1, w2, kv scalars or 0-d tensors.

    W = w1 (J^-2/3 I1 - 3) + w2 (J^-4/3 I2 - 3) + kv/2 (J - 1)^2
    (NeoHookean: w1 = mu/2, w2 = 0, kv = kappa; MooneyRivlin: C10, C01, 2/D1)."""
    pts, w = el.gauss_rule(etype)
    _, dN = el.shape_at(etype, pts)
    dN, w = _t(dN), _t(w)
    conn = torch.as_tensor(np.asarray(conn, dtype=np.int64))
    Xe, Ue = X[conn], U[conn]                                          # [E, nen, 3]
    Jm = torch.einsum('eai,gma->geim', Xe, dN)
    detJ = torch.linalg.det(Jm)
    invJ = torch.linalg.inv(Jm)
    dNdX = torch.einsum('gemi,gma->geia', invJ, dN)
    F = torch.eye(3, dtype=X.dtype) + torch.einsum('eaI,geia->geIi', Ue, dNdX)
    C = F.transpose(-1, -2) @ F
    I1 = torch.diagonal(C, dim1=-2, dim2=-1).sum(-1)
    I2 = 0.5 * (I1 * I1 - torch.diagonal(C @ C, dim1=-2, dim2=-1).sum(-1))
    J = torch.linalg.det(F)
    W = (w1 * (J ** (-2.0 / 3.0) * I1 - 3.0) + w2 * (J ** (-4.0 / 3.0) * I2 - 3.0)
         + 0.5 * kv * (J - 1.0) ** 2)
    return torch.sum(w[:, None] * detJ * W)


def internal_force(etype, conn, X, U, w1, w2, kv, create_graph=False):
    """f_int [N,3] = dPi/dU."""
    if not U.requires_grad:
        U = U.detach().requires_grad_(True)
    Pi = strain_energy(etype, conn, X, U, w1, w2, kv)
    return torch.autograd.grad(Pi, U, create_graph=create_graph)[0]


def internal_force_vjp(etype, conn, X, U, params, lam):
    """Gradients of lam . f_int(X, U; w1, w2, kv) w.r.t. X, U and (w1, w2, kv)."""
    X = _t(X).requires_grad_(True)
    U = _t(U).requires_grad_(True)
    th = [torch.tensor(float(p), dtype=torch.float64, requires_grad=True) for p in params]
    f = internal_force(etype, conn, X, U, th[0], th[1], th[2], create_graph=True)
    phi = torch.sum(f * _t(lam))
    gX, gU, *gth = torch.autograd.grad(phi, [X, U] + th)
    return gX.numpy(), gU.numpy(), np.array([g.item() for g in gth])


def pressure_volume(stype, conn, x):
    """V(x) = sum over facets and Gauss points of w/3 det[x, x_xi, x_eta] (pressure.py:95-105)."""
    N, dN, w = osurf.surf_shape(stype)                                 # N [g,a], dN [g,m,a]
    N, dN, w = _t(N), _t(dN), _t(w)
    conn = torch.as_tensor(np.asarray(conn, dtype=np.int64))
    xe = x[conn]                                                       # [F, nen, 3]
    r = torch.einsum('ga,fai->gfi', N, xe)
    r1 = torch.einsum('ga,fai->gfi', dN[:, 0], xe)
    r2 = torch.einsum('ga,fai->gfi', dN[:, 1], xe)
    det = torch.einsum('gfi,gfi->gf', r, torch.cross(r1, r2, dim=-1))
    return torch.sum(w[:, None] * det / 3.0)


def pressure_hvp(stype, conn, x, p, lam):
    """p * Hess(V)(x) . lam  [N,3] (the pressure part of the tangent times lam)."""
    x = _t(x).requires_grad_(True)
    V = pressure_volume(stype, conn, x)
    g = torch.autograd.grad(V, x, create_graph=True)[0]
    hv = torch.autograd.grad(torch.sum(g * _t(lam)), x)[0]
    return p * hv.numpy()
