"""Per-group kernel calls behind the reference-compatible element / load API.

These serve code written against the reference's per-group methods
(Element_3D.structural_Force, potential_Energy, Pressure.get_stiffness, ...);
the Assembly and the solvers use the fused DevicePlan path instead.
"""
import numpy as np
import torch

from . import _native as nat
from ._native import F32, F64, call, ptr
from .device import DevicePlan, Group


def require_identity(R):
    if R is None:
        return
    R = torch.as_tensor(R, dtype=torch.float64).cpu()
    if not torch.allclose(R, torch.eye(3, dtype=torch.float64)):
        raise NotImplementedError('instance rotations are not supported on the accelerated path')


def as_u(RGC):
    return torch.as_tensor(RGC).to('cuda', torch.float64).contiguous().reshape(-1)


def _conn32(elems):
    return torch.as_tensor(elems).to('cuda', torch.int32).contiguous()


def elem_prepare(etype, ngp, elems, X):
    E = elems.shape[0]
    dN = torch.empty((ngp, E, 3, etype), dtype=torch.float64, device='cuda')
    w = torch.empty((ngp, E), dtype=torch.float64, device='cuda')
    N = torch.empty((ngp, E, etype), dtype=torch.float64, device='cuda')
    call('fea_elem_prepare', etype, E, _conn32(elems), X, dN, w, N, nat.stream_ptr())
    return dN, w, N


def group_plan(elem):
    """Plan over the part's nodes with every DOF free and this single group:
    its CSR is the reference's coalesced group pattern (C3base.py:114-145)."""
    X = torch.as_tensor(elem.part.nodes).detach().cpu().numpy()
    g = Group('elem', elem.etype, elem.conn_np(), params=elem.kernel_params())
    return DevicePlan(X, np.ones(X.size, dtype=bool), [g])


def coalesce_map(elem, plan):
    """Index of each [a,j,b,k,e] entry in the coalesced pattern (C3base.py:129-141)."""
    coo = plan.csr_indices_coo()
    n3 = plan.n_nodes * 3
    keys = coo[0] * n3 + coo[1]
    c = elem._elems.to('cuda')
    nen = elem.etype
    A = c.t().reshape(nen, 1, 1, 1, -1)
    B = c.t().reshape(1, 1, nen, 1, -1)
    j = torch.arange(3, device='cuda').reshape(1, 3, 1, 1, 1)
    k = torch.arange(3, device='cuda').reshape(1, 1, 1, 3, 1)
    q = ((A * 3 + j) * n3 + (B * 3 + k)).reshape(-1)
    return torch.searchsorted(keys, q)


def elem_residual_per_node(elem, U, dtype=F64):
    """R_e in the reference's [a, j, e] layout (C3base.py:424-431)."""
    E, nen = elem._elems.shape[0], elem.etype
    ident = torch.arange(E * nen, dtype=torch.int32, device='cuda')
    out = torch.empty(E * nen * 3, dtype=torch.float64, device='cuda')
    X = elem._nodes().reshape(-1)
    call('fea_elem_residual', nen, dtype, E, _conn32(elem._elems), X, U,
         elem.kernel_params(), ident, out, nat.stream_ptr())
    return out.view(E, nen, 3).permute(1, 2, 0).reshape(-1)


def elem_energy(elem, U):
    E = elem._elems.shape[0]
    nb = nat.lib().fea_elem_energy_blocks(E)
    part = torch.zeros(max(1, nb), dtype=torch.float64, device='cuda')
    out = torch.zeros(1, dtype=torch.float64, device='cuda')
    s = nat.stream_ptr()
    call('fea_elem_energy', elem.etype, E, _conn32(elem._elems), elem._nodes().reshape(-1), U,
         elem.kernel_params(), part, s)
    call('fea_reduce_sum', nb, part, out, s)
    return out[0]


def surface_pressure_per_facet(stype, conn, X, U, pressure, onlyforce):
    """Pressure.get_stiffness values in the reference layouts (pressure.py:118-132):
    F values [e, i, a] = -p dV/dx and K values [e, i, a, j, b] = -p d2V/dx2."""
    E = conn.shape[0]
    n = stype
    c32 = _conn32(conn)
    dr = torch.arange(E * n, dtype=torch.int32, device='cuda')
    rs = torch.empty(E * n * 3, dtype=torch.float64, device='cuda')
    dk = ks = None
    if not onlyforce:
        dk = torch.arange(E * n * n, dtype=torch.int32, device='cuda')
        ks = torch.empty(E * n * n * 9, dtype=torch.float64, device='cuda')
    call('fea_surface_pressure', n, F64, E, c32, X, U, float(-pressure),
         dk, ks, dr, rs, nat.stream_ptr())
    Fv = rs.view(E, n, 3).permute(0, 2, 1).reshape(-1)
    if onlyforce:
        return Fv, None
    Kv = ks.view(E, n, n, 3, 3).permute(0, 3, 1, 4, 2).reshape(-1)
    return Fv, Kv


def surface_volume(stype, conn, X, U):
    E = conn.shape[0]
    nb = nat.lib().fea_surface_blocks(E)
    part = torch.zeros(max(1, nb), dtype=torch.float64, device='cuda')
    out = torch.zeros(1, dtype=torch.float64, device='cuda')
    s = nat.stream_ptr()
    call('fea_surface_volume', stype, E, _conn32(conn), X, U, part, s)
    call('fea_reduce_sum', nb, part, out, s)
    return out[0]


def surface_traction_force(stype, conn, X, traction):
    """Nodal dead-load forces f [e, i, a] of each facet (positive external force)."""
    E = conn.shape[0]
    n = stype
    dr = torch.arange(E * n, dtype=torch.int32, device='cuda')
    rs = torch.empty(E * n * 3, dtype=torch.float64, device='cuda')
    t = np.asarray(traction, dtype=np.float64)
    U0 = torch.zeros_like(X)
    call('fea_surface_traction', n, E, _conn32(conn), X, U0, t, -1.0,
         dr, rs, None, nat.stream_ptr())
    return rs.view(E, n, 3).permute(0, 2, 1).reshape(-1)
