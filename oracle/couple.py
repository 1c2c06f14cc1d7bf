"""Couple constraints and reference points, restated from the reference (FP64, CPU).

Test infrastructure only (see oracle/__init__.py).

A model (see oracle/assembly.py) with reference points carries two more keys:

    'rps':     [{'node': (x, y, z), 'fixed': 6 bools (Boundary_Condition_RP, optional),
                 'force': (fx, fy, fz) (Concentrate_Force, optional),
                 'moment': (mx, my, mz) (Moment, optional)}]      in name-sorted order
    'couples': [{'nodes': int array of coupled node ids, 'rp': index into 'rps'}]

Each reference point owns 6 generalized coordinates (v, z): a translation v and a rotation
vector z (reference_points.py:5-24).  A Couple slaves its nodes to the point:
u_b = v + R(z) y_b - y_b with y_b = X_b - X_rp (couple.py:36-43), R(z) the Rodrigues
rotation of couple.py:272-292.  The reduced generalized coordinates are the free node DOFs
(in node order) followed by the free reference-point DOFs (assembly.py:97-185: instance
blocks first, then reference points).

The reduced residual and tangent are the chain rule of the total energy through
u(GC):  R_red = T^T R_full  and  K_red = T^T K_full T + sum_b R_b . d2u_b/dz2,
T = du/dGC.  This is what Couple.modify_R_K (couple.py:66-184) assembles entry by entry
(its (reference point, row) entries are the transposes of its (row, reference point)
entries, i.e. it assumes K_full symmetric, which the hyperelastic, pressure and traction
terms are).  Reference-point loads enter as constant generalized forces with potential
F.v + M.z (concentrate_force.py:31-43, moment.py:31-43).
"""
import numpy as np
import scipy.sparse as sp

from . import assembly as asm

_EPS = np.zeros((3, 3, 3))
_EPS[0, 1, 2] = _EPS[1, 2, 0] = _EPS[2, 0, 1] = 1.0
_EPS[0, 2, 1] = _EPS[2, 1, 0] = _EPS[1, 0, 2] = -1.0


def rotate(z, y):
    """Couple._rotation3d (couple.py:272-292): R(z) y for y [n,3], theta = |z| + 1e-20."""
    y = np.asarray(y, dtype=np.float64).reshape(-1, 3)
    z = np.asarray(z, dtype=np.float64)
    theta = np.linalg.norm(z) + 1e-20
    w = z / theta
    c, s = np.cos(theta), np.sin(theta)
    return y * c + np.cross(w[None, :], y) * s + w[None, :] * (y @ w)[:, None] * (1.0 - c)


def rotation_jacobians(z, y):
    """Couple._calculate_Ydotz (couple.py:186-270): dY/dz [n,3,3] and d2Y/dz2 [n,3,3,3] of
    Y(z) = R(z) y - y, by the chain rule through the axis w = z / theta and the angle
    theta = |z| + 1e-20 (indices [node, component i, z_p(, z_q)])."""
    y = np.asarray(y, dtype=np.float64).reshape(-1, 3)
    z = np.asarray(z, dtype=np.float64)
    theta = np.linalg.norm(z) + 1e-20
    w = z / theta
    c, s = np.cos(theta), np.sin(theta)
    wy = y @ w                                         # [n]
    wxy = np.cross(w[None, :], y)                      # [n,3]
    I = np.eye(3)
    # derivatives of Y(w, theta) = y cos + (w x y) sin + w (w.y)(1 - cos) - y
    d_t = -y * s + wxy * c + w[None, :] * wy[:, None] * s                       # dY_i/dtheta
    d_tt = -y * c - wxy * s + w[None, :] * wy[:, None] * c                      # d2Y_i/dtheta2
    eps_y = np.einsum('ilk,nk->nil', _EPS, y)                                   # d(w x y)_i/dw_l
    wy_term = I[None] * wy[:, None, None] + w[None, :, None] * y[:, None, :]    # d(w_i (w.y))/dw_l
    d_w = eps_y * s + wy_term * (1.0 - c)                                       # [n,i,l]
    d_wt = eps_y * c + wy_term * s                                              # d2Y_i/dw_l dtheta
    d_ww = (I[None, :, :, None] * y[:, None, None, :] +
            I[None, :, None, :] * y[:, None, :, None]) * (1.0 - c)              # [n,i,l,m]
    # derivatives of (w, theta) with respect to z
    w_z = I / theta - np.outer(z, z) / theta ** 3                               # dw_l/dz_p
    t_z = w                                                                     # dtheta/dz_p
    t_zz = w_z                                                                  # d2theta/dz_p dz_q
    w_zz = 3.0 * np.einsum('l,p,q->lpq', z, z, z) / theta ** 5
    zt3 = z / theta ** 3
    for i in range(3):
        w_zz[i, i, :] -= zt3
        w_zz[i, :, i] -= zt3
        w_zz[:, i, i] -= zt3
    Yd = np.einsum('nil,lp->nip', d_w, w_z) + np.einsum('ni,p->nip', d_t, t_z)
    Ydd = (np.einsum('nilm,lp,mq->nipq', d_ww, w_z, w_z)
           + np.einsum('nil,lp,q->nipq', d_wt, w_z, t_z)
           + np.einsum('nil,lq,p->nipq', d_wt, w_z, t_z)
           + np.einsum('ni,p,q->nipq', d_tt, t_z, t_z)
           + np.einsum('ni,pq->nipq', d_t, t_zz)
           + np.einsum('nil,lpq->nipq', d_w, w_zz))
    return Yd, Ydd


def coupled_nodes(model):
    out = np.zeros(model['nodes'].shape[0], dtype=bool)
    for c in model.get('couples', []):
        out[np.asarray(c['nodes'])] = True
    return out


def layout(model):
    """(node free mask [3N], reference-point free mask [6 n_rp]) after elements -> loads ->
    constraints -> boundarys (assembly.py:157-168; couple.py:57-64; boundary_condition_rp.py:29-31)."""
    node_free = asm.free_mask(model).reshape(-1, 3).copy()
    node_free[coupled_nodes(model)] = False
    rps = model.get('rps', [])
    rp_free = np.zeros((len(rps), 6), dtype=bool)
    for k, rp in enumerate(rps):
        if rp.get('force') is not None:
            rp_free[k, :3] = True
        if rp.get('moment') is not None:
            rp_free[k, 3:] = True
    for c in model.get('couples', []):
        rp_free[c['rp'], :] = True
    for k, rp in enumerate(rps):
        if rp.get('fixed') is not None:
            rp_free[k, np.asarray(rp['fixed'], dtype=bool)] = False
    return node_free.reshape(-1), rp_free.reshape(-1)


def n_gc(model):
    nf, rf = layout(model)
    return int(nf.sum() + rf.sum())


def state(model, GC):
    """_GC2RGC with Couple.modify_RGC (assembly.py:470-492, couple.py:36-43):
    (nodal displacement [N,3], reference-point coordinates [n_rp,6])."""
    nf, rf = layout(model)
    n1 = int(nf.sum())
    U = np.zeros(nf.shape[0])
    U[nf] = GC[:n1]
    rp = np.zeros(rf.shape[0])
    rp[rf] = GC[n1:]
    U = U.reshape(-1, 3)
    rp = rp.reshape(-1, 6)
    X = model['nodes']
    for c in model.get('couples', []):
        nodes = np.asarray(c['nodes'])
        Xr = np.asarray(model['rps'][c['rp']]['node'], dtype=np.float64)
        y = X[nodes] - Xr
        v, z = rp[c['rp'], :3], rp[c['rp'], 3:]
        U[nodes] = v[None, :] + rotate(z, y) - y
    return U, rp


def _load_vector(model):
    f = np.zeros(6 * len(model.get('rps', [])))
    for k, rp in enumerate(model.get('rps', [])):
        if rp.get('force') is not None:
            f[6 * k:6 * k + 3] = rp['force']
        if rp.get('moment') is not None:
            f[6 * k + 3:6 * k + 6] = rp['moment']
    return f


def jacobian(model, rp):
    """T = du/dGC [3N, n_gc] and the per-couple second derivatives (nodes, Ydd)."""
    nf, rf = layout(model)
    N3 = nf.shape[0]
    n1 = int(nf.sum())
    red_rp = np.full(rf.shape[0], -1)
    red_rp[rf] = n1 + np.arange(int(rf.sum()))
    rows, cols, vals = [np.nonzero(nf)[0]], [np.arange(n1)], [np.ones(n1)]
    second = []
    X = model['nodes']
    for c in model.get('couples', []):
        k = c['rp']
        nodes = np.asarray(c['nodes'])
        y = X[nodes] - np.asarray(model['rps'][k]['node'], dtype=np.float64)
        Yd, Ydd = rotation_jacobians(rp[k, 3:], y)
        for i in range(3):
            for d in range(6):
                col = red_rp[6 * k + d]
                if col < 0 or (d < 3 and d != i):
                    continue
                val = np.ones(nodes.shape[0]) if d < 3 else Yd[:, i, d - 3]
                rows.append(3 * nodes + i)
                cols.append(np.full(nodes.shape[0], col))
                vals.append(val)
        second.append((nodes, k, Ydd))
    T = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N3, n1 + int(rf.sum()))).tocsr()
    return T, second, red_rp


def assemble_reduced(model, GC):
    """assemble_Stiffness_Matrix with constraints (assembly.py:283-387, couple.py:66-184):
    R_red [n_gc] and K_red (csr, sorted) at the generalized coordinates GC."""
    U, rp = state(model, GC)
    R_full, coo = asm.assemble_full(model, U)
    K_full = sp.coo_matrix((coo[2], (coo[0], coo[1])), shape=(R_full.shape[0],) * 2).tocsr()
    T, second, red_rp = jacobian(model, rp)
    f = _load_vector(model)
    R = T.T @ R_full
    sel = red_rp >= 0
    R[red_rp[sel]] -= f[sel]
    K = (T.T @ K_full @ T).tolil()
    for nodes, k, Ydd in second:
        G = np.einsum('ni,nipq->pq', R_full.reshape(-1, 3)[nodes], Ydd)
        for p in range(3):
            for q in range(3):
                a, b = red_rp[6 * k + 3 + p], red_rp[6 * k + 3 + q]
                if a >= 0 and b >= 0:
                    K[a, b] += G[p, q]
    K = K.tocsr()
    K.sum_duplicates()
    K.sort_indices()
    return R, K


def assemble_force(model, GC):
    U, rp = state(model, GC)
    R_full, _ = asm.assemble_full(model, U, onlyforce=True)
    T, _, red_rp = jacobian(model, rp)
    R = T.T @ R_full
    f = _load_vector(model)
    sel = red_rp >= 0
    R[red_rp[sel]] -= f[sel]
    return R


def total_energy(model, GC):
    """_total_Potential_Energy (assembly.py:389-415) with the reference-point load potentials."""
    U, rp = state(model, GC)
    return asm.total_energy(model, U) - float(_load_vector(model) @ rp.reshape(-1))
