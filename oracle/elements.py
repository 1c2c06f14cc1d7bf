"""Solid elements C3D4 / C3D8 / C3D10: FP64 numpy restatement of the reference.

Test infrastructure only (see oracle/__init__.py).  Layout conventions follow
the reference exactly so that flattened outputs can be compared entry by entry:

* dN/dX           [g, E, 3, nen]   (C3base.py:175, :208)
* w*detJ (signed) [g, E]           (C3base.py:206, :214)
* R_e flattened over [a, j, e]      (C3base.py:424-426, :431)
* K_e flattened over [a, j, b, k, e] (C3base.py:435-439)
"""
import numpy as np

from . import materials as _mat

# Monomial exponents of the reference's 3-D polynomial basis
# (C3base.py:40-60 docstring and :283-334): 1,g,h,r,gh,hr,rg,g2,h2,r2,
# g2h,h2g,h2r,r2h,r2g,g2r,ghr,g3,h3,r3.
MONO3 = np.array([
    (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1),
    (2, 0, 0), (0, 2, 0), (0, 0, 2), (2, 1, 0), (1, 2, 0), (0, 2, 1), (0, 1, 2),
    (1, 0, 2), (2, 0, 1), (1, 1, 1), (3, 0, 0), (0, 3, 0), (0, 0, 3)], dtype=np.int64)


def _c3d8_table():
    """brick.py:52-142 — N_a = (1+g g_a)(1+h h_a)(1+r r_a)/8, node order brick.py:13-22."""
    signs = np.array([(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1),
                      (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)], dtype=np.float64)
    t = np.zeros((8, 20))
    for a, (ga, ha, ra) in enumerate(signs):
        t[a, 0] = 1.0
        t[a, 1], t[a, 2], t[a, 3] = ga, ha, ra
        t[a, 4], t[a, 5], t[a, 6] = ga * ha, ha * ra, ra * ga
        t[a, 16] = ga * ha * ra
    return t * 0.125


SHAPE_TABLES = {
    # tetrahedral.py:30-33
    'C3D4': np.array([[1.0, -1.0, -1.0, -1.0], [0.0, 1.0, 0.0, 0.0],
                      [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0]]),
    # tetrahedral.py:94-105 (mid-node order 4(01) 5(12) 6(02) 7(03) 8(13) 9(23))
    'C3D10': np.array([[1., -3., -3., -3., 4., 4., 4., 2., 2., 2.],
                       [0., -1., 0., 0., 0., 0., 0., 2., 0., 0.],
                       [0., 0., -1., 0., 0., 0., 0., 0., 2., 0.],
                       [0., 0., 0., -1., 0., 0., 0., 0., 0., 2.],
                       [0., 4., 0., 0., -4., 0., -4., -4., 0., 0.],
                       [0., 0., 0., 0., 4., 0., 0., 0., 0., 0.],
                       [0., 0., 4., 0., -4., -4., 0., 0., -4., 0.],
                       [0., 0., 0., 4., 0., -4., -4., 0., 0., -4.],
                       [0., 0., 0., 0., 0., 0., 4., 0., 0., 0.],
                       [0., 0., 0., 0., 0., 4., 0., 0., 0., 0.]]),
    'C3D8': _c3d8_table(),
}


def gauss_rule(etype):
    """Gauss points [g,3] and weights [g] (tetrahedral.py:38-40, :107-116; brick.py:144-160)."""
    if etype == 'C3D4':
        return np.array([[0.25, 0.25, 0.25]]), np.array([1.0 / 6.0])
    if etype == 'C3D10':
        alpha, beta = 0.58541020, 0.13819660   # truncated constants, tetrahedral.py:112-113
        p = np.array([[beta, beta, beta], [alpha, beta, beta],
                      [beta, alpha, beta], [beta, beta, alpha]])
        return p, np.full(4, 1.0 / 24.0)
    if etype == 'C3D8':
        q = 1.0 / np.sqrt(3.0)
        p = np.array([[-q, -q, -q], [q, -q, -q], [q, q, -q], [-q, q, -q],
                      [-q, -q, q], [q, -q, q], [q, q, q], [-q, q, q]])
        return p, np.ones(8)
    raise ValueError(etype)


def shape_at(etype, pts):
    """N [g, nen] and dN/dxi [g, 3, nen] at local points (C3base.py:190-199, :211)."""
    tab = SHAPE_TABLES[etype]
    nm = tab.shape[1]
    ex = MONO3[:nm]
    pts = np.asarray(pts, dtype=np.float64)
    mono = np.prod(pts[:, None, :] ** ex[None, :, :], axis=2)          # [g, nm]
    dmono = np.zeros((pts.shape[0], 3, nm))
    for m in range(3):
        e2 = ex.copy()
        coef = e2[:, m].astype(np.float64)
        e2[:, m] = np.maximum(e2[:, m] - 1, 0)
        dmono[:, m, :] = coef[None, :] * np.prod(pts[:, None, :] ** e2[None, :, :], axis=2)
    N = mono @ tab.T                                                   # [g, nen]
    dN = np.einsum('gmb,ab->gma', dmono, tab)                          # [g, 3, nen]
    return N, dN


NEN = {'C3D4': 4, 'C3D8': 8, 'C3D10': 10}


def pre_load_gaussian(etype, nodes, conn):
    """Element_3D._pre_load_gaussian (C3base.py:155-216) for full-order elements.

    Returns dNdX [g,E,3,nen], w*detJ [g,E] (signed), N [g,E,nen].
    """
    pts, w = gauss_rule(etype)
    N, dN = shape_at(etype, pts)
    X = nodes[conn]                                                    # [E, nen, 3]
    # J[g,e,i,m] = sum_a X[e,a,i] dN_a/dxi_m   (C3base.py:199-202)
    J = np.einsum('eai,gma->geim', X, dN)
    detJ = np.linalg.det(J)
    invJ = np.linalg.inv(J)                                            # [g,e,m,i]
    dNdX = np.einsum('gemi,gma->geia', invJ, dN)                       # C3base.py:208
    wdet = detJ * w[:, None]                                           # C3base.py:214
    N0 = np.broadcast_to(N[:, None, :], (len(w), conn.shape[0], N.shape[1])).copy()
    return dNdX, wdet, N0


def deformation_gradient(dNdX, conn, U):
    """F[g,e,I,i] = delta + sum_a U[conn[e,a], I] dNdX[g,e,i,a]   (C3base.py:455-464)."""
    Ue = U[conn]                                                       # [E, nen, 3]
    F = np.einsum('geia,eaI->geIi', dNdX, Ue)
    F[..., 0, 0] += 1.0
    F[..., 1, 1] += 1.0
    F[..., 2, 2] += 1.0
    return F


def structural_force(dNdX, wdet, conn, U, material, onlyforce=False):
    """Element_3D.structural_Force (C3base.py:413-452) with rotation = identity.

    Returns (R_e flattened [a,j,e], K_e flattened [a,j,b,k,e] or None).
    """
    F = deformation_gradient(dNdX, conn, U)
    s, C = material.constitutive(F)                                    # hyperelastic.py:35-84
    R = np.einsum('geij,geia,ge->aje', s, dNdX, wdet)                  # C3base.py:424-426
    if onlyforce:
        return R.reshape(-1), None
    K = np.einsum('geijkl,gelb,geia,ge->ajbke', C, dNdX, dNdX, wdet)   # C3base.py:435-439
    return R.reshape(-1), K.reshape(-1)


def potential_energy(dNdX, wdet, conn, U, material):
    """Element_3D.potential_Energy (C3base.py:387-411)."""
    F = deformation_gradient(dNdX, conn, U)
    W = material.energy(F)
    return float(np.einsum('ge,ge->', W, wdet))


def mass_values(N, wdet, rho):
    """Consistent element mass expanded to [a,j,b,k,e] (C3base.py:361-374)."""
    Me = np.einsum('gea,geb,ge->abe', N, N, wdet * rho)
    nen, E = Me.shape[0], Me.shape[2]
    full = np.zeros((nen, 3, nen, 3, E))
    for d in range(3):
        full[:, d, :, d, :] = Me
    return full.reshape(-1)


def force_indices(conn):
    """_indices_force: DOF index of each R_e entry in [a,j,e] order (C3base.py:148-153)."""
    nen = conn.shape[1]
    idx = conn.T[:, None, :] * 3 + np.arange(3)[None, :, None]
    return idx.reshape(-1).astype(np.int64)


def coo_indices(conn):
    """Row/col DOF of each K_e entry in [a,j,b,k,e] order (C3base.py:114-122)."""
    nen, E = conn.shape[1], conn.shape[0]
    A = conn.T[:, None, None, None, :]
    B = conn.T[None, None, :, None, :]
    j = np.arange(3)[None, :, None, None, None]
    k = np.arange(3)[None, None, None, :, None]
    rows = np.broadcast_to(A * 3 + j, (nen, 3, nen, 3, E)).reshape(-1)
    cols = np.broadcast_to(B * 3 + k, (nen, 3, nen, 3, E)).reshape(-1)
    return rows.astype(np.int64), cols.astype(np.int64)


def group_pattern(conn):
    """Element-group coalesced COO pattern (C3base.py:114-145).

    Returns (_indices_matrix [2, nnz_grp], _index_matrix_coalesce [9 nen^2 E]):
    unique (row, col) pairs sorted by row*scaler+col, and for each element entry
    the position of its pair — the reference builds the same with argsort +
    unique_consecutive + inverse permutation.
    """
    rows, cols = coo_indices(conn)
    scaler = max(rows.max(), cols.max()) + 1
    key = rows * scaler + cols
    uniq, inv = np.unique(key, return_inverse=True)
    ind = np.stack([uniq // scaler, uniq % scaler]).astype(np.int64)
    return ind, inv.astype(np.int64)


def required_dofs(conn, n_nodes):
    """Element_3D.set_required_DoFs for full-order elements (C3base.py:499-504)."""
    m = np.zeros((n_nodes, 3), dtype=bool)
    m[np.unique(conn)] = True
    return m


def make_material(kind, *params):
    if kind == 'NH':
        return _mat.NeoHookean(*params)
    if kind == 'MR':
        return _mat.MooneyRivlin(*params)
    raise ValueError(kind)
