"""Surface facets T3 / T6 / Q4, follower pressure and dead-load traction.

Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np

# 2-D monomials (basesurface.py:62-70): 1, g, h, gh, g2, h2, g2h, gh2
MONO2 = np.array([(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1), (1, 2)], dtype=np.int64)

SURF_TABLES = {
    'T3': np.array([[1.0, -1.0, -1.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]),        # triangle.py:22-26
    'T6': np.array([[1.0, -3.0, -3.0, 4.0, 2.0, 2.0],                             # triangle.py:67-73
                    [0.0, -1.0, 0.0, 0.0, 2.0, 0.0],
                    [0.0, 0.0, -1.0, 0.0, 0.0, 2.0],
                    [0.0, 4.0, 0.0, -4.0, -4.0, 0.0],
                    [0.0, 0.0, 0.0, 4.0, 0.0, 0.0],
                    [0.0, 0.0, 4.0, -4.0, 0.0, -4.0]]),
    'Q4': np.array([[0.25, -0.25, -0.25, 0.25],                                   # quadrilateral.py:22-27
                    [0.25, 0.25, -0.25, -0.25],
                    [0.25, 0.25, 0.25, 0.25],
                    [0.25, -0.25, 0.25, -0.25]]),
}

SURF_NEN = {'T3': 3, 'T6': 6, 'Q4': 4}


def surf_gauss(stype):
    """Gauss points and weights (triangle.py:37-40, :83-86; quadrilateral.py:35-40)."""
    if stype in ('T3', 'T6'):
        return (np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]]),
                np.array([1 / 6, 1 / 6, 1 / 6]))
    if stype == 'Q4':
        q = 1.0 / np.sqrt(3.0)
        return np.array([[-q, -q], [q, -q], [q, q], [-q, q]]), np.ones(4)
    raise ValueError(stype)


def surf_shape_at(stype, pts):
    """N [g,a] and dN [g,m,a] of a facet type at parametric points pts [g,2].

    N_a = sum_b table[a,b] * monomial_b(pts) over the monomials of
    basesurface.py:122-132; dN from the derivative of each monomial
    (basesurface.py:152-170 restated on exponents)."""
    tab = SURF_TABLES[stype]
    nm = tab.shape[1]
    ex = MONO2[:nm]
    pts = np.asarray(pts, dtype=np.float64)
    mono = np.prod(pts[:, None, :] ** ex[None], axis=2)
    dmono = np.zeros((pts.shape[0], 2, nm))
    for m in range(2):
        e2 = ex.copy()
        coef = e2[:, m].astype(np.float64)
        e2[:, m] = np.maximum(e2[:, m] - 1, 0)
        dmono[:, m] = coef[None] * np.prod(pts[:, None, :] ** e2[None], axis=2)
    N = mono @ tab.T
    dN = np.einsum('gmb,ab->gma', dmono, tab)
    return N, dN


def surf_shape(stype):
    """shape_function_gaussian = [N [g,a], dN [g,m,a]] (basesurface.py:122-141)."""
    pts, w = surf_gauss(stype)
    N, dN = surf_shape_at(stype, pts)
    return N, dN, w


def reference_area_det(stype, nodes, sconn):
    """det_Jacobian = sqrt(det(J^T J)) on the undeformed surface (basesurface.py:144-150)."""
    _, dN, _ = surf_shape(stype)
    J = np.einsum('gma,eai->geim', dN, nodes[sconn])
    return np.sqrt(np.linalg.det(np.einsum('geim,gein->gemn', J, J)))


def _ra(stype, pos, sconn):
    """r_added_gaussian[g,e,i,m] = [x, x_xi, x_eta] (pressure.py:95-105)."""
    N, dN, w = surf_shape(stype)
    shp = np.concatenate([N[:, None, :], dN], axis=1)                  # [g,3,a]
    ra = np.einsum('gma,eai->geim', shp, pos[sconn])
    return ra, shp, w


def pressure_volume(stype, pos, sconn):
    """V = 1/3 sum_g w_g det[x, x_xi, x_eta] (pressure.py:149-157)."""
    ra, _, w = _ra(stype, pos, sconn)
    return float((w[:, None] * np.linalg.det(ra)).sum() / 3.0)


def pressure_force_stiffness(stype, pos, sconn, pressure, onlyforce=False):
    """Pressure.get_stiffness (pressure.py:80-132).

    Returns (F values [e,i,a] flattened = -p dV/dx, K values [e,i,a,j,b]
    flattened = -p d2V/dx2 or None).  Indices: force_idx / stiff_idx below.
    """
    ra, shp, w = _ra(stype, pos, sconn)
    det = np.linalg.det(ra)
    inv = np.linalg.inv(ra)                                            # [g,e,m,i]
    det_radra = det[..., None, None] * np.swapaxes(inv, -1, -2)        # [g,e,i,m]  :112
    Vdra = w[:, None, None, None] * det_radra / 3.0                    # :114
    Vdot = np.einsum('gma,geim->eia', shp, Vdra)                       # :118
    if onlyforce:
        return -pressure * Vdot.reshape(-1), None
    part_I = np.einsum('geim,genj->geimjn', det_radra, inv)            # :123
    part_II = np.transpose(part_I, (0, 1, 2, 5, 4, 3))                 # :124
    Vdra2 = w[:, None, None, None, None, None] * (part_I - part_II) / 3.0
    Vdot2 = np.einsum('gma,gnb,geimjn->eiajb', shp, shp, Vdra2)        # :127
    return -pressure * Vdot.reshape(-1), -pressure * Vdot2.reshape(-1)


def pressure_force_idx(sconn):
    """_Vdot_indices layout [e,i,a] (pressure.py:59)."""
    return (sconn[:, None, :] * 3 + np.arange(3)[None, :, None]).reshape(-1).astype(np.int64)


def pressure_stiff_idx(sconn):
    """_Vdot_2_indices layout [e,i,a,j,b] (pressure.py:62-72)."""
    E, n = sconn.shape
    A = sconn[:, None, :, None, None]
    B = sconn[:, None, None, None, :]
    i = np.arange(3)[None, :, None, None, None]
    j = np.arange(3)[None, None, None, :, None]
    rows = np.broadcast_to(A * 3 + i, (E, 3, n, 3, n)).reshape(-1)
    cols = np.broadcast_to(B * 3 + j, (E, 3, n, 3, n)).reshape(-1)
    return np.stack([rows, cols]).astype(np.int64)


def traction_force(stype, nodes, sconn, traction):
    """Dead-load uniform traction (new; no reference counterpart).

    f[e,a,i] = sum_g N_a(g) t_i w_g detJ0(g,e) with the reference-area Gauss
    rule of basesurface.py:144-150.  Returns nodal forces [n_nodes, 3].
    """
    N, _, w = surf_shape(stype)
    dA = reference_area_det(stype, nodes, sconn) * w[:, None]          # [g,e]
    t = np.asarray(traction, dtype=np.float64)
    fe = np.einsum('ga,ge,i->eai', N, dA, t)
    f = np.zeros_like(nodes)
    np.add.at(f, sconn.reshape(-1), fe.reshape(-1, 3))
    return f
