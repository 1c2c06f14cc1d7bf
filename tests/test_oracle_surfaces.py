"""Pin the oracle's surface restatement (oracle/surfaces.py) with closed forms.

The oracle restates the reference's facet tables and Gauss rules
(FEA/assemble/elements/C3/surfaces/triangle.py:3-99, quadrilateral.py:4-57),
the reference-area Jacobian (basesurface.py:113-150), the follower-pressure
volume functional and its derivatives (FEA/assemble/loads/pressure.py:80-159)
and the dead-load traction.  Nothing here retypes those formulas: each check
compares them with a value the mathematics fixes independently.

* Tables: N_a(node_b) = delta_ab at the facet's own nodes (this fixes every
  coefficient of a table given its node order), derivatives against central
  differences of N, reproduction of linear fields.
* Gauss rules: exact integrals of monomials over the reference triangle
  (a! b! / (a+b+2)!) and square, up to the degree each rule integrates.
* Closed box surfaces (T3, T6 with curved-parameter midside nodes, Q4 with
  non-parallelogram facets) off the origin, nodes jittered inside each face:
  V = 1/3 sum_g w det[x, x_xi, x_eta] equals the box volume (divergence
  theorem; the integrand is a polynomial the rule integrates exactly on flat
  facets), sum w detJ0 equals the surface area, the traction resultant equals
  t * area and its moment equals (first moment of area) x t.
* Pressure force -p dV/dx: zero net force on a closed surface, zero net
  moment on any surface (V is invariant under rotations about the origin),
  sum over nodes = -p/3 * (vector area) on an open surface, central
  differences of V for the force and of the force for the stiffness.
"""
import math

import numpy as np
import pytest

from oracle import surfaces as osurf

ORIGIN = np.array([0.7, -1.3, 2.1])
SIZE = np.array([1.0, 0.8, 1.2])

NODE_COORDS = {
    'T3': [[0, 0], [1, 0], [0, 1]],
    'T6': [[0, 0], [1, 0], [0, 1], [0.5, 0], [0.5, 0.5], [0, 0.5]],
    'Q4': [[-1, -1], [1, -1], [1, 1], [-1, 1]],
}


@pytest.mark.parametrize('stype', ['T3', 'T6', 'Q4'])
def test_shape_tables_kronecker_partition_derivatives(stype):
    xi = np.array(NODE_COORDS[stype], dtype=float)
    N, _ = osurf.surf_shape_at(stype, xi)
    assert np.abs(N - np.eye(len(xi))).max() < 1e-15
    rng = np.random.default_rng(1)
    pts = rng.uniform(0.05, 0.45, size=(9, 2)) if stype != 'Q4' else rng.uniform(-0.9, 0.9, size=(9, 2))
    N, dN = osurf.surf_shape_at(stype, pts)
    assert np.abs(N.sum(1) - 1.0).max() < 1e-14
    assert np.abs(dN.sum(2)).max() < 1e-13
    # isoparametric reproduction of the parametric coordinates themselves
    assert np.abs(N @ xi - pts).max() < 1e-14
    h = 1e-6
    for m in range(2):
        e = np.zeros(2)
        e[m] = h
        fd = (osurf.surf_shape_at(stype, pts + e)[0] - osurf.surf_shape_at(stype, pts - e)[0]) / (2 * h)
        assert np.abs(fd - dN[:, m]).max() < 1e-9


def _exact_triangle(a, b):
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


def _exact_square(a, b):
    one = lambda k: 0.0 if k % 2 else 2.0 / (k + 1)
    return one(a) * one(b)


@pytest.mark.parametrize('stype', ['T3', 'T6', 'Q4'])
def test_gauss_rule_exactness(stype):
    pts, w = osurf.surf_gauss(stype)
    if stype == 'Q4':
        degs = [(a, b) for a in range(4) for b in range(4)]
        exact = _exact_square
    else:
        degs = [(a, b) for a in range(3) for b in range(3) if a + b <= 2]
        exact = _exact_triangle
    for a, b in degs:
        q = float((w * pts[:, 0] ** a * pts[:, 1] ** b).sum())
        assert abs(q - exact(a, b)) < 1e-15, (stype, a, b, q, exact(a, b))


def box_surface(stype, n=3, jitter=0.12, seed=0, origin=ORIGIN, size=SIZE):
    """Closed outward-oriented surface of the box origin + [0, size], n x n cells per face.

    Nodes on a lattice of step size / (2n) (T6 uses the half-steps for its midside
    nodes); nodes strictly inside a face are moved within the face plane by up to
    `jitter` lattice steps, so every facet stays flat and the facets still tile
    each face, while the T6 and Q4 maps become non-affine.
    """
    L = 2 * n
    key_id, pos, conn = {}, [], []
    rng = np.random.default_rng(seed)

    def node(I):
        I = tuple(int(v) for v in I)
        if I not in key_id:
            x = origin + size * np.array(I, dtype=float) / L
            on = [v in (0, L) for v in I]
            if sum(on) == 1:                              # face interior: move in-plane
                d = rng.uniform(-jitter, jitter, 3) * size / L
                d[on.index(True)] = 0.0
                x = x + d
            key_id[I] = len(pos)
            pos.append(x)
        return key_id[I]

    for k in range(3):
        for side in (0, 1):
            u, v = ((k + 1) % 3, (k + 2) % 3) if side else ((k + 2) % 3, (k + 1) % 3)

            def lat(i, j):                                # lattice point, i along u, j along v (half-steps)
                I = np.zeros(3, dtype=int)
                I[k] = side * L
                I[u], I[v] = i, j
                return node(I)

            for i in range(0, L, 2):
                for j in range(0, L, 2):
                    c00, c10, c11, c01 = lat(i, j), lat(i + 2, j), lat(i + 2, j + 2), lat(i, j + 2)
                    if stype == 'Q4':
                        conn.append([c00, c10, c11, c01])
                    elif stype == 'T3':
                        conn += [[c00, c10, c11], [c00, c11, c01]]
                    else:
                        conn.append([c00, c10, c11, lat(i + 1, j), lat(i + 2, j + 1), lat(i + 1, j + 1)])
                        conn.append([c00, c11, c01, lat(i + 1, j + 1), lat(i + 1, j + 2), lat(i, j + 1)])
    return np.array(pos), np.array(conn, dtype=np.int64)


def _box_area_and_moment(origin=ORIGIN, size=SIZE):
    """Surface area and first moment of area (integral of x dA) of the box surface."""
    area, moment = 0.0, np.zeros(3)
    for k in range(3):
        fa = size[(k + 1) % 3] * size[(k + 2) % 3]
        c = origin + 0.5 * size
        for side in (0, 1):
            cf = c.copy()
            cf[k] = origin[k] + side * size[k]
            area += fa
            moment += fa * cf
    return area, moment


@pytest.mark.parametrize('stype', ['T3', 'T6', 'Q4'])
def test_closed_box_volume_area_traction(stype):
    pos, conn = box_surface(stype)
    V = osurf.pressure_volume(stype, pos, conn)
    assert abs(V - np.prod(SIZE)) < 1e-13 * np.prod(SIZE), V
    area, moment = _box_area_and_moment()
    _, _, w = osurf.surf_shape(stype)
    A0 = float((osurf.reference_area_det(stype, pos, conn) * w[:, None]).sum())
    assert abs(A0 - area) < 1e-13 * area, (A0, area)
    t = np.array([0.3, -1.1, 0.45])
    f = osurf.traction_force(stype, pos, conn, t)
    assert np.abs(f.sum(0) - t * area).max() < 1e-13 * area
    if stype != 'T6':
        # sum_a x_a (x) f_a = integral x dA (x) t: the rule integrates x * detJ0 exactly
        # for affine triangles and bilinear quadrilaterals (not for curved-parameter T6)
        M = np.einsum('ai,aj->ij', pos, f)
        assert np.abs(M - np.outer(moment, t)).max() < 1e-12 * np.abs(np.outer(moment, t)).max()


def test_closed_box_t6_affine_moment():
    pos, conn = box_surface('T6', jitter=0.0)
    area, moment = _box_area_and_moment()
    t = np.array([0.3, -1.1, 0.45])
    f = osurf.traction_force('T6', pos, conn, t)
    M = np.einsum('ai,aj->ij', pos, f)
    assert np.abs(M - np.outer(moment, t)).max() < 1e-12 * np.abs(np.outer(moment, t)).max()


def _nodal(conn, vals, n):
    F = np.zeros(3 * n)
    np.add.at(F, osurf.pressure_force_idx(conn), vals)
    return F.reshape(n, 3)


def _deform(pos, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    return pos + amp * rng.standard_normal(pos.shape)


@pytest.mark.parametrize('stype', ['T3', 'T6', 'Q4'])
def test_closed_pressure_balance_and_derivatives(stype):
    pos, conn = box_surface(stype, n=2)
    x = _deform(pos, 7)
    p = 0.37
    Fv, Kv = osurf.pressure_force_stiffness(stype, x, conn, p)
    F = _nodal(conn, Fv, len(x))
    scale = np.abs(F).max()
    assert np.abs(F.sum(0)).max() < 1e-13 * scale * len(x)          # closed surface: no net force
    assert np.abs(np.cross(x, F).sum(0)).max() < 1e-13 * scale * len(x)  # rotation invariance of V
    # force = -p dV/dx by central differences of V
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(4):
        d = rng.standard_normal(x.shape)
        fd = (osurf.pressure_volume(stype, x + h * d, conn) - osurf.pressure_volume(stype, x - h * d, conn)) / (2 * h)
        assert abs(-p * fd - (F * d).sum()) < 1e-8 * np.abs(F).sum()
    # stiffness = dF/dx by central differences of the force
    idx = osurf.pressure_stiff_idx(conn)
    n3 = 3 * len(x)
    K = np.zeros((n3, n3))
    np.add.at(K, (idx[0], idx[1]), Kv)
    assert np.abs(K - K.T).max() < 1e-12 * np.abs(K).max()
    for _ in range(3):
        d = rng.standard_normal(x.shape)
        Fp = _nodal(conn, osurf.pressure_force_stiffness(stype, x + h * d, conn, p, onlyforce=True)[0], len(x))
        Fm = _nodal(conn, osurf.pressure_force_stiffness(stype, x - h * d, conn, p, onlyforce=True)[0], len(x))
        fd = ((Fp - Fm) / (2 * h)).reshape(-1)
        assert np.abs(fd - K @ d.reshape(-1)).max() < 1e-7 * np.abs(K).max()


@pytest.mark.parametrize('stype', ['T3', 'T6', 'Q4'])
def test_open_surface_net_force_is_a_third_of_vector_area(stype):
    """Sum_a dV/dx_a = d V / d(translation) = 1/3 sum_g w (x_xi x x_eta): a third of the
    vector area, whatever the surface (here one face of the box, deformed)."""
    pos, conn = box_surface(stype, n=2)
    top = np.array([np.all(np.isclose(pos[c, 2], ORIGIN[2] + SIZE[2])) for c in conn])
    sub = conn[top]
    area = SIZE[0] * SIZE[1]
    p = 0.5
    F = _nodal(sub, osurf.pressure_force_stiffness(stype, pos, sub, p, onlyforce=True)[0], len(pos))
    assert np.abs(F.sum(0) - (-p / 3.0) * area * np.array([0.0, 0.0, 1.0])).max() < 1e-14
    # rotation invariance holds on an open surface as well
    x = _deform(pos, 5)
    F = _nodal(sub, osurf.pressure_force_stiffness(stype, x, sub, p, onlyforce=True)[0], len(x))
    assert np.abs(np.cross(x, F).sum(0)).max() < 1e-13 * np.abs(F).max() * len(x)
