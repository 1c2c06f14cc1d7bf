"""Oracle checks for node-to-surface penalty contact (CPU)."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import contact as oc


def plate(n=6, tilt=0.0, seed=0):
    """n x n unit square of T3 facets at z ~ 0, normals +z, optional random z tilt."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing='ij')
    nodes = np.stack([X.ravel(), Y.ravel(), tilt * rng.uniform(-1, 1, X.size)], axis=1)
    nid = lambda i, j: i * (n + 1) + j
    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            tris += [(a, b, c), (a, c, d)]                     # counter-clockwise from +z
    return nodes, np.array(tris, dtype=np.int64)


def with_slaves(nodes, pts):
    x = np.concatenate([nodes, pts], axis=0)
    slave = np.arange(nodes.shape[0], x.shape[0])
    return x, slave


def test_pairs_match_kdtree():
    rng = np.random.default_rng(1)
    nodes, conn = plate(12, tilt=0.02)
    pts = np.concatenate([rng.uniform(0, 1, (300, 2)), rng.uniform(-0.1, 0.1, (300, 1))], axis=1)
    x, slave = with_slaves(nodes, pts)
    r = 0.13
    pairs = oc.candidate_pairs(slave, conn, x, r)
    c = oc.centroids(conn, x)
    d = np.linalg.norm(x[slave][:, None] - c[None], axis=2)
    near = np.abs(d - r) < 1e-9
    tree = cKDTree(c)
    got = {(int(s), int(f)) for s, f in pairs if not near[s, f]}
    ref = {(s, f) for s, fl in enumerate(tree.query_ball_point(x[slave], r)) for f in fl if not near[s, f]}
    assert got == ref and len(got) > 1000


def test_flat_plate_closed_form():
    nodes, conn = plate(2)
    p = np.array([[0.3, 0.2, -0.01]])
    x, slave = with_slaves(nodes, p)
    k = 1e3
    pairs = oc.candidate_pairs(slave, conn, x, 1.0)
    fc, chosen, e = oc.penalty_forces(slave, conn, x, pairs, k, 0.5)
    fc = fc.reshape(-1, 3)
    assert np.allclose(fc[slave[0]], [0.0, 0.0, k * 0.01], rtol=1e-13, atol=0)
    assert np.isclose(e, 0.5 * k * 1e-4, rtol=1e-13)
    tri = x[conn[chosen[0]]]
    # the facet holding (0.3, 0.2) in the lower-left cell, lower triangle (a, b, c)
    assert np.all(tri[:, :2].min(0) <= [0.3, 0.2]) and np.all(tri[:, :2].max(0) >= [0.3, 0.2])
    w = -fc[conn[chosen[0]], 2] / (k * 0.01)
    assert np.allclose(w.sum(), 1.0) and np.allclose(w @ tri[:, :2], [0.3, 0.2])


def test_outside_or_too_deep_gives_no_force():
    nodes, conn = plate(2)
    x, slave = with_slaves(nodes, np.array([[0.3, 0.2, 0.01], [1.3, 0.2, -0.01], [0.3, 0.2, -0.6]]))
    pairs = oc.candidate_pairs(slave, conn, x, 2.0)
    fc, chosen, e = oc.penalty_forces(slave, conn, x, pairs, 1e3, 0.5)
    assert np.all(fc == 0.0) and np.all(chosen == -1) and e == 0.0


def _penetrating(seed):
    rng = np.random.default_rng(seed)
    nodes, conn = plate(5, tilt=0.01, seed=seed)
    pts = np.concatenate([rng.uniform(0.05, 0.95, (40, 2)), rng.uniform(-0.03, -0.02, (40, 1))], axis=1)
    return with_slaves(nodes, pts), conn


@pytest.mark.parametrize('seed', [0, 1])
def test_force_is_minus_energy_gradient(seed):
    (x, slave), conn = _penetrating(seed)
    k, depth, r = 1e3, 0.2, 0.5
    pairs = oc.candidate_pairs(slave, conn, x, r)
    fc, chosen, _ = oc.penalty_forces(slave, conn, x, pairs, k, depth)
    assert np.all(chosen >= 0)
    h = 1e-7
    grad = np.zeros(x.size)
    for i in range(x.size):
        xp, xm = x.copy().reshape(-1), x.copy().reshape(-1)
        xp[i] += h
        xm[i] -= h
        ep = oc.penalty_forces(slave, conn, xp.reshape(-1, 3), pairs, k, depth)
        em = oc.penalty_forces(slave, conn, xm.reshape(-1, 3), pairs, k, depth)
        assert np.array_equal(ep[1], chosen) and np.array_equal(em[1], chosen)
        grad[i] = (ep[2] - em[2]) / (2 * h)
    assert np.abs(fc + grad).max() < 1e-6 * np.abs(fc).max()


def test_force_and_moment_balance():
    (x, slave), conn = _penetrating(3)
    pairs = oc.candidate_pairs(slave, conn, x, 0.5)
    fc, chosen, _ = oc.penalty_forces(slave, conn, x, pairs, 1e3, 0.2)
    f = fc.reshape(-1, 3)
    scale = np.abs(f).max()
    assert np.abs(f.sum(0)).max() < 1e-12 * scale * len(slave)
    assert np.abs(np.cross(x, f).sum(0)).max() < 1e-12 * scale * len(slave)


def test_tributary_areas_partition_the_surface():
    nodes, conn = plate(4, tilt=0.0)
    ids, area = oc.tributary_areas(conn, nodes)
    assert np.array_equal(ids, np.arange(nodes.shape[0]))
    assert np.isclose(area.sum(), 1.0, rtol=1e-14)
    # interior node of a uniform right-triangle grid: 6 adjacent facets of area h^2/2
    h = 0.25
    assert np.isclose(area[1 * 5 + 1], 6 * (h * h / 2) / 3, rtol=1e-14)
    assert np.isclose(area[0], 2 * (h * h / 2) / 3, rtol=1e-14)     # corner on the cut diagonal
    assert np.isclose(area[4], 1 * (h * h / 2) / 3, rtol=1e-14)     # corner in one facet


def test_area_scales_force_and_energy():
    (x, slave), conn = _penetrating(4)
    pairs = oc.candidate_pairs(slave, conn, x, 0.5)
    area = np.linspace(0.5, 2.0, len(slave))
    f1, c1, e1 = oc.penalty_forces(slave, conn, x, pairs, 1e3, 0.2)
    fa, ca, ea = oc.penalty_forces(slave, conn, x, pairs, 1e3, 0.2, area)
    assert np.array_equal(c1, ca)
    assert np.allclose(fa.reshape(-1, 3)[slave], f1.reshape(-1, 3)[slave] * area[:, None], rtol=1e-14, atol=0)
    fa2, _, ea2 = oc.penalty_forces(slave, conn, x, pairs, 2e3, 0.2, 0.5 * area)
    assert np.allclose(fa2, fa, rtol=1e-14, atol=1e-300) and np.isclose(ea2, ea, rtol=1e-14)


def test_valley_crease_has_no_dead_zone():
    """Two facets forming a valley along y (outward normals tilted towards each
    other, like the dent a slave presses into the master surface): a node just
    below the valley line projects outside both facets; the second tier still
    assigns it to one of them, with the force along that facet's normal."""
    s = -0.05                                           # valley depth under the rims
    nodes = np.array([[-1.0, 0.0, 0.0], [0.0, 0.0, s], [0.0, 1.0, s], [-1.0, 1.0, 0.0],
                      [1.0, 0.0, 0.0], [1.0, 1.0, 0.0]])
    conn = np.array([[0, 1, 2], [1, 4, 5]], dtype=np.int64)
    # make both normals point up: (b - a) x (c - a) with z > 0
    for f in range(2):
        a, b, c = nodes[conn[f]]
        if np.cross(b - a, c - a)[2] < 0:
            conn[f] = conn[f][[0, 2, 1]]
    p = np.array([[0.0, 0.5, s - 0.01]])
    x, slave = with_slaves(nodes, p)
    pairs = oc.candidate_pairs(slave, conn, x, 5.0)
    ws = [oc._facet_terms(x[slave[0]], *x[conn[f]])[2] for f in range(2)]
    assert all(w.min() < -oc.INSIDE_TOL for w in ws)   # outside both: the dead zone
    fc, chosen, e = oc.penalty_forces(slave, conn, x, pairs, 1e3, 0.5)
    assert chosen[0] in (0, 1) and e > 0
    g, n, _ = oc._facet_terms(x[slave[0]], *x[conn[chosen[0]]])
    assert np.allclose(fc.reshape(-1, 3)[slave[0]], -1e3 * g * n, rtol=1e-14)
    assert np.abs(fc.reshape(-1, 3).sum(0)).max() < 1e-12
