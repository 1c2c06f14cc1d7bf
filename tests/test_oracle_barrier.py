"""Pins for the oracle's restatement of the reference's Gauss-pair contact barrier (CPU).

Gradient and Hessian against central differences of the energy (so a dropped term, a sign
or a transposed index fails), invariance under rigid motions and under swapping the two
facets of a pair, the pair set against brute force, and the self-contact ratio's closed
values (0 for coplanar neighbours, 1 for facing parallel plates).
"""
import numpy as np
import pytest

from oracle import barrier as ob
from tests.helpers import rel_err


def plates(n=3, gap=0.6, tilt=0.0, jitter=0.05, seed=0):
    """Two triangulated n x n plates: the lower at z = 0 (normals +z), the upper at z = gap
    (normals -z), the upper tilted by `tilt` radians about the x axis."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 2.0, n + 1)
    P = np.array([[x, y, 0.0] for y in g for x in g])
    tri = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = j * (n + 1) + i, j * (n + 1) + i + 1, (j + 1) * (n + 1) + i, (j + 1) * (n + 1) + i + 1
            tri += [[a, b, d], [a, d, c]]
    tri = np.array(tri)
    lower = P + jitter * rng.uniform(-1, 1, P.shape) * np.array([1, 1, 0.2])
    c, s = np.cos(tilt), np.sin(tilt)
    R = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    upper = (P - [1, 1, 0]) @ R.T + [1, 1, gap] + jitter * rng.uniform(-1, 1, P.shape) * np.array([1, 1, 0.2])
    X = np.concatenate([lower, upper])
    conn = np.concatenate([tri, tri[:, [0, 2, 1]] + P.shape[0]])      # upper facets face down
    return X, conn


def _all(X, conn, x, pairs, ratio, order=2):
    return ob.barrier((X, X), x, conn, x, conn, pairs, ratio, order=order)


@pytest.mark.parametrize('tilt,gap', [(0.0, 0.6), (0.9, 0.9), (1.2, 1.0)])
def test_barrier_derivatives_match_central_differences(tilt, gap):
    """Initial configuration: facing plates (ratio 1); current: the upper plate tilted so
    n_p . n_q runs through the angle factor's transition (-0.8, 0) and the Gauss distances
    through the distance factor's (0.75, 1.5)."""
    X, conn = plates(gap=gap)
    x, _ = plates(tilt=tilt, gap=gap, seed=3)
    x = x + 0.03 * np.random.default_rng(1).standard_normal(x.shape)
    pairs = ob.candidate_pairs(ob.centroids(x, conn), 3.0)
    ratio = ob.self_ratio(X, conn, pairs)
    keep = ratio.sum((1, 2)) > 0
    pairs, ratio = pairs[keep], ratio[keep]
    E, grad, H = _all(X, conn, x, pairs, ratio)
    assert E.sum() > 0
    h = 1e-6
    checked = 0
    for p in np.argsort(-np.abs(grad).max(1))[:8]:
        nodes = np.concatenate([conn[pairs[p, 0]], conn[pairs[p, 1]]])
        fd_g = np.zeros(18)
        fd_H = np.zeros((18, 18))
        for k in range(18):
            xp, xm = x.copy(), x.copy()
            xp[nodes[k // 3], k % 3] += h
            xm[nodes[k // 3], k % 3] -= h
            Ep = _all(X, conn, xp, pairs[p:p + 1], ratio[p:p + 1], order=1)
            Em = _all(X, conn, xm, pairs[p:p + 1], ratio[p:p + 1], order=1)
            fd_g[k] = (Ep[0][0] - Em[0][0]) / (2 * h)
            fd_H[:, k] = (Ep[1][0] - Em[1][0]) / (2 * h)
        assert rel_err(grad[p], fd_g) < 1e-7
        assert rel_err(H[p], fd_H) < 1e-6
        assert np.abs(H[p] - H[p].T).max() < 1e-10 * np.abs(H[p]).max()
        checked += 1
    assert checked == 8


def test_barrier_is_invariant_under_rigid_motion_and_swap():
    X, conn = plates(gap=0.9)
    x, _ = plates(tilt=1.0, gap=0.9, seed=4)
    pairs = ob.candidate_pairs(ob.centroids(x, conn), 3.0)
    ratio = ob.self_ratio(X, conn, pairs)
    E = _all(X, conn, x, pairs, ratio, order=0)
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    E2 = _all(X, conn, x @ Q.T + [3.0, -1.0, 2.0], pairs, ratio, order=0)
    assert rel_err(E2, E) < 1e-10
    sw = pairs[:, ::-1].copy()
    E3 = _all(X, conn, x, sw, ob.self_ratio(X, conn, sw), order=0)
    assert rel_err(E3, E) < 1e-12


def test_pairs_match_brute_force():
    X, conn = plates(n=4, gap=0.7)
    c = ob.centroids(X, conn)
    pairs = ob.candidate_pairs(c, 1.3)
    d = np.linalg.norm(c[:, None] - c[None], axis=2)
    i, j = np.nonzero(np.triu(d <= 1.3, k=1))
    assert np.array_equal(pairs, np.stack([i, j], 1))


def test_self_ratio_closed_values():
    X, conn = plates(n=2, gap=0.5, jitter=0.0)
    F = conn.shape[0] // 2
    pairs = np.array([[0, 1], [0, F]])              # coplanar neighbours; lower facing upper
    r = ob.self_ratio(X, conn, pairs)
    assert np.all(r[0] == 0.0)                      # dn = 0: T = 0 < 0.8
    # flat facing facets: dn = 2 e_z, so T = 2 gap / |dy| for each Gauss pair
    y, _ = ob.facet_geometry(X, conn)
    dy = y[0][:, None, :] - y[F][None, :, :]
    T = 2 * 0.5 / np.linalg.norm(dy, axis=-1)
    t = np.clip((T - 0.8) / 0.4, 0, 1)
    assert np.abs(r[1] - t ** 3 * (6 * t ** 2 - 15 * t + 10)).max() < 1e-15
    assert r[1].max() == 1.0
