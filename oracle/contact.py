"""Oracle for node-to-surface penalty contact (test infrastructure only).

The contact law on the accelerated path is new: the reference's contact
(FEA/assemble/loads/contact.py:11-1217) is a smooth barrier between Gauss
points of two surfaces with a scipy cKDTree broad phase (contact.py:86-122,
:385-401), whose force-only branch used by explicit dynamics does not run
(contact.py:658-659 and :1140-1142 return inside an unfinished branch).  This
module restates the replacement law in numpy, in the operation order the
kernels use where the result is discrete (candidate sets):

  candidate pair (s, f)  iff  |x_s - c_f|^2 <= r^2,  c_f = ((x_a + x_b) + x_c) / 3,
                              |d|^2 = (dx*dx + dy*dy) + dz*dz   (IEEE, no FMA)
  penetration             g = (x_s - x_a) . n,  n = unit((x_b - x_a) x (x_c - x_a)),
                          -depth_max < g < 0;  v = smallest barycentric coordinate
                          of the projection on the facet plane
  chosen facet            tier 0 (v >= -1e-10): largest g; only if tier 0 is
                          empty, tier 1 (-0.1 <= v < -1e-10): largest v, which
                          covers the dead zone under concave creases
                          (ties: smallest facet id)
  forces                  slave  F = -k A_s g n,  facet vertex i  -w_i F  (w barycentric,
                          A_s the slave's tributary area, 1 when not given)
  energy                  k A_s g^2 / 2

Pinned by: scipy cKDTree (pair sets), a closed-form flat-plate case, force =
-grad(energy) by finite differences, and balance of force and moment.
"""
import numpy as np

INSIDE_TOL = 1e-10      # barycentric tolerance of the inside test (edges / vertices count as inside)
EXTEND_TOL = 0.1        # second tier: projections at most this far outside (barycentric)


def positions(X, u):
    return np.asarray(X, dtype=np.float64).reshape(-1, 3) + np.asarray(u, dtype=np.float64).reshape(-1, 3)


def centroids(conn, x):
    t = x[conn]
    return ((t[:, 0] + t[:, 1]) + t[:, 2]) / 3.0


def candidate_pairs(slave, conn, x, radius):
    """All (slave position s, facet f) with |x_s - c_f|^2 <= radius^2, sorted; brute force."""
    c = centroids(conn, x)
    p = x[slave]
    d = p[:, None, :] - c[None, :, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    s, f = np.nonzero(d2 <= radius * radius)
    return np.stack([s, f], axis=1).astype(np.int64)


def _facet_terms(p, a, b, c):
    e1, e2, ap = b - a, c - a, p - a
    n = np.cross(e1, e2)
    nn = np.sqrt(n @ n)
    if not nn > 0.0:
        return None
    n = n / nn
    g = ap @ n
    q = ap - g * n
    d00, d01, d11 = e1 @ e1, e1 @ e2, e2 @ e2
    d20, d21 = q @ e1, q @ e2
    den = d00 * d11 - d01 * d01
    xi = (d11 * d20 - d01 * d21) / den
    eta = (d00 * d21 - d01 * d20) / den
    return g, n, np.array([1.0 - xi - eta, xi, eta])


def penalty_forces(slave, conn, x, pairs, penalty, depth_max, area=None):
    """Nodal contact forces fc[3N] (external), chosen facet per slave (-1: none), energy."""
    N = x.shape[0]
    fc = np.zeros((N, 3))
    chosen = -np.ones(len(slave), dtype=np.int64)
    energy = 0.0
    by_s = {}
    for s, f in pairs:
        by_s.setdefault(int(s), []).append(int(f))
    for s, fl in by_s.items():
        p = x[slave[s]]
        best = None
        for f in sorted(fl):
            t = _facet_terms(p, *x[conn[f]])
            if t is None:
                continue
            g, n, w = t
            if not (g < 0.0 and g > -depth_max):
                continue
            v = min(w[0], min(w[1], w[2]))
            if not v >= -EXTEND_TOL:
                continue
            tier = 0 if v >= -INSIDE_TOL else 1
            key = g if tier == 0 else v
            if best is None or tier < best[0] or (tier == best[0] and key > best[1]):
                best = (tier, key, f, g, n, w)
        if best is None:
            continue
        _, _, f, g, n, w = best
        chosen[s] = f
        ks = penalty * (1.0 if area is None else area[s])
        F = -ks * g * n
        fc[slave[s]] += F
        for i in range(3):
            fc[conn[f][i]] -= w[i] * F
        energy += 0.5 * ks * g * g
    return fc.reshape(-1), chosen, energy


def tributary_areas(tri, X):
    """Per-node third of the adjacent T3 facet areas: (nodes ascending, areas)."""
    P = np.asarray(X, dtype=np.float64)[tri]
    a = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
    nodes = np.unique(tri)
    area = np.zeros(nodes.shape[0])
    for t, at in zip(tri, a):
        for v in t:
            area[np.searchsorted(nodes, v)] += at / 3.0
    return nodes, area


def contact_force(slave, conn, X, u, radius, penalty, depth_max, area=None):
    x = positions(X, u)
    pairs = candidate_pairs(slave, conn, x, radius)
    return penalty_forces(slave, conn, x, pairs, penalty, depth_max, area)[0]


def contact_energy(slave, conn, X, u, radius, penalty, depth_max, area=None):
    x = positions(X, u)
    pairs = candidate_pairs(slave, conn, x, radius)
    return penalty_forces(slave, conn, x, pairs, penalty, depth_max, area)[2]
