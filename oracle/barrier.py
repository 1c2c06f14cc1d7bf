"""The reference's Gauss-point-pair contact barrier (ContactSelf / Contact), restated (FP64, CPU).

Test infrastructure only (see oracle/__init__.py).

For T3 facets p, q (3 Gauss points each, (1/6,1/6), (2/3,1/6), (1/6,2/3), weight 1/6,
triangle.py:35-41) and every Gauss pair (g on p, G on q), with current Gauss positions y,
unit normals n (x_xi x x_eta normalised, basesurface.py:215-229) and reference-area weights
w = |X_xi x X_eta| / 6 (basesurface.py:150):

    dy = y_g - y_G,  dn = n_p - n_q,  D = dn . dy / 2,  M = n_p . n_q,  L = |dy|
    g  = c_g exp(k_g D)                                   (contact.py:431-432)
    f  = S(clamp((s_f - M) / (s_f - e_f)))                (contact.py:426-429)
    h  = S(clamp((h_t - L) / (r_h h_t)))                  (contact.py:435-438)
    phi = ratio w_g w_G g f h,   S(t) = t^3 (6 t^2 - 15 t + 10)

summed over the candidate facet pairs: facet centroids (mean of the Gauss points) within
r = search_ratio * h_t at the current configuration (scipy cKDTree.query_pairs,
contact.py:86-122).  Self contact (ContactSelf) weighs each Gauss pair by a ratio from
the *initial* configuration, S(clamp((-(n0_g - n0_G) . dy0/|dy0| - 0.8) / 0.4)), and drops
pairs whose ratios are all zero (contact.py:302-343); the two-surface Contact uses ratio 1.
The contact energy enters the total potential energy with a plus sign (the load returns
-sum phi, contact.py:386-387, and the assembly subtracts load potentials); its gradient and
Hessian enter R and K the same way (contact.py:702-705).

Derivatives here are this restatement's own (chain rule through D, M, L and the normal
map n(x)); they are pinned against central differences of the energy.
"""
import numpy as np
import scipy.spatial

XI = np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]])
NG = np.stack([1.0 - XI[:, 0] - XI[:, 1], XI[:, 0], XI[:, 1]], axis=1)    # N_a(xi_g) [g, a]
DXI = np.array([-1.0, 1.0, 0.0])                                           # dN_a/dxi
DETA = np.array([-1.0, 0.0, 1.0])                                          # dN_a/deta
WG = 1.0 / 6.0

DEFAULTS = dict(c_g=1e-5, k_g=20.0, h_t=1.5, r_h=0.5, s_f=-0.0, e_f=-0.8, ign_min=0.8, ign_max=1.2,
                search_ratio=2.0)


def smooth(t):
    t = np.clip(t, 0.0, 1.0)
    return t ** 3 * (6 * t ** 2 - 15 * t + 10)


def smooth_d(t):
    """(S'(t), S''(t)) with zero outside (0, 1), as contact.py:547-555."""
    inside = (t > 0.0) & (t < 1.0)
    d1 = np.where(inside, 30 * t ** 2 * (t - 1) ** 2, 0.0)
    d2 = np.where(inside, 60 * t * (t - 1) * (2 * t - 1), 0.0)
    return d1, d2


def facet_geometry(x, conn):
    """Gauss positions [F, 3, 3] (facet, gauss, xyz), normal vector N = x_xi x x_eta [F, 3]."""
    xe = x[conn]                                                          # [F, a, 3]
    y = np.einsum('ga,fai->fgi', NG, xe)
    u = np.einsum('a,fai->fi', DXI, xe)
    v = np.einsum('a,fai->fi', DETA, xe)
    return y, np.cross(u, v)


def centroids(x, conn):
    y, _ = facet_geometry(x, conn)
    return y.mean(axis=1)


def candidate_pairs(c1, r, c2=None):
    """ContactBase._filter_point_pairs (contact.py:86-122): (i, j) index pairs, i < j for one
    surface; for two surfaces i indexes c1 and j indexes c2."""
    if c2 is None:
        pts = c1
    else:
        pts = np.concatenate([c1, c2])
    pairs = scipy.spatial.cKDTree(pts).query_pairs(r, output_type='ndarray').reshape(-1, 2)
    pairs = np.sort(pairs, axis=1)
    if c2 is not None:
        n1 = c1.shape[0]
        pairs = pairs[(pairs[:, 0] < n1) & (pairs[:, 1] >= n1)]
        pairs[:, 1] -= n1
    return pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]


def self_ratio(X0, conn, pairs, prm=DEFAULTS):
    """ContactSelf ratio (contact.py:312-343) of each pair [P, g, G] from the initial geometry."""
    y, N = facet_geometry(X0, conn)
    n = N / np.linalg.norm(N, axis=1, keepdims=True)
    i, j = pairs[:, 0], pairs[:, 1]
    dy = y[i][:, :, None, :] - y[j][:, None, :, :]                        # [P, g, G, 3]
    dm = (n[i] - n[j])[:, None, None, :]
    dx = dy / np.linalg.norm(dy, axis=-1, keepdims=True)
    T = -(dm * dx).sum(-1)
    return smooth((T - prm['ign_min']) / (prm['ign_max'] - prm['ign_min']))


def _pair_setup(X0, x1, conn1, x2, conn2, pairs):
    y1, N1 = facet_geometry(x1, conn1)
    y2, N2 = facet_geometry(x2, conn2)
    _, A1 = facet_geometry(X0[0], conn1)
    _, A2 = facet_geometry(X0[1], conn2)
    w1 = np.linalg.norm(A1, axis=1) * WG                                  # |X_xi x X_eta| w_g
    w2 = np.linalg.norm(A2, axis=1) * WG
    i, j = pairs[:, 0], pairs[:, 1]
    return y1[i], N1[i], y2[j], N2[j], (w1[i] * w2[j])[:, None, None] * np.ones((1, 3, 3))


def barrier(X0, x1, conn1, x2, conn2, pairs, ratio, prm=DEFAULTS, order=2):
    """Energy per pair [P]; with order >= 1 the gradient [P, 18] over (nodes of p, nodes of q)
    x xyz, with order 2 also the Hessian [P, 18, 18].  X0 = (reference coords of surface 1's
    instance, of surface 2's), x1, x2 current coordinates."""
    yp, Np, yq, Nq, w = _pair_setup(X0, x1, conn1, x2, conn2, pairs)
    P = pairs.shape[0]
    lp, lq = np.linalg.norm(Np, axis=1), np.linalg.norm(Nq, axis=1)
    n_p, n_q = Np / lp[:, None], Nq / lq[:, None]
    dy = yp[:, :, None, :] - yq[:, None, :, :]                            # [P, g, G, 3]
    dn = (n_p - n_q)[:, None, None, :]
    D = 0.5 * (dn * dy).sum(-1)
    M = (n_p * n_q).sum(-1)[:, None, None] * np.ones((1, 3, 3))
    L = np.linalg.norm(dy, axis=-1)
    sf = (prm['s_f'] - prm['e_f'])
    tm = (prm['s_f'] - M) / sf
    th = (prm['h_t'] - L) / (prm['r_h'] * prm['h_t'])
    f, h = smooth(tm), smooth(th)
    g = prm['c_g'] * np.exp(prm['k_g'] * D)
    K = ratio * w
    phi = K * g * f * h
    E = phi.sum((1, 2))
    if order == 0:
        return E
    f1, f2 = smooth_d(tm)
    f1, f2 = -f1 / sf, f2 / sf ** 2                                       # df/dM, d2f/dM2
    h1, h2 = smooth_d(th)
    hs = prm['r_h'] * prm['h_t']
    h1, h2 = -h1 / hs, h2 / hs ** 2                                       # dh/dL, d2h/dL2
    g1, g2 = prm['k_g'] * g, prm['k_g'] ** 2 * g
    # per Gauss pair gradients in the 12 variables (y_g, y_G, n_p, n_q)
    Lsafe = np.where(L > 0, L, 1.0)
    dhat = dy / Lsafe[..., None]
    Z = np.zeros_like(dy)
    gD = np.concatenate([0.5 * dn * np.ones_like(dy), -0.5 * dn * np.ones_like(dy), 0.5 * dy, -0.5 * dy], -1)
    gM = np.concatenate([Z, Z, n_q[:, None, None, :] * np.ones_like(dy), n_p[:, None, None, :] * np.ones_like(dy)], -1)
    gL = np.concatenate([dhat, -dhat, Z, Z], -1)
    cD, cM, cL = K * g1 * f * h, K * g * f1 * h, K * g * f * h1
    grad12 = cD[..., None] * gD + cM[..., None] * gM + cL[..., None] * gL  # [P, g, G, 12]
    # chain to nodal DOFs: y_g = sum_a N_a(xi_g) x_a ; n = n(x) through N = x_xi x x_eta
    Jn_p = _normal_jacobian(x1[conn1[pairs[:, 0]]], n_p, lp)               # [P, 3, 9]
    Jn_q = _normal_jacobian(x2[conn2[pairs[:, 1]]], n_q, lq)
    J = np.zeros((P, 3, 3, 12, 18))
    I3 = np.eye(3)
    for gi in range(3):
        for a in range(3):
            J[:, gi, :, 0:3, 3 * a:3 * a + 3] += NG[gi, a] * I3
            J[:, :, gi, 3:6, 9 + 3 * a:9 + 3 * a + 3] += NG[gi, a] * I3
    J[:, :, :, 6:9, 0:9] = Jn_p[:, None, None]
    J[:, :, :, 9:12, 9:18] = Jn_q[:, None, None]
    grad = np.einsum('pgGk,pgGkx->px', grad12, J)
    if order == 1:
        return E, grad
    # Hessian in the 12 variables
    H12 = (K * g2 * f * h)[..., None, None] * np.einsum('...i,...j->...ij', gD, gD)
    H12 += (K * g * f2 * h)[..., None, None] * np.einsum('...i,...j->...ij', gM, gM)
    H12 += (K * g * f * h2)[..., None, None] * np.einsum('...i,...j->...ij', gL, gL)
    for c, a, b in ((K * g1 * f1 * h, gD, gM), (K * g1 * f * h1, gD, gL), (K * g * f1 * h1, gM, gL)):
        H12 += c[..., None, None] * (np.einsum('...i,...j->...ij', a, b) + np.einsum('...i,...j->...ij', b, a))
    HD = np.zeros((12, 12))                                               # d2D / dvar2
    for r0, c0, s in ((0, 6, 0.5), (0, 9, -0.5), (3, 6, -0.5), (3, 9, 0.5)):
        HD[r0:r0 + 3, c0:c0 + 3] = s * I3
        HD[c0:c0 + 3, r0:r0 + 3] = s * I3
    HM = np.zeros((12, 12))
    HM[6:9, 9:12] = I3
    HM[9:12, 6:9] = I3
    HLb = (I3[None, None, None] - np.einsum('...i,...j->...ij', dhat, dhat)) / Lsafe[..., None, None]
    HL = np.zeros(dy.shape[:-1] + (12, 12))
    HL[..., 0:3, 0:3] = HLb
    HL[..., 3:6, 3:6] = HLb
    HL[..., 0:3, 3:6] = -HLb
    HL[..., 3:6, 0:3] = -HLb
    H12 += cD[..., None, None] * HD + cM[..., None, None] * HM + cL[..., None, None] * HL
    H = np.einsum('pgGkx,pgGkl,pgGly->pxy', J, H12, J)
    # second derivatives of the normals, contracted with the total gradient in n
    bp = grad12[..., 6:9].sum((1, 2))
    bq = grad12[..., 9:12].sum((1, 2))
    H[:, 0:9, 0:9] += _normal_second(x1[conn1[pairs[:, 0]]], n_p, lp, bp)
    H[:, 9:18, 9:18] += _normal_second(x2[conn2[pairs[:, 1]]], n_q, lq, bq)
    return E, grad, H


def _cross_mat(v):
    """[v]x with [v]x w = v x w, v [..., 3]."""
    z = np.zeros(v.shape[:-1])
    return np.stack([np.stack([z, -v[..., 2], v[..., 1]], -1), np.stack([v[..., 2], z, -v[..., 0]], -1),
                     np.stack([-v[..., 1], v[..., 0], z], -1)], -2)


def _dN_dx(xe):
    """dN/dx_a [P, 3(N comp), 9]: N = u x v, u = x_xi, v = x_eta; dN = du x v + u x dv."""
    u = np.einsum('a,pai->pi', DXI, xe)
    v = np.einsum('a,pai->pi', DETA, xe)
    out = np.zeros((xe.shape[0], 3, 9))
    for a in range(3):
        out[:, :, 3 * a:3 * a + 3] = -DXI[a] * _cross_mat(v) + DETA[a] * _cross_mat(u)
    return out


def _normal_jacobian(xe, n, l):
    Pm = (np.eye(3)[None] - np.einsum('pi,pj->pij', n, n)) / l[:, None, None]
    return np.einsum('pij,pjx->pix', Pm, _dN_dx(xe))


def _normal_second(xe, n, l, b):
    """sum_i b_i d2 n_i / dx dx [P, 9, 9]."""
    dN = _dN_dx(xe)
    bn = (b * n).sum(-1)
    I3 = np.eye(3)
    B2 = (-(np.einsum('pk,pj->pjk', b, n) + np.einsum('pj,pk->pjk', b, n) + bn[:, None, None] * I3[None])
          + 3 * bn[:, None, None] * np.einsum('pj,pk->pjk', n, n)) / (l ** 2)[:, None, None]
    out = np.einsum('pjx,pjk,pky->pxy', dN, B2, dN)
    c = np.einsum('pij,pj->pi', (I3[None] - np.einsum('pi,pj->pij', n, n)) / l[:, None, None], b)
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1.0
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1.0
    coef = np.outer(DXI, DETA) - np.outer(DETA, DXI)                      # [a, b]
    blk = np.einsum('ab,gAB,pg->paAbB', coef, eps, c)                     # [P, a, alpha, b, beta]
    return out + blk.reshape(-1, 9, 9)
