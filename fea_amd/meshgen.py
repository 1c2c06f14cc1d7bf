"""Structured mesh generators for the benchmark / test workloads.

* ``kuhn_tet_block``  — box split into nx*ny*nz hex cells, each cut into 6
  C3D4 tets sharing the cell's (0,0,0)-(1,1,1) diagonal (Kuhn split).
* ``hex_block``       — box of C3D8 bricks in the node order of the reference's
  C3D8 (brick.py:13-22).
* ``tet10_from_tet4`` — C3D10 by adding mid-edge nodes in the reference order
  4(01) 5(12) 6(02) 7(03) 8(13) 9(23) (tetrahedral.py:80-81).
* ``block_with_void`` — 2x1x1 block around a perturbed spherical void, tets
  from cubed-sphere hexahedral patches (conforming, C3D4 or C3D10).

Every generator returns a dict with ``nodes`` [N,3] float64, ``conn`` [E,nen]
int64, ``face_sets`` {'x0','x1','y0','y1','z0','z1': [(elem_idx ndarray,
face_id)]} in the reference's surface-set format (part.py:166-178) and
``node_sets`` of the same names.
"""
import itertools

import numpy as np

# local face definitions (node positions) of the reference elements
TET_FACES = {0: (0, 2, 1), 1: (0, 1, 3), 2: (1, 2, 3), 3: (0, 3, 2)}               # tetrahedral.py:14-19
HEX_FACES = {0: (0, 3, 2, 1), 1: (4, 5, 6, 7), 2: (0, 1, 5, 4), 3: (1, 2, 6, 5),
             4: (2, 3, 7, 6), 5: (0, 4, 7, 3)}                                       # brick.py:30-35


def _grid_nodes(nx, ny, nz, size, origin):
    xs = origin[0] + np.linspace(0.0, size[0], nx + 1)
    ys = origin[1] + np.linspace(0.0, size[1], ny + 1)
    zs = origin[2] + np.linspace(0.0, size[2], nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing='ij')
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def _nid(i, j, k, nx, ny):
    return i + (nx + 1) * (j + (ny + 1) * k)


def _box_sets(nodes, conn, faces, size, origin):
    lo = np.asarray(origin, dtype=np.float64)
    hi = lo + np.asarray(size, dtype=np.float64)
    tol = 1e-9 * max(size)
    node_sets, face_sets = {}, {}
    for ax, name in enumerate('xyz'):
        for side, val in (('0', lo[ax]), ('1', hi[ax])):
            on = np.abs(nodes[:, ax] - val) <= tol
            node_sets[name + side] = np.nonzero(on)[0]
            entries = []
            for fid, loc in faces.items():
                hit = np.all(on[conn[:, list(loc)]], axis=1)
                idx = np.nonzero(hit)[0]
                if idx.size:
                    entries.append((idx, fid))
            face_sets[name + side] = entries
    return node_sets, face_sets


def _jitter_interior(nodes, nx, ny, nz, size, origin, jitter, seed):
    if jitter <= 0:
        return nodes
    h = np.array([size[0] / nx, size[1] / ny, size[2] / nz])
    lo = np.asarray(origin, dtype=np.float64)
    hi = lo + np.asarray(size, dtype=np.float64)
    tol = 1e-9 * max(size)
    interior = np.all((nodes > lo + tol) & (nodes < hi - tol), axis=1)
    rng = np.random.default_rng(seed)
    d = rng.uniform(-1.0, 1.0, size=(int(interior.sum()), 3)) * (jitter * h)[None, :]
    out = nodes.copy()
    out[interior] += d
    return out


def kuhn_tet_block(nx, ny, nz, size=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
                   jitter=0.0, seed=0):
    """Box of 6*nx*ny*nz positively oriented C3D4 tets (Kuhn split)."""
    nodes = _grid_nodes(nx, ny, nz, size, origin)
    perms = list(itertools.permutations(range(3)))
    corner = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    local = []
    for p in perms:
        e = np.eye(3, dtype=np.int64)
        v0 = np.zeros(3, dtype=np.int64)
        v1 = e[p[0]]
        v2 = v1 + e[p[1]]
        v3 = np.ones(3, dtype=np.int64)
        sign = np.linalg.det(np.stack([v1 - v0, v2 - v0, v3 - v0], axis=1).astype(float))
        local.append([v0, v2, v1, v3] if sign < 0 else [v0, v1, v2, v3])
    local = np.array(local)                                    # [6, 4, 3] offsets
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')
    cells = np.stack([I.ravel(order='F'), J.ravel(order='F'), K.ravel(order='F')], axis=1)
    conn = np.empty((cells.shape[0], 6, 4), dtype=np.int64)
    for t in range(6):
        for a in range(4):
            off = local[t, a]
            conn[:, t, a] = _nid(cells[:, 0] + off[0], cells[:, 1] + off[1], cells[:, 2] + off[2], nx, ny)
    conn = conn.reshape(-1, 4)
    node_sets, face_sets = _box_sets(nodes, conn, TET_FACES, size, origin)
    nodes = _jitter_interior(nodes, nx, ny, nz, size, origin, jitter, seed)
    return {'nodes': nodes, 'conn': conn, 'node_sets': node_sets, 'face_sets': face_sets,
            'etype': 'C3D4'}


def hex_block(nx, ny, nz, size=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), jitter=0.0, seed=0):
    """Box of nx*ny*nz C3D8 bricks (node order brick.py:13-22)."""
    nodes = _grid_nodes(nx, ny, nz, size, origin)
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')
    i, j, k = I.ravel(order='F'), J.ravel(order='F'), K.ravel(order='F')
    offs = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    conn = np.stack([_nid(i + a, j + b, k + c, nx, ny) for a, b, c in offs], axis=1).astype(np.int64)
    node_sets, face_sets = _box_sets(nodes, conn, HEX_FACES, size, origin)
    nodes = _jitter_interior(nodes, nx, ny, nz, size, origin, jitter, seed)
    return {'nodes': nodes, 'conn': conn, 'node_sets': node_sets, 'face_sets': face_sets,
            'etype': 'C3D8'}


def tet10_from_tet4(mesh):
    """Quadratic C3D10 mesh from a C3D4 mesh (straight-edged mid nodes)."""
    nodes, conn = mesh['nodes'], mesh['conn']
    edges = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]                   # tetrahedral.py:80-81
    pairs = np.concatenate([np.sort(conn[:, list(e)], axis=1) for e in edges], axis=0)
    uniq, inv = np.unique(pairs, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    mid = 0.5 * (nodes[uniq[:, 0]] + nodes[uniq[:, 1]])
    n0 = nodes.shape[0]
    new_nodes = np.concatenate([nodes, mid], axis=0)
    E = conn.shape[0]
    mids = inv.reshape(len(edges), E).T + n0
    conn10 = np.concatenate([conn, mids], axis=1).astype(np.int64)
    out = dict(mesh)
    out.update({'nodes': new_nodes, 'conn': conn10, 'etype': 'C3D10'})
    # mid nodes on a box face belong to that face's node set
    sets = {}
    for name, idx in mesh['node_sets'].items():
        on = np.zeros(n0, dtype=bool)
        on[idx] = True
        both = on[uniq[:, 0]] & on[uniq[:, 1]]
        sets[name] = np.concatenate([idx, n0 + np.nonzero(both)[0]])
    out['node_sets'] = sets
    return out


def _void_radius(d, r0, amp, waves, phases):
    """r(d) = r0 + amp * mean_k cos(2 pi k_k . d + psi_k) on unit directions d."""
    arg = 2.0 * np.pi * (d @ waves.T) + phases[None, :]
    return r0 + amp * np.cos(arg).mean(axis=1)


def block_with_void(m=10, k=5, ns=5, *, r0=0.3, amp=0.02, n_modes=4, seed=0, quadratic=True):
    """2 x 1 x 1 block with a central void (centre (1, .5, .5), radius r0
    perturbed by ``n_modes`` seeded low-order Fourier modes of amplitude amp).

    Hexahedral blocks: the cube [0.5, 1.5] x [0, 1]^2 around the void is six
    cubed-sphere patches (m x m cells on each cube face, k radial layers from
    the void to the face); the slabs [0, 0.5] and [1.5, 2] are ns x m x m
    cells.  Each hexahedron is split into 12 tets around an added centre node,
    every quad face cut along the diagonal through its smallest global node id,
    so neighbouring cells agree and the mesh is conforming.  With ``quadratic``
    the result is straight-sided C3D10 (corner nodes of the void on its surface).
    Default sizes give 48000 tets.  Sets: box faces 'x0'..'z1' and 'void'."""
    rng = np.random.default_rng(seed)
    waves = rng.integers(-2, 3, size=(n_modes, 3)).astype(np.float64)
    waves[np.all(waves == 0, axis=1), 0] = 1.0
    phases = rng.uniform(0.0, 2.0 * np.pi, n_modes)
    c = np.array([1.0, 0.5, 0.5])
    pts, hexes = [], []

    def add_block(P):                          # P [a+1, b+1, c+1, 3] structured node grid
        A, B, C = P.shape[0] - 1, P.shape[1] - 1, P.shape[2] - 1
        base = sum(p.shape[0] for p in pts)
        pts.append(P.reshape(-1, 3))
        idx = lambda i, j, l: base + (i * (B + 1) + j) * (C + 1) + l
        I, J, L = np.meshgrid(np.arange(A), np.arange(B), np.arange(C), indexing='ij')
        I, J, L = I.ravel(), J.ravel(), L.ravel()
        cell = [idx(I, J, L), idx(I + 1, J, L), idx(I + 1, J + 1, L), idx(I, J + 1, L),
                idx(I, J, L + 1), idx(I + 1, J, L + 1), idx(I + 1, J + 1, L + 1), idx(I, J + 1, L + 1)]
        hexes.append(np.stack(cell, axis=1))

    st = np.linspace(-1.0, 1.0, m + 1)
    rho = np.linspace(0.0, 1.0, k + 1)
    for ax in range(3):
        for sgn in (-1.0, 1.0):
            S, T = np.meshgrid(st, st, indexing='ij')
            face = np.zeros(S.shape + (3,))
            o1, o2 = [a for a in range(3) if a != ax]
            face[..., ax] = sgn
            face[..., o1] = S
            face[..., o2] = T
            d = face / np.linalg.norm(face, axis=-1, keepdims=True)
            r = _void_radius(d.reshape(-1, 3), r0, amp, waves, phases).reshape(S.shape)
            inner = c + r[..., None] * d
            outer = c + 0.5 * face
            P = inner[:, :, None, :] + rho[None, None, :, None] * (outer - inner)[:, :, None, :]
            add_block(P)
    for x0 in (0.0, 1.5):
        xs = np.linspace(x0, x0 + 0.5, ns + 1)
        ys = np.linspace(0.0, 1.0, m + 1)
        X, Y, Z = np.meshgrid(xs, ys, ys, indexing='ij')
        add_block(np.stack([X, Y, Z], axis=-1))
    P = np.concatenate(pts)
    H = np.concatenate(hexes)
    key = np.round(P * 1e9).astype(np.int64)
    uniq, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    nodes = P[first]
    H = inv[H]
    # orient every hex positively (patch cells can come out mirrored)
    Q = nodes[H]
    vol = np.einsum('ei,ei->e', Q[:, 1] - Q[:, 0], np.cross(Q[:, 3] - Q[:, 0], Q[:, 4] - Q[:, 0]))
    neg = vol < 0
    H[neg] = H[neg][:, [0, 3, 2, 1, 4, 7, 6, 5]]
    n0 = nodes.shape[0]
    centres = nodes[H].mean(axis=1)
    nodes = np.concatenate([nodes, centres])
    cid = n0 + np.arange(H.shape[0])
    tets = []
    for f in HEX_FACES.values():
        q = H[:, list(f)]                                        # [E, 4]
        first_min = np.argmin(q, axis=1)
        # diagonal through the smallest id: triangles (q0, q1, q2), (q0, q2, q3) with q0 = min
        roll = np.stack([np.roll(row, -s) for row, s in zip(q, first_min)]) if q.shape[0] else q
        for tri in ((0, 1, 2), (0, 2, 3)):
            tets.append(np.stack([roll[:, tri[0]], roll[:, tri[1]], roll[:, tri[2]], cid], axis=1))
    conn = np.concatenate(tets).astype(np.int64)
    T = nodes[conn]
    vol = np.einsum('ei,ei->e', T[:, 1] - T[:, 0], np.cross(T[:, 2] - T[:, 0], T[:, 3] - T[:, 0]))
    neg = vol < 0
    conn[neg] = conn[neg][:, [0, 2, 1, 3]]
    size = (2.0, 1.0, 1.0)
    node_sets, face_sets = _box_sets(nodes, conn, TET_FACES, size, (0.0, 0.0, 0.0))
    dist = np.linalg.norm(nodes - c, axis=1)
    dirs = (nodes - c) / np.maximum(dist, 1e-300)[:, None]
    on_void = np.abs(dist - _void_radius(dirs, r0, amp, waves, phases)) < 1e-9
    node_sets['void'] = np.nonzero(on_void)[0]
    entries = []
    for fid, loc in TET_FACES.items():
        idx = np.nonzero(np.all(on_void[conn[:, list(loc)]], axis=1))[0]
        if idx.size:
            entries.append((idx, fid))
    face_sets['void'] = entries
    mesh = {'nodes': nodes, 'conn': conn, 'node_sets': node_sets, 'face_sets': face_sets, 'etype': 'C3D4',
            'void': {'centre': c, 'r0': r0, 'amp': amp, 'waves': waves, 'phases': phases}}
    if quadratic:
        # straight-sided C3D10: moving the void edges' mid nodes onto the curved
        # surface inverts Gauss points of the thin tets next to the void
        mesh = tet10_from_tet4(mesh)
    return mesh
