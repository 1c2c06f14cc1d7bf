"""Tile tables of the explicit loop's residual (fea_elem_residual_tiles /
fea_explicit_update_tiles).

The elements of each group are put in Morton order of their centroids (so that a tile
of consecutive elements covers a compact piece of the mesh and few nodes), cut into
tiles of TE elements, and every (tile, free node) pair gets one partial sum: the
float64 sum, in element order, of the tile's element forces on that node.  A node's
force is then the sum of its partials (about two per node on a structured tet mesh
with TE = 512), in group and tile order.  This replaces the reference's scatter-add of
element forces (C3base.py:428-431 through assembly.py:347-380) for the explicit loop.
"""
import numpy as np


def morton_order(points, bits=10):
    """Stable permutation that sorts ``points`` [n, 3] along a Morton (Z-order) curve of
    their bounding box quantised to 2^bits cells per axis."""
    p = np.asarray(points, dtype=np.float64)
    if p.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    lo, span = p.min(0), np.ptp(p, axis=0)
    span = np.where(span > 0, span, 1.0)
    q = np.minimum(((p - lo) / span * (1 << bits)).astype(np.int64), (1 << bits) - 1)
    code = np.zeros(p.shape[0], dtype=np.int64)
    for b in range(bits):
        for d in range(3):
            code |= ((q[:, d] >> b) & 1) << (3 * b + d)
    return np.argsort(code, kind='stable')


def group_tiles(conn, X, free_node, tile_size, partial_base=0):
    """Tables of one element group.

    conn [E, nen] node ids, X [N, 3] reference coordinates, free_node [N] bool (a node
    with at least one free DOF).  Returns a dict with
      order       [E]        element permutation (tile kernel element i = group element order[i]),
      tile_lbase  [T + 1]    first partial of each tile (global index, from partial_base),
      pslot_ptr   [P + 1]    each partial's range in pslot (relative to this group),
      pslot       [Q] uint16 local slots e_local * nen + a, element order within a partial,
      pnode       [P]        the node of each partial,
      tnode_ptr   [T + 1], tnode: the nodes each tile's elements touch (sorted),
      lconn       [E * nen] uint16: element i's nodes as indices into its tile's list,
      max_tile_nodes: the largest tile's node count.
    """
    conn = np.asarray(conn, dtype=np.int64)
    E, nen = conn.shape
    TE = int(tile_size)
    if TE * nen > 65536:
        raise ValueError('tile slots exceed uint16')
    order = morton_order(X[conn].mean(axis=1)) if E else np.zeros(0, dtype=np.int64)
    c = conn[order]
    T = (E + TE - 1) // TE
    e = np.repeat(np.arange(E, dtype=np.int64), nen)
    node = c.reshape(-1)
    slot = (e % TE) * nen + np.tile(np.arange(nen, dtype=np.int64), E)
    keep = np.asarray(free_node, dtype=bool)[node]
    tile = e[keep] // TE
    N = max(1, int(X.shape[0]))
    key = tile * N + node[keep]
    uniq, inv = np.unique(key, return_inverse=True)     # partials sorted by (tile, node)
    inv = inv.reshape(-1)
    P = uniq.shape[0]
    lbase = np.searchsorted(uniq // N, np.arange(T + 1)) + partial_base
    o = np.argsort(inv, kind='stable')                  # keeps element order inside a partial
    pslot = slot[keep][o].astype(np.uint16)
    pslot_ptr = np.concatenate([[0], np.cumsum(np.bincount(inv, minlength=P))])
    # every node a tile's elements touch (free or not), sorted, and each element's nodes as
    # indices into its tile's list (the kernel stages these nodes' X and U in LDS)
    tkey = np.unique(e // TE * N + node)
    tnode = tkey % N
    tnode_ptr = np.searchsorted(tkey // N, np.arange(T + 1))
    lconn = (np.searchsorted(tkey, e // TE * N + node) - tnode_ptr[e // TE]).astype(np.uint16)
    return {'order': order, 'te': TE, 'tile_lbase': lbase.astype(np.int64),
            'pslot_ptr': pslot_ptr.astype(np.int64), 'pslot': pslot, 'pnode': (uniq % N).astype(np.int64),
            'tnode_ptr': tnode_ptr.astype(np.int64), 'tnode': tnode.astype(np.int64), 'lconn': lconn,
            'max_tile_nodes': int(np.diff(tnode_ptr).max()) if T else 0}


def node_partials(pnodes, n_nodes):
    """CSR node -> partials (partials of a node in ascending partial index: group, tile)."""
    pn = np.asarray(pnodes, dtype=np.int64)
    o = np.argsort(pn, kind='stable')
    ptr = np.concatenate([[0], np.cumsum(np.bincount(pn, minlength=n_nodes))])
    return ptr.astype(np.int64), o.astype(np.int64)


def emulate(groups, tables, elem_forces, n_nodes):
    """numpy emulation of the kernels' sums (for tests): ``elem_forces[g]`` [E, nen, 3] in the
    group's own element order; returns the per-node force [n_nodes, 3] summed partial by partial."""
    partial = []
    for g, t, F in zip(groups, tables, elem_forces):
        nen = g.shape[1]
        Fo = np.asarray(F, dtype=np.float64)[t['order']]
        TE = t['te']
        lb = t['tile_lbase']
        base = lb[0]
        P = t['pnode'].shape[0]
        out = np.zeros((P, 3))
        # tile of each partial
        tile_of = np.searchsorted(lb, np.arange(base, base + P), side='right') - 1
        for p in range(P):
            tt = tile_of[p]
            for q in range(t['pslot_ptr'][p], t['pslot_ptr'][p + 1]):
                sl = int(t['pslot'][q])
                out[p] += Fo[tt * TE + sl // nen, sl % nen]
        partial.append(out)
    partial = np.concatenate(partial) if partial else np.zeros((0, 3))
    pn = np.concatenate([t['pnode'] for t in tables]) if tables else np.zeros(0, dtype=np.int64)
    ptr, lst = node_partials(pn, n_nodes)
    f = np.zeros((n_nodes, 3))
    for A in range(n_nodes):
        for s in range(ptr[A], ptr[A + 1]):
            f[A] += partial[lst[s]]
    return f
