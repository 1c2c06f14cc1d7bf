"""CPU checks of the explicit loop's tile tables (fea_amd/solver/_explicit_tiles.py): the
partials cover every (element, free node) incidence once, their sums reproduce a plain
scatter-add of element forces, and the Morton order is a permutation that keeps tiles
compact."""
import numpy as np
import pytest

from fea_amd import meshgen
from fea_amd.solver._explicit_tiles import emulate, group_tiles, morton_order, node_partials


def _mesh(kind):
    if kind == 'tet':
        return meshgen.kuhn_tet_block(6, 5, 4, jitter=0.1, seed=3)
    return meshgen.hex_block(6, 5, 4, jitter=0.1, seed=3)


@pytest.mark.parametrize('kind,te', [('tet', 64), ('tet', 512), ('hex', 32)])
def test_tile_sums_equal_scatter_add(kind, te):
    m = _mesh(kind)
    conn, X = m['conn'], m['nodes']
    N = X.shape[0]
    free = np.ones(N, dtype=bool)
    free[m['node_sets']['z0']] = False
    F = np.random.default_rng(1).standard_normal(conn.shape + (3,))
    t = group_tiles(conn, X, free, te)
    ref = np.zeros((N, 3))
    np.add.at(ref, conn.reshape(-1), F.reshape(-1, 3))
    ref[~free] = 0.0
    assert np.abs(emulate([conn], [t], [F], N) - ref).max() < 1e-13
    # every (element, free node) incidence in exactly one partial
    nen = conn.shape[1]
    assert t['pslot'].shape[0] == int(free[conn].sum())
    lb = t['tile_lbase']
    for tt in range(lb.shape[0] - 1):
        sl = np.concatenate([t['pslot'][t['pslot_ptr'][p]:t['pslot_ptr'][p + 1]] for p in range(lb[tt], lb[tt + 1])]
                            or [np.zeros(0, np.uint16)])
        assert np.unique(sl).shape[0] == sl.shape[0]
        assert sl.max(initial=0) < te * nen
        # the partial's node is the node of each of its slots
        c = conn[t['order']]
        for p in range(lb[tt], lb[tt + 1]):
            s = t['pslot'][t['pslot_ptr'][p]:t['pslot_ptr'][p + 1]].astype(np.int64)
            assert np.all(c[tt * te + s // nen, s % nen] == t['pnode'][p])
            assert np.all(np.diff(s) > 0)                     # element order inside a partial


def test_two_groups_partials_in_group_order():
    a, b = _mesh('tet'), _mesh('tet')
    N = a['nodes'].shape[0]
    X = a['nodes']
    free = np.ones(N, dtype=bool)
    ta = group_tiles(a['conn'], X, free, 64, 0)
    tb = group_tiles(b['conn'][::-1].copy(), X, free, 64, ta['pnode'].shape[0])
    ptr, lst = node_partials(np.concatenate([ta['pnode'], tb['pnode']]), N)
    assert np.all(np.diff(ptr) >= 2)
    for A in range(N):
        assert np.all(np.diff(lst[ptr[A]:ptr[A + 1]]) > 0)


def test_morton_order_is_a_compact_permutation():
    m = meshgen.kuhn_tet_block(16, 16, 16)
    cent = m['nodes'][m['conn']].mean(axis=1)
    o = morton_order(cent)
    assert np.array_equal(np.sort(o), np.arange(cent.shape[0]))
    # 512-element tiles of the Morton order touch fewer distinct nodes than of the mesh order
    def touched(order):
        c = m['conn'][order]
        return sum(np.unique(c[i:i + 512]).shape[0] for i in range(0, c.shape[0], 512))
    assert touched(o) < 0.9 * touched(np.arange(cent.shape[0]))


def test_tile_node_lists_and_local_connectivity():
    m = _mesh('tet')
    conn, X = m['conn'], m['nodes']
    t = group_tiles(conn, X, np.ones(X.shape[0], dtype=bool), 64)
    c = conn[t['order']]
    for i in range(c.shape[0]):
        tt = i // 64
        nodes = t['tnode'][t['tnode_ptr'][tt]:t['tnode_ptr'][tt + 1]]
        assert np.all(np.diff(nodes) > 0)
        assert np.array_equal(nodes[t['lconn'][4 * i:4 * i + 4]], c[i])
    assert t['max_tile_nodes'] == int(np.diff(t['tnode_ptr']).max())
