"""Host-side Couple maps (fea_amd.couple.couple_maps) against the oracle, on the CPU.

The maps are built from the extended CSR (coupled node DOFs kept as rows); the
condensation the HIP kernels perform is replayed here in numpy over those maps, and the
reduced residual / tangent must equal the oracle's T^T R and T^T K T + R d2u/dz2.  This
pins the map construction (slots, pairs, transposes, the point block, fixed point DOFs).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from fea_amd.couple import couple_maps
from oracle import assembly as oasm
from oracle import couple as cp
from tests.test_oracle_couple import _rand_gc, couple_model


def _ext_system(om, GC):
    """Extended CSR pattern, values and residual (coupled node DOFs kept) at GC."""
    ext = oasm.free_mask(om)                                 # free before the Couple removes its nodes
    dof_ext = np.full(ext.shape[0], -1)
    dof_ext[ext] = np.arange(int(ext.sum()))
    U, rp = cp.state(om, GC)
    R_full, (r, c, v) = oasm.assemble_full(om, U)
    keep = ext[r] & ext[c]
    K = sp.coo_matrix((v[keep], (dof_ext[r[keep]], dof_ext[c[keep]])), shape=(int(ext.sum()),) * 2).tocsr()
    K.sum_duplicates()
    K.sort_indices()
    return dof_ext, K, R_full[ext], U, rp


def _replay(m, K_ext, R_ext, ydot, ydot2, load_scale=1.0):
    """numpy replay of fea_couple_condense over the maps."""
    n_red, nnz = m['n_red'], m['nnz']
    val = np.zeros(nnz)
    R = np.zeros(n_red)
    sel = m['ff_src'] >= 0
    val[sel] = K_ext[m['ff_src'][sel]]
    R[:m['n_red_nodes']] = R_ext[m['rf_src']]

    def T(q, d):
        return (1.0 if q % 3 == d else 0.0) if d < 3 else ydot[q, d - 3]
    for p in range(m['pair_ptr'].shape[0] - 1):
        acc = np.zeros(6)
        for e in range(m['pair_ptr'][p], m['pair_ptr'][p + 1]):
            for d in range(6):
                acc[d] += K_ext[m['pe_slot'][e]] * T(m['pe_dof'][e], d)
        for d in range(6):
            for s in (m['pair_row_slot'][p, d], m['pair_col_slot'][p, d]):
                if s >= 0:
                    val[s] = acc[d]
    nc = m['cnode_ptr'].shape[0] - 1
    for k in range(nc):
        blk = np.zeros((6, 6))
        rv = np.zeros(6)
        for q in range(3 * m['cnode_ptr'][k], 3 * m['cnode_ptr'][k + 1]):
            r = R_ext[m['crow_ext'][q]]
            t = np.array([T(q, d) for d in range(6)])
            u = np.zeros(6)
            for e in range(m['cr_ptr'][q], m['cr_ptr'][q + 1]):
                u += K_ext[m['ce_slot'][e]] * np.array([T(m['ce_dof'][e], d) for d in range(6)])
            blk += np.outer(t, u)
            blk[3:, 3:] += r * ydot2[q]
            rv += r * t
        for j in range(36):
            s = m['rp_slot'][36 * k + j]
            if s >= 0:
                val[s] = blk[j // 6, j % 6]
        for d in range(6):
            row = m['red_rp'][6 * k + d]
            if row >= 0:
                R[row] = rv[d] - load_scale * m['load'][6 * k + d]
    return sp.csr_matrix((val, m['col'], m['rowptr']), shape=(n_red, n_red)), R


@pytest.mark.parametrize('etype,fixed_rp', [('C3D8', None), ('C3D4', [False, True, False, True, False, False]),
                                            ('C3D10', None)])
def test_replayed_condensation_matches_oracle(etype, fixed_rp):
    om, mesh = couple_model(etype, pressure=0.05, fixed_rp=fixed_rp)
    GC = _rand_gc(om, 0.02, 4)
    dof_ext, K_ext, R_ext, U, rp = _ext_system(om, GC)
    node_free, rp_free = cp.layout(om)
    r0 = om['rps'][0]
    rps = [{'node': r0['node'], 'free': rp_free[:6], 'load': np.concatenate([r0['force'], r0['moment']])}]
    couples = [(np.asarray(mesh['node_sets']['x1']), 0)]
    m = couple_maps(om['nodes'], node_free, dof_ext, K_ext.indptr, K_ext.indices, couples, rps)
    y = om['nodes'][m['cnode']] - np.asarray(r0['node'])
    Yd, Ydd = cp.rotation_jacobians(rp[0, 3:], y)
    K, R = _replay(m, K_ext.data, R_ext, Yd.reshape(-1, 3), Ydd.reshape(-1, 3, 3))
    R_o, K_o = cp.assemble_reduced(om, GC)
    assert m['n_red'] == cp.n_gc(om)
    assert np.abs(R - R_o).max() < 1e-13 * np.abs(R_o).max()
    assert np.abs((K - K_o).toarray()).max() < 1e-13 * np.abs(K_o.data).max()
    # the pattern holds every structural entry of the oracle's product, and nothing else
    Ko = K_o.tocoo()
    big = np.abs(Ko.data) > 1e-14 * np.abs(Ko.data).max()
    ours = set(zip(np.repeat(np.arange(m['n_red']), np.diff(m['rowptr'])).tolist(), m['col'].tolist()))
    assert set(zip(Ko.row[big].tolist(), Ko.col[big].tolist())) <= ours
    assert m['diag'].min() >= 0
    # node rows of one node share their columns, so the node-grouped SpMV applies
    assert m['grp'] is not None
