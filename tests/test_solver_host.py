"""Host-side argument handling of the static solver (no GPU calls)."""
import pytest

from fea_amd.solver.static_implicit import predictor_order


@pytest.mark.parametrize('arg,order', [(False, 0), (True, 1), ('none', 0), ('linear', 1), ('quadratic', 2),
                                       ('cubic', 3), (0, 0), (4, 4), (6, 6)])
def test_predictor_order(arg, order):
    assert predictor_order(arg) == order


@pytest.mark.parametrize('arg', ['quartic', 7, -1, 2.0, None])
def test_predictor_order_rejects(arg):
    with pytest.raises(ValueError):
        predictor_order(arg)


def test_auto_predictor_caps_at_six():
    assert predictor_order('auto') == 6


def test_triple_columns_index():
    """The compact column array read by the triple-blocked SpMV: entry rowptr[head] / 9 + j of
    a group is the first column of the head row's neighbour triple j."""
    import numpy as np
    import scipy.sparse as sp
    from fea_amd import meshgen
    from fea_amd.device import row_group_tables, triple_columns
    m = meshgen.hex_block(3, 2, 2)
    N = m['nodes'].shape[0]
    e = m['conn']
    r = np.repeat(e, e.shape[1], axis=1).ravel()
    c = np.tile(e, (1, e.shape[1])).ravel()
    B = sp.csr_matrix((np.ones(r.size), (r, c)), shape=(N, N))
    A = sp.kron(B, np.ones((3, 3))).tocsr()
    A.sort_indices()
    grp, triple = row_group_tables(A.indptr, A.indices, np.full(N, 3))
    assert triple
    tc = triple_columns(A.indptr, A.indices, grp)
    assert tc.size * 9 == A.nnz
    for g in range(N):
        s0, L = A.indptr[grp[g]], A.indptr[grp[g] + 1] - A.indptr[grp[g]]
        assert np.array_equal(tc[s0 // 9 + np.arange(L // 3)], A.indices[s0 + 3 * np.arange(L // 3)])
    with pytest.raises(ValueError):
        triple_columns(A.indptr, A.indices, np.arange(N + 1))      # one-row groups: not triples
    # 3-row groups with row lengths a multiple of 3 whose second run of columns is not
    # consecutive: not triple-structured; the same rows with consecutive runs are
    for cols, want in (([0, 1, 2, 3, 5, 7], False), ([0, 1, 2, 4, 5, 6], True)):
        rowptr = np.arange(0, 6 * 9 + 1, 6)
        grp, triple = row_group_tables(rowptr, np.tile(cols, 9), np.full(3, 3))
        assert grp is not None and triple == want
