"""CPU checks of the C ABI: every symbol of include/fea_amd.h is exported, and the
host-side pattern builder reproduces the reference CSR bit-exactly."""
import ctypes
import os
import re

import numpy as np
import pytest

from fea_amd import _native, meshgen
from oracle import assembly as oasm
from tests.helpers import nh, oracle_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, 'include', 'fea_amd.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(fea_\w+)\s*\(', txt)))


def test_library_exports_every_header_symbol():
    L = _native.lib()
    names = header_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.fea_version() >= 1


def test_binding_covers_header():
    assert set(header_functions()) <= set(_native.exported_symbols())


TORCH_OPS = ['elem_residual_tangent', 'elem_tangent_direct', 'elem_residual', 'elem_energy', 'gather_csr', 'gather_residual', 'spmv',
             'pcg_start', 'pcg_iterate', 'jacobi_inv', 'block_jacobi_inv']


def test_torch_operators_registered():
    """The TORCH_LIBRARY(fea, ...) library loads and registers every operator with a schema;
    there is no CPU kernel, so CPU tensors are refused (no silent host fallback)."""
    import torch
    O = _native.ops()
    for name in TORCH_OPS:
        schema = getattr(O, name).default._schema
        assert schema.name == 'fea::' + name
    x = torch.zeros(3, dtype=torch.float64)
    rowptr = torch.tensor([0, 1, 2, 3], dtype=torch.int32)
    col = torch.tensor([0, 1, 2], dtype=torch.int32)
    with pytest.raises((NotImplementedError, RuntimeError)):
        O.spmv(1, 1, 0, None, rowptr, col, x, x, x.clone())


def build_pattern(model):
    """Call fea_pattern_create directly (host code, no GPU)."""
    L = _native.lib()
    free = oasm.free_mask(model)
    N = model['nodes'].shape[0]
    dof = np.full(3 * N, -1, dtype=np.int32)
    dof[free] = np.arange(int(free.sum()), dtype=np.int32)
    conns = [np.ascontiguousarray(g['conn'], dtype=np.int32) for g in model['groups']]
    conns += [np.ascontiguousarray(p['conn'], dtype=np.int32) for p in model['pressures']]
    ptrs = (ctypes.c_void_p * len(conns))(*[c.ctypes.data for c in conns])
    ne = np.array([c.shape[0] for c in conns], dtype=np.int32)
    nen = np.array([c.shape[1] for c in conns], dtype=np.int32)
    ink = np.ones(len(conns), dtype=np.int32)
    h = ctypes.c_void_p()
    _native.check(L.fea_pattern_create(N, dof.ctypes.data, len(conns), ptrs, ne.ctypes.data, nen.ctypes.data,
                                       ink.ctypes.data, ctypes.byref(h)), 'create')
    sizes = np.zeros(5, dtype=np.int64)
    L.fea_pattern_sizes(h, sizes.ctypes.data)
    n_free, nnz, nb, nks, nrs = [int(s) for s in sizes]
    rowptr = np.zeros(n_free + 1, dtype=np.int32)
    col = np.zeros(nnz, dtype=np.int32)
    diag = np.zeros(n_free, dtype=np.int32)
    L.fea_pattern_export_csr(h, rowptr.ctypes.data, col.ctypes.data, diag.ctypes.data)
    blk = [np.zeros(nb, dtype=np.int32) for _ in range(3)]
    kseg = np.zeros(nb + 1, dtype=np.int32)
    rseg = np.zeros(N + 1, dtype=np.int32)
    L.fea_pattern_export_blocks(h, blk[0].ctypes.data, blk[1].ctypes.data, blk[2].ctypes.data,
                                kseg.ctypes.data, rseg.ctypes.data)
    groups = []
    for gi, c in enumerate(conns):
        dk = np.zeros(c.shape[0] * c.shape[1] ** 2, dtype=np.int32)
        dr = np.zeros(c.shape[0] * c.shape[1], dtype=np.int32)
        cp = np.zeros(c.shape[0] * c.shape[1] ** 2, dtype=np.int32)
        L.fea_pattern_export_group(h, gi, dk.ctypes.data, dr.ctypes.data)
        _native.check(L.fea_pattern_export_group_colpos(h, gi, cp.ctypes.data), 'colpos')
        groups.append((dk, dr, cp))
    L.fea_pattern_destroy(h)
    return dict(n_free=n_free, nnz=nnz, rowptr=rowptr, col=col, diag=diag, blk=blk, kseg=kseg, rseg=rseg,
                groups=groups, dof=dof, nks=nks, nrs=nrs)


def cases():
    c1 = meshgen.kuhn_tet_block(3, 3, 3, jitter=0.05, seed=1)
    c2 = meshgen.hex_block(4, 2, 2, size=(4, 1, 1))
    c3 = meshgen.tet10_from_tet4(meshgen.kuhn_tet_block(2, 2, 2))
    return [oracle_model(c1, nh(1, 100), 'z0', pressure=0.2, load_set='z1'),
            oracle_model(c2, nh(1, 100), 'x0'),
            oracle_model(c3, nh(1, 10), 'z0', pressure=0.1, load_set='z1')]


@pytest.mark.parametrize('ci', range(3))
def test_pattern_csr_bit_exact(ci):
    model = cases()[ci]
    P = build_pattern(model)
    indptr, indices = oasm.csr_pattern_bruteforce(model)
    assert np.array_equal(P['rowptr'], indptr)
    assert np.array_equal(P['col'], indices)
    # the reference reduction + coalesce (assembly.py:371-380) gives the same structure
    U = np.zeros_like(model['nodes'])
    _, K = oasm.assemble_reduced(model, U)
    assert np.array_equal(K.indptr, indptr) and np.array_equal(K.indices, indices)
    # diag slots point at the diagonal
    rows = np.repeat(np.arange(P['n_free']), np.diff(P['rowptr']))
    assert np.array_equal(rows[P['diag']], P['col'][P['diag']])


@pytest.mark.parametrize('ci', range(3))
def test_staging_maps_are_a_permutation(ci):
    model = cases()[ci]
    P = build_pattern(model)
    dks = np.concatenate([g[0] for g in P['groups']])
    used = dks[dks >= 0]
    assert np.array_equal(np.sort(used), np.arange(P['nks']))
    drs = np.concatenate([g[1] for g in P['groups']])
    usedr = drs[drs >= 0]
    assert np.array_equal(np.sort(usedr), np.arange(P['nrs']))
    # every K staging slot of block q lies inside kseg[q]..kseg[q+1] and pairs (A,B) match
    conn = model['groups'][0]['conn']
    nen = conn.shape[1]
    dk = P['groups'][0][0].reshape(-1, nen, nen)
    blk_of_slot = np.searchsorted(P['kseg'], np.arange(P['nks']), side='right') - 1
    e, a, b = np.nonzero(dk >= 0)
    q = blk_of_slot[dk[e, a, b]]
    assert np.array_equal(P['blk'][0][q], conn[e, a]) and np.array_equal(P['blk'][1][q], conn[e, b])
    # inside each block's segment the slots follow (group, element, a, b) order: the fixed
    # order of the staging reduction
    key, slot = [], []
    for gi, g in enumerate(P['groups']):
        d = g[0]
        t = np.nonzero(d >= 0)[0]
        key.append(gi * (1 << 40) + t)
        slot.append(d[t])
    key, slot = np.concatenate(key), np.concatenate(slot)
    order = np.argsort(slot)
    k_sorted, blk_sorted = key[order], blk_of_slot[slot[order]]
    same = blk_sorted[1:] == blk_sorted[:-1]
    assert np.all(k_sorted[1:][same] > k_sorted[:-1][same])


def test_pattern_rejects_bad_dof_map():
    L = _native.lib()
    dof = np.array([1, 0, -1], dtype=np.int32)
    conn = np.zeros((0, 4), dtype=np.int32)
    ptrs = (ctypes.c_void_p * 1)(conn.ctypes.data)
    h = ctypes.c_void_p()
    rc = L.fea_pattern_create(1, dof.ctypes.data, 1, ptrs, np.array([0], np.int32).ctypes.data,
                              np.array([4], np.int32).ctypes.data, np.array([1], np.int32).ctypes.data,
                              ctypes.byref(h))
    assert rc != 0 and b'dof_map' in L.fea_last_error()


def test_pattern_empty_model():
    model = {'nodes': np.zeros((0, 3)), 'fixed': np.zeros((0, 3), bool), 'groups': [], 'pressures': []}
    L = _native.lib()
    h = ctypes.c_void_p()
    rc = L.fea_pattern_create(0, None, 0, None, None, None, None, ctypes.byref(h))
    assert rc == 0
    sizes = np.zeros(5, dtype=np.int64)
    L.fea_pattern_sizes(h, sizes.ctypes.data)
    assert sizes.sum() == 0
    L.fea_pattern_destroy(h)


def color_elements(conn, n_nodes):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    color = np.zeros(conn.shape[0], dtype=np.int32)
    nc = ctypes.c_int32(0)
    _native.check(_native.lib().fea_color_elements(n_nodes, conn.shape[0], conn.shape[1], conn.ctypes.data,
                                                   color.ctypes.data, ctypes.byref(nc)), 'color')
    return color, nc.value


@pytest.mark.parametrize('ci', range(3))
def test_element_colours_and_direct_columns(ci):
    """fea_color_elements: no two elements of a colour share a node (a structured brick
    grid gets its 8 parity colours); fea_pattern_export_group_colpos: rowptr[row] +
    colpos + the free-column offset of every (element, a, b) block is the CSR entry of
    that row and column -- the addresses fea_elem_tangent_direct adds into."""
    model = cases()[ci]
    conn = model['groups'][0]['conn']
    N = model['nodes'].shape[0]
    color, nc = color_elements(conn, N)
    assert color.min() == 0 and color.max() == nc - 1
    for c in range(nc):
        nodes = conn[color == c].reshape(-1)
        assert np.unique(nodes).size == nodes.size
    if ci == 1:
        assert nc == 8
    P = build_pattern(model)
    dof, rowptr, col = P['dof'], P['rowptr'], P['col']
    dk, _, cp = P['groups'][0]
    nen = conn.shape[1]
    cp = cp.reshape(-1, nen, nen)
    assert np.array_equal(cp >= 0, dk.reshape(-1, nen, nen) >= 0)
    checked = 0
    for e in range(conn.shape[0]):
        for a in range(nen):
            for b in range(nen):
                if cp[e, a, b] < 0:
                    continue
                A, B = conn[e, a], conn[e, b]
                fc = [dof[3 * B + k] for k in range(3) if dof[3 * B + k] >= 0]
                for j in range(3):
                    r = dof[3 * A + j]
                    if r < 0:
                        continue
                    base = rowptr[r] + cp[e, a, b]
                    assert list(col[base:base + len(fc)]) == fc
                    checked += 1
    assert checked > 0
