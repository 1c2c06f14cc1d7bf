"""CPU checks of the solve server's host logic: the request stream is
reproducible and follows its stated distributions, a pressure request's load
is the exact follower pressure, and the admission footprint estimate bounds
the exact device sizes of a request (computed here with the host pattern
builder, no GPU)."""
import numpy as np
import pytest

from fea_amd import serving
from oracle import surfaces as osf
from tests.helpers import facets, nh, oracle_model
from tests.test_abi_pattern import build_pattern


def test_stream_is_reproducible_and_in_range():
    a = serving.random_requests(200, 5.0, seed=4)
    b = serving.random_requests(200, 5.0, seed=4)
    assert [r.spec for r in a] == [r.spec for r in b]
    assert [r.arrival for r in a] == [r.arrival for r in b]
    arr = np.array([r.arrival for r in a])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    # Poisson stream: mean gap 1/qps within sampling error (199 gaps, CV 1)
    assert abs(np.diff(arr).mean() - 0.2) < 0.2 * 4 / np.sqrt(199)
    E = np.array([r.n_elements() for r in a])
    assert E.min() >= 0.5 * 2e3 and E.max() <= 1.5 * 2e5
    # log-uniform: median near the geometric mean of the range
    assert 0.5 * 2e4 < np.median(E) < 2.0 * 2e4
    types = {r.spec['etype'] for r in a}
    assert types == {'C3D4', 'C3D8'}
    for r in a:
        m = r.spec['material']
        if m[0] == 'NH':
            assert 0.5 <= m[1] <= 1.5 and 50 <= m[2] <= 150
        else:
            mu = 2 * (m[1] + m[2])
            assert 0.5 - 1e-12 <= mu <= 1.5 + 1e-12 and 50 <= 2 / m[3] <= 150
        kind, val = r.spec['load']
        if kind == 'pressure':
            assert val > 0          # compressive follower pressure
        else:
            assert np.count_nonzero(val) == 1


def _functional_force(mesh, set_name, p):
    conn, st = facets(mesh, set_name)
    F, _ = osf.pressure_force_stiffness(st, mesh['nodes'], conn, p, onlyforce=True)
    f = np.zeros(mesh['nodes'].size)
    np.add.at(f, osf.pressure_force_idx(conn), F)
    return f.reshape(-1, 3)


def _consistent_force(mesh, faces, p):
    """sum over faces of the dead traction -p n_out: a uniform inward pressure."""
    g = np.zeros_like(mesh['nodes'])
    for face in faces:
        n = np.zeros(3)
        n['xyz'.index(face[0])] = 1.0 if face[1] == '1' else -1.0
        conn, st = facets(mesh, face)
        g += osf.traction_force(st, mesh['nodes'], conn, -p * n)
    return g


@pytest.mark.parametrize('etype,fixed', [('C3D8', 'y0'), ('C3D4', 'z1')])
def test_pressure_request_load_is_the_follower_pressure(etype, fixed):
    """A pressure request loads the five unclamped faces: the volume functional's
    nodal forces equal the consistent nodal forces of a uniform inward pressure on
    every node that can move (its edge terms act on the clamped perimeter only).
    On the one face opposite the clamp alone they do not: its perimeter nodes carry
    the functional's line loads."""
    p = 0.7
    spec = {'etype': etype, 'cells': (5, 4, 3), 'size': (2.0, 1.5, 1.0), 'material': ('NH', 1.0, 100.0),
            'fixed': fixed, 'load': ('pressure', p), 'load_steps': 2}
    mesh = serving.request_mesh(spec)
    free = np.ones(mesh['nodes'].shape[0], dtype=bool)
    free[mesh['node_sets'][fixed]] = False
    faces = [f for f in serving._OPPOSITE if f != fixed]
    f = _functional_force(mesh, serving.LOADED_SET, p)
    g = _consistent_force(mesh, faces, p)
    assert np.abs(f[free] - g[free]).max() < 1e-12 * np.abs(g).max()
    assert np.abs(mesh['nodes'].mean(axis=0)).max() < 1e-12          # origin inside the block
    face = serving._OPPOSITE[fixed]
    f1 = _functional_force(mesh, face, p)
    g1 = _consistent_force(mesh, [face], p)
    assert np.abs(f1[free] - g1[free]).max() > 0.1 * np.abs(g1).max()


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_footprint_estimate_bounds_the_exact_sizes(seed):
    r = serving.random_requests(3, 1.0, seed=seed, e_range=(500, 5000))[seed % 3]
    spec = r.spec
    mesh = serving.request_mesh(spec)
    kind, val = spec['load']
    om = oracle_model(mesh, nh(1.0, 100.0), spec['fixed'], pressure=val if kind == 'pressure' else None,
                      traction=val if kind == 'traction' else None, load_set=serving.LOADED_SET)
    pat = build_pattern(om)
    N = mesh['nodes'].shape[0]
    E, nen = mesh['conn'].shape
    exact = (9 * pat['nks'] * 8 + pat['nnz'] * (4 + 4) + 4 * (pat['n_free'] + 1) + 4 * pat['n_free']
             + E * nen * nen * 4 + 2 * E * nen * 4 + 3 * N * 8 * 4)
    est = serving.estimate_footprint(spec, 'mixed')
    assert exact <= est <= 2.0 * exact
