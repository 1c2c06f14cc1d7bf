"""CPU checks of the solve server's host logic: the request stream is
reproducible and follows its stated distributions, and the admission
footprint estimate bounds the exact device sizes of a request (computed here
with the host pattern builder, no GPU)."""
import numpy as np
import pytest

from fea_amd import serving
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


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_footprint_estimate_bounds_the_exact_sizes(seed):
    r = serving.random_requests(3, 1.0, seed=seed, e_range=(500, 5000))[seed % 3]
    spec = r.spec
    mesh = serving.request_mesh(spec)
    face = serving._OPPOSITE[spec['fixed']]
    kind, val = spec['load']
    om = oracle_model(mesh, nh(1.0, 100.0), spec['fixed'], pressure=val if kind == 'pressure' else None,
                      traction=val if kind == 'traction' else None, load_set=face)
    pat = build_pattern(om)
    N = mesh['nodes'].shape[0]
    E, nen = mesh['conn'].shape
    exact = (9 * pat['nks'] * 8 + pat['nnz'] * (4 + 4) + 4 * (pat['n_free'] + 1) + 4 * pat['n_free']
             + E * nen * nen * 4 + 2 * E * nen * 4 + 3 * N * 8 * 4)
    est = serving.estimate_footprint(spec, 'mixed')
    assert exact <= est <= 2.0 * exact
