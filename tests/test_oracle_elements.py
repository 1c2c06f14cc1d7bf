"""Pin the oracle's element restatement with closed forms and derivative checks."""
import numpy as np
import pytest

from fea_amd import meshgen
from oracle import elements as el
from oracle import materials as om


def meshes():
    t4 = meshgen.kuhn_tet_block(2, 2, 2, size=(1.0, 0.8, 1.2), jitter=0.05, seed=3)
    h8 = meshgen.hex_block(2, 2, 2, size=(1.5, 1.0, 0.7), jitter=0.05, seed=4)
    t10 = meshgen.tet10_from_tet4(meshgen.kuhn_tet_block(2, 1, 1, jitter=0.05, seed=5))
    return [t4, h8, t10]


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
def test_shape_partition_of_unity(etype):
    pts = np.random.default_rng(0).uniform(0, 0.3, size=(7, 3))
    N, dN = el.shape_at(etype, pts)
    assert np.allclose(N.sum(1), 1.0, atol=1e-14)
    assert np.allclose(dN.sum(2), 0.0, atol=1e-13)


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
def test_shape_nodal_interpolation(etype):
    # N_a(xi_b) = delta_ab at the element's own nodes
    nodes = {'C3D4': [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
             'C3D8': [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
             'C3D10': [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [.5, .5, 0],
                       [0, .5, 0], [0, 0, .5], [.5, 0, .5], [0, .5, .5]]}[etype]
    N, _ = el.shape_at(etype, np.array(nodes, dtype=float))
    assert np.allclose(N, np.eye(len(nodes)), atol=1e-14)


@pytest.mark.parametrize('mi', range(3))
def test_volume_and_patch(mi):
    m = meshes()[mi]
    X, conn = m['nodes'], m['conn']
    dNdX, w, N = el.pre_load_gaussian(m['etype'], X, conn)
    # sum w*detJ = volume of the box (jitter keeps the boundary; straight edges)
    lo, hi = X.min(0), X.max(0)
    assert abs(w.sum() - np.prod(hi - lo)) < 1e-12
    # affine displacement u = (A - I) X gives F = A exactly at every Gauss point
    A = np.array([[1.1, 0.05, -0.02], [0.03, 0.95, 0.04], [-0.01, 0.02, 1.05]])
    U = X @ (A - np.eye(3)).T
    F = el.deformation_gradient(dNdX, conn, U)
    assert np.abs(F - A).max() < 1e-12


def _setup(mi, mat):
    m = meshes()[mi]
    dNdX, w, _ = el.pre_load_gaussian(m['etype'], m['nodes'], m['conn'])
    rng = np.random.default_rng(11 + mi)
    U = 0.03 * rng.standard_normal(m['nodes'].shape)
    return m, dNdX, w, U


MATS = [om.NeoHookean(1.0, 100.0), om.MooneyRivlin(0.5, 0.1, 0.01)]


@pytest.mark.parametrize('mi', range(3))
@pytest.mark.parametrize('mat', MATS, ids=['NH', 'MR'])
def test_residual_translation_invariance_and_energy_gradient(mi, mat):
    m, dNdX, w, U = _setup(mi, mat)
    conn = m['conn']
    R, _ = el.structural_force(dNdX, w, conn, U, mat, onlyforce=True)
    nen, E = conn.shape[1], conn.shape[0]
    Re = R.reshape(nen, 3, E)
    # sum_a R_a = 0 per element (rigid translation does no work)
    assert np.abs(Re.sum(0)).max() < 1e-10 * np.abs(Re).max()
    # residual = gradient of the element energy (complex step on U of a single element)
    Rglob = np.zeros(U.size)
    np.add.at(Rglob, el.force_indices(conn), R)
    h = 1e-30
    rng = np.random.default_rng(0)
    for dof in rng.choice(U.size, 6, replace=False):
        Uc = U.astype(complex)
        Uc.reshape(-1)[dof] += 1j * h
        F = el.deformation_gradient(dNdX, conn, Uc)
        Ec = np.einsum('ge,ge->', mat.energy(F), w)
        assert abs(Ec.imag / h - Rglob[dof]) <= 1e-10 * max(1.0, np.abs(Rglob).max())


@pytest.mark.parametrize('mi', range(3))
@pytest.mark.parametrize('mat', MATS, ids=['NH', 'MR'])
def test_tangent_is_residual_derivative_and_symmetric(mi, mat):
    m, dNdX, w, U = _setup(mi, mat)
    conn = m['conn']
    R, K = el.structural_force(dNdX, w, conn, U, mat)
    rows, cols = el.coo_indices(conn)
    n = U.size
    Kd = np.zeros((n, n))
    np.add.at(Kd, (rows, cols), K)
    assert np.abs(Kd - Kd.T).max() <= 1e-11 * np.abs(Kd).max()
    h = 1e-30
    rng = np.random.default_rng(1)
    for dof in rng.choice(n, 5, replace=False):
        Uc = U.astype(complex)
        Uc.reshape(-1)[dof] += 1j * h
        Rc, _ = el.structural_force(dNdX, w, conn, Uc, mat, onlyforce=True)
        col = np.zeros(n, dtype=complex)
        np.add.at(col, el.force_indices(conn), Rc)
        assert np.abs(col.imag / h - Kd[:, dof]).max() <= 1e-10 * np.abs(Kd).max()


def test_group_pattern_matches_bruteforce():
    m = meshes()[0]
    ind, inv = el.group_pattern(m['conn'])
    rows, cols = el.coo_indices(m['conn'])
    assert np.array_equal(ind[0][inv], rows) and np.array_equal(ind[1][inv], cols)
    keys = ind[0] * (rows.max() + 1) + ind[1]
    assert np.all(np.diff(keys) > 0)


def test_void_block_mesh_is_conforming_and_valid():
    from fea_amd import meshgen
    m = meshgen.block_with_void(4, 2, 2, seed=3, quadratic=False)
    T = m['nodes'][m['conn']]
    vol = np.einsum('ei,ei->e', T[:, 1] - T[:, 0], np.cross(T[:, 2] - T[:, 0], T[:, 3] - T[:, 0])) / 6
    assert vol.min() > 0
    faces = np.sort(np.concatenate([m['conn'][:, f] for f in ([0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3])]), axis=1)
    u, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2                                   # every face shared by at most two tets
    P = m['nodes'][u[cnt == 1]].mean(1)
    on_box = np.any((np.abs(P) < 1e-9) | (np.abs(P - [2.0, 1.0, 1.0]) < 1e-9), axis=1)
    n_void = sum(len(i) for i, _ in m['face_sets']['void'])
    assert (~on_box).sum() == n_void > 0                    # the rest of the boundary is the void
    # volume: block minus the (perturbed) void, within the faceting error
    assert abs(vol.sum() - (2.0 - 4.0 / 3.0 * np.pi * 0.3 ** 3)) < 0.02
    v = m['node_sets']['void']
    c, vd = m['void']['centre'], m['void']
    d = m['nodes'][v] - c
    r = np.linalg.norm(d, axis=1)
    rr = meshgen._void_radius(d / r[:, None], vd['r0'], vd['amp'], vd['waves'], vd['phases'])
    assert np.abs(r - rr).max() < 1e-12                     # void corner nodes on the perturbed sphere
    assert np.ptp(r) > 0.01                                  # and the perturbation is there
    m10 = meshgen.block_with_void(4, 2, 2, seed=3, quadratic=True)
    _, wdet, _ = el.pre_load_gaussian('C3D10', m10['nodes'], m10['conn'])
    assert wdet.min() > 0                                    # every C3D10 Gauss point valid
