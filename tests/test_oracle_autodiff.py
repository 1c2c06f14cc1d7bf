"""Pins the autograd oracle used by the adjoint tests: its internal force equals
the numpy restatement of the reference's force (which the element tests pin),
its pressure Hessian-vector product equals the restated pressure stiffness,
and its vector-Jacobian products match central finite differences."""
import numpy as np
import pytest
import torch

from oracle import assembly as oasm
from oracle import autodiff as ad
from oracle import materials as omat
from oracle import surfaces as osurf
from fea_amd import meshgen
from tests.helpers import facets

MATS = {'nh': (omat.NeoHookean(0.8, 40.0), (0.4, 0.0, 40.0)),
        'mr': (omat.MooneyRivlin(0.3, 0.12, 0.05), (0.3, 0.12, 40.0))}


def small_mesh(etype, seed=0):
    if etype == 'C3D8':
        m = meshgen.hex_block(2, 2, 1, jitter=0.15, seed=seed)
    else:
        m = meshgen.kuhn_tet_block(2, 1, 1, jitter=0.15, seed=seed)
        if etype == 'C3D10':
            m = meshgen.tet10_from_tet4(m)
    rng = np.random.default_rng(seed)
    m['nodes'] = m['nodes'] + 0.03 * rng.standard_normal(m['nodes'].shape)   # boundary too
    return m


def model_of(mesh, material):
    N = mesh['nodes'].shape[0]
    return {'nodes': mesh['nodes'], 'fixed': np.zeros((N, 3), dtype=bool), 'pressures': [], 'tractions': [],
            'groups': [{'etype': mesh['etype'], 'conn': mesh['conn'], 'material': material}]}


def oracle_fint(mesh, material, U):
    return oasm.assemble_force(model_of(mesh, material), U).reshape(-1, 3)


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
@pytest.mark.parametrize('mat', ['nh', 'mr'])
def test_autograd_force_equals_restated_reference_force(etype, mat):
    mesh = small_mesh(etype)
    material, th = MATS[mat]
    U = 0.05 * np.random.default_rng(1).standard_normal(mesh['nodes'].shape)
    f = ad.internal_force(etype, mesh['conn'], torch.as_tensor(mesh['nodes']), torch.as_tensor(U), *th)
    ref = oracle_fint(mesh, material, U)
    assert np.abs(f.numpy() - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize('etype', ['C3D4', 'C3D10'])
def test_vjp_matches_central_differences(etype):
    mesh = small_mesh(etype, seed=2)
    material, th = MATS['mr']
    rng = np.random.default_rng(3)
    U = 0.05 * rng.standard_normal(mesh['nodes'].shape)
    lam = rng.standard_normal(U.shape)
    gX, gU, gth = ad.internal_force_vjp(etype, mesh['conn'], mesh['nodes'], U, th, lam)
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(U.shape)
        def phi(Xs, Us, mat=material):
            m = dict(mesh)
            m['nodes'] = Xs
            return np.sum(lam * oracle_fint(m, mat, Us))
        fdX = (phi(mesh['nodes'] + h * d, U) - phi(mesh['nodes'] - h * d, U)) / (2 * h)
        fdU = (phi(mesh['nodes'], U + h * d) - phi(mesh['nodes'], U - h * d)) / (2 * h)
        assert abs(fdX - np.sum(gX * d)) < 1e-6 * max(1.0, abs(fdX))
        assert abs(fdU - np.sum(gU * d)) < 1e-6 * max(1.0, abs(fdU))
    # material parameters: MooneyRivlin(C10, C01, D1) with kv = 2/D1
    c10, c01, d1 = 0.3, 0.12, 0.05
    def phi_p(c10_, c01_, d1_):
        return np.sum(lam * oracle_fint(mesh, omat.MooneyRivlin(c10_, c01_, d1_), U))
    hp = 1e-6
    fd10 = (phi_p(c10 + hp, c01, d1) - phi_p(c10 - hp, c01, d1)) / (2 * hp)
    fd01 = (phi_p(c10, c01 + hp, d1) - phi_p(c10, c01 - hp, d1)) / (2 * hp)
    fdD1 = (phi_p(c10, c01, d1 + hp * d1) - phi_p(c10, c01, d1 - hp * d1)) / (2 * hp * d1)
    assert abs(fd10 - gth[0]) < 1e-6 * max(1.0, abs(fd10))
    assert abs(fd01 - gth[1]) < 1e-6 * max(1.0, abs(fd01))
    assert abs(fdD1 - gth[2] * (-2.0 / d1 ** 2)) < 1e-6 * max(1.0, abs(fdD1))


@pytest.mark.parametrize('etype', ['C3D4', 'C3D10', 'C3D8'])
def test_pressure_hvp_equals_restated_stiffness(etype):
    mesh = small_mesh(etype, seed=4)
    conn, stype = facets(mesh, 'z1')
    rng = np.random.default_rng(5)
    x = mesh['nodes'] + 0.05 * rng.standard_normal(mesh['nodes'].shape)
    lam = rng.standard_normal(x.shape)
    p = 0.3
    hv = ad.pressure_hvp(stype, conn, x, p, lam)
    _, K = osurf.pressure_force_stiffness(stype, x, conn, p)
    idx = osurf.pressure_stiff_idx(conn)
    N = x.shape[0]
    Kd = np.zeros((3 * N, 3 * N))
    np.add.at(Kd, (idx[0], idx[1]), -K)             # the load's stiffness enters the tangent with - sign
    ref = (Kd @ lam.reshape(-1)).reshape(-1, 3)
    assert np.abs(hv - ref).max() < 1e-12 * np.abs(ref).max()
