"""GPU parity of the element kernels against the oracle (reference einsum path).

FP64: residual / tangent within 1e-10 relative (max-norm over the group);
FP32 kernels within 2e-5 relative.  The kernels use an (S, Voigt C) total-
Lagrangian form, the oracle the reference's (P^T, dP^T/dF) einsums, so
agreement is between two independent derivations.
"""
import numpy as np
import pytest
import torch

from fea_amd import meshgen
from fea_amd.assemble import Part, elements, materials
from oracle import elements as oel
from tests.helpers import rel_err, mr, nh

pytestmark = pytest.mark.gpu


def mk(etype, seed):
    if etype == 'C3D4':
        m = meshgen.kuhn_tet_block(3, 2, 2, jitter=0.08, seed=seed)
    elif etype == 'C3D8':
        m = meshgen.hex_block(3, 2, 2, size=(1.5, 1.0, 1.0), jitter=0.08, seed=seed)
    else:
        m = meshgen.tet10_from_tet4(meshgen.kuhn_tet_block(2, 2, 1, jitter=0.08, seed=seed))
    return m


MATS = {'NH': (lambda: materials.NeoHookean(1.0, 100.0), lambda: nh(1.0, 100.0)),
        'MR': (lambda: materials.MooneyRivlin(0.5, 0.1, 0.01), lambda: mr(0.5, 0.1, 0.01))}


def element_group(m, mat):
    part = Part(torch.as_tensor(m['nodes']))
    e = elements.initialize_element(m['etype'], torch.arange(m['conn'].shape[0]), torch.as_tensor(m['conn']), part)
    e.set_materials(mat)
    e.initialize()
    return e


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
@pytest.mark.parametrize('mname', ['NH', 'MR'])
def test_structural_force_fp64(gpu, etype, mname):
    m = mk(etype, 7)
    e = element_group(m, MATS[mname][0]())
    omat = MATS[mname][1]()
    rng = np.random.default_rng(3)
    U = 0.05 * rng.standard_normal(m['nodes'].shape)
    idx_f, Rv, idx_k, Kv = e.structural_Force(torch.as_tensor(U))
    dNdX, w, _ = oel.pre_load_gaussian(m['etype'], m['nodes'], m['conn'])
    R_o, K_o = oel.structural_force(dNdX, w, m['conn'], U, omat)
    assert np.array_equal(idx_f.cpu().numpy(), oel.force_indices(m['conn']))
    assert rel_err(Rv.cpu().numpy(), R_o) < 1e-10
    ind, inv = oel.group_pattern(m['conn'])
    assert np.array_equal(idx_k.cpu().numpy(), ind)
    Kc = np.zeros(ind.shape[1])
    np.add.at(Kc, inv, K_o)
    assert rel_err(Kv.cpu().numpy(), Kc) < 1e-10
    assert np.array_equal(e._index_matrix_coalesce.cpu().numpy(), inv)


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
def test_gauss_geometry_and_energy(gpu, etype):
    m = mk(etype, 9)
    e = element_group(m, materials.MooneyRivlin(0.5, 0.1, 0.01))
    dNdX, w, N = oel.pre_load_gaussian(m['etype'], m['nodes'], m['conn'])
    assert rel_err(e.shape_function_d1_gaussian.cpu().numpy(), dNdX) < 1e-13
    assert rel_err(e.gaussian_weight.cpu().numpy(), w) < 1e-13
    assert rel_err(e.shape_function_d0_gaussian.cpu().numpy(), N) < 1e-13
    U = 0.05 * np.random.default_rng(2).standard_normal(m['nodes'].shape)
    E_gpu = float(e.potential_Energy(torch.as_tensor(U)))
    E_o = oel.potential_energy(dNdX, w, m['conn'], U, mr(0.5, 0.1, 0.01))
    assert abs(E_gpu - E_o) <= 1e-11 * abs(E_o)


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
def test_fp32_kernels(gpu, etype):
    from fea_amd.device import DevicePlan, Group
    from fea_amd._native import F32, F64
    m = mk(etype, 4)
    N = m['nodes'].shape[0]
    prm = materials.MooneyRivlin(0.5, 0.1, 0.01).kernel_params()
    plan = DevicePlan(m['nodes'], np.ones(3 * N, bool), [Group('elem', m['conn'].shape[1], m['conn'], params=prm)])
    U = torch.as_tensor(0.05 * np.random.default_rng(5).standard_normal(3 * N), device='cuda')
    R64, K64 = plan.tangent(U, 1.0, F64)
    R64, K64 = R64.clone(), K64.clone()
    _, K32 = plan.tangent(U, 1.0, F32, with_residual=False)
    R32 = plan.residual(U, 1.0, F32).clone()
    assert rel_err(K32.double().cpu().numpy(), K64.cpu().numpy()) < 2e-5
    assert rel_err(R32.cpu().numpy(), R64.cpu().numpy()) < 2e-5


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
@pytest.mark.parametrize('scale', [0.01, 0.03])
def test_residual_kernel_fp64(gpu, etype, scale):
    """The residual-only kernel (fea_elem_residual in FP64, the explicit loop's element pass)
    against the oracle's assembled internal force, at moderate and large strains (no element
    inverted: the oracle's forces are finite)."""
    from fea_amd.device import DevicePlan, Group
    from fea_amd._native import F64
    m = mk(etype, 6)
    N = m['nodes'].shape[0]
    plan = DevicePlan(m['nodes'], np.ones(3 * N, bool),
                      [Group('elem', m['conn'].shape[1], m['conn'], params=materials.NeoHookean(1.0, 100.0).kernel_params())])
    U = scale * np.random.default_rng(11).standard_normal((N, 3))
    R = plan.residual(torch.as_tensor(U.reshape(-1), device='cuda'), 1.0, F64).cpu().numpy()
    dNdX, w, _ = oel.pre_load_gaussian(m['etype'], m['nodes'], m['conn'])
    R_o, _ = oel.structural_force(dNdX, w, m['conn'], U, nh(1.0, 100.0))
    Ra = np.zeros(3 * N)
    np.add.at(Ra, oel.force_indices(m['conn']), R_o)
    assert np.isfinite(Ra).all()
    assert rel_err(R, Ra) < 1e-10


def test_rigid_rotation_is_stress_free(gpu):
    # u = (Q - I) X for a proper rotation Q: zero residual (objectivity of the kernels)
    m = mk('C3D8', 1)
    e = element_group(m, materials.MooneyRivlin(0.5, 0.1, 0.01))
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    U = m['nodes'] @ (Q - np.eye(3)).T
    _, Rv = e.structural_Force(torch.as_tensor(U), if_onlyforce=True)
    assert np.abs(Rv.cpu().numpy()).max() < 1e-12
