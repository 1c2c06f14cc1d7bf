"""GPU parity of the hand-written adjoint: element VJPs (d/dX, d/dU, d/dtheta)
and the pressure Hessian-vector product against the torch-autograd oracle,
and the compliance shape / material gradient through the adjoint solve against
central finite differences of full Newton solves."""
import numpy as np
import pytest
import torch

from fea_amd import adjoint, solver, workloads
from fea_amd.device import DevicePlan, Group
from oracle import autodiff as ad
from tests.helpers import facets
from tests.test_oracle_autodiff import small_mesh

pytestmark = pytest.mark.gpu

PARAMS = {'nh': (0.4, 0.0, 40.0), 'mr': (0.3, 0.12, 40.0)}
NTYPE = {'C3D4': 4, 'C3D8': 8, 'C3D10': 10}


def plan_for(mesh, params, pressure=None, fixed=None):
    N = mesh['nodes'].shape[0]
    free = np.ones(3 * N, dtype=bool)
    if fixed is not None:
        free.reshape(-1, 3)[fixed] = False
    groups = [Group('elem', NTYPE[mesh['etype']], mesh['conn'], params=np.array(params))]
    if pressure is not None:
        conn, _ = facets(mesh, 'z1')

        class _P:
            pass
        ld = _P()
        ld.pressure = pressure
        groups.append(Group('pressure', conn.shape[1], conn, load=ld))
    return DevicePlan(mesh['nodes'], free, groups)


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
@pytest.mark.parametrize('mat', ['nh', 'mr'])
def test_elem_vjp_matches_autograd(gpu, etype, mat):
    mesh = small_mesh(etype, seed=7)
    th = PARAMS[mat]
    p = plan_for(mesh, th, fixed=mesh['node_sets']['x0'])      # constrained nodes keep their rows
    rng = np.random.default_rng(8)
    U = 0.05 * rng.standard_normal(mesh['nodes'].shape)
    lam = rng.standard_normal(U.shape)
    cu = lambda a: torch.as_tensor(a.reshape(-1), device='cuda').contiguous()
    gx, gu, gth = adjoint.elem_vjp(p, cu(U), cu(lam), want_x=True, want_u=True, want_theta=True)
    ox, ou, oth = ad.internal_force_vjp(etype, mesh['conn'], mesh['nodes'], U, th, lam)
    tol = 1e-11
    assert np.abs(gx.cpu().numpy() - ox.reshape(-1)).max() < tol * np.abs(ox).max()
    assert np.abs(gu.cpu().numpy() - ou.reshape(-1)).max() < tol * np.abs(ou).max()
    assert np.abs(gth.cpu().numpy()[0] - oth).max() < tol * np.abs(oth).max()


@pytest.mark.parametrize('etype', ['C3D4', 'C3D10', 'C3D8'])
def test_pressure_hvp_matches_autograd(gpu, etype):
    mesh = small_mesh(etype, seed=9)
    p = plan_for(mesh, PARAMS['nh'], pressure=0.3)
    rng = np.random.default_rng(10)
    U = 0.05 * rng.standard_normal(mesh['nodes'].shape)
    lam = rng.standard_normal(U.shape)
    cu = lambda a: torch.as_tensor(a.reshape(-1), device='cuda').contiguous()
    hv = adjoint.pressure_hvp(p, cu(U), cu(lam)).cpu().numpy()
    conn, stype = facets(mesh, 'z1')
    ref = ad.pressure_hvp(stype, conn, mesh['nodes'] + U, 0.3, lam).reshape(-1)
    assert np.abs(hv - ref).max() < 1e-12 * np.abs(ref).max()


def test_internal_force_function_backward(gpu):
    mesh = small_mesh('C3D10', seed=7)
    th = PARAMS['mr']
    p = plan_for(mesh, th)
    rng = np.random.default_rng(12)
    U0 = 0.02 * rng.standard_normal(mesh['nodes'].shape)
    X = torch.as_tensor(mesh['nodes'].reshape(-1), device='cuda').clone().requires_grad_(True)
    U = torch.as_tensor(U0.reshape(-1), device='cuda').clone().requires_grad_(True)
    theta = torch.tensor([th], dtype=torch.float64, requires_grad=True)
    f = adjoint.InternalForce.apply(X, U, theta, p)
    ref = ad.internal_force('C3D10', mesh['conn'], torch.as_tensor(mesh['nodes']), torch.as_tensor(U0), *th)
    assert torch.isfinite(ref).all()
    assert np.abs(f.detach().cpu().numpy() - ref.numpy().reshape(-1)).max() < 1e-12 * ref.abs().max().item()
    lam = rng.standard_normal(U0.shape)
    torch.sum(f * torch.as_tensor(lam.reshape(-1), device='cuda')).backward()
    ox, ou, oth = ad.internal_force_vjp('C3D10', mesh['conn'], mesh['nodes'], U0, th, lam)
    assert np.abs(X.grad.cpu().numpy() - ox.reshape(-1)).max() < 1e-11 * np.abs(ox).max()
    assert np.abs(U.grad.cpu().numpy() - ou.reshape(-1)).max() < 1e-11 * np.abs(ou).max()
    assert np.abs(theta.grad.numpy()[0] - oth).max() < 1e-11 * np.abs(oth).max()


def _void_model(m=3, k=2, ns=2, pressure=0.05):
    fe, mesh = workloads.void_block_pressure(m, k, ns, pressure=pressure, seed=1)
    fe.solver = solver.StaticImplicitSolver(precision='fp64', rel_tol=1e-12, load_steps=2, cg_tol=1e-26)
    fe.initialize()
    return fe, mesh


def test_compliance_shape_gradient_matches_finite_differences(gpu):
    fe, mesh = _void_model()
    J0, gX, gth = adjoint.compliance_gradient(fe)
    assert J0 > 0
    plan = fe.assembly.plan
    X0 = plan.X.reshape(-1).clone()
    rng = np.random.default_rng(13)
    h = 1e-6
    for _ in range(6):
        d = torch.as_tensor(rng.uniform(-1, 1, X0.numel()), device='cuda')
        Jp = adjoint.compliance(fe, X=X0 + h * d)[0].item()
        Jm = adjoint.compliance(fe, X=X0 - h * d)[0].item()
        fd = (Jp - Jm) / (2 * h)
        an = float(torch.dot(gX, d))
        assert abs(fd - an) < 2e-6 * abs(an) + 1e-9 * J0, (fd, an)
    plan.set_coordinates(X0)
    # a solve started from the base equilibrium at the full load lands on the same J
    GC_base = fe.assembly.GC.clone()
    d = torch.as_tensor(rng.uniform(-1, 1, X0.numel()), device='cuda')
    Jc = adjoint.compliance(fe, X=X0 + 1e-4 * d)[0].item()
    Jw = adjoint.compliance(fe, X=X0 + 1e-4 * d, GC0=GC_base)[0].item()
    assert abs(Jw - Jc) < 1e-11 * abs(Jc)
    plan.set_coordinates(X0)
    # material parameters: theta = (mu/2, 0, kappa)
    th0 = adjoint._theta_of(plan)
    for j, hj in ((0, 1e-6), (2, 1e-4)):
        e = torch.zeros_like(th0)
        e[0, j] = hj
        Jp = adjoint.compliance(fe, X=X0, theta=th0 + e)[0].item()
        Jm = adjoint.compliance(fe, X=X0, theta=th0 - e)[0].item()
        fd = (Jp - Jm) / (2 * hj)
        assert abs(fd - float(gth[0, j])) < 2e-6 * abs(fd) + 1e-9 * J0, (j, fd, float(gth[0, j]))
    adjoint._set_theta(plan, th0)


def test_multigrid_adjoint_gradient_matches_block_jacobi(gpu):
    """With preconditioner='amg' the forward solve and the adjoint solve share the model's
    multigrid hierarchy: the compliance and its shape / material gradients match those of
    the node-block Jacobi solves (both PCGs to 1e-26 on the scaled residual ratio)."""
    out = []
    for pc in ('block', 'amg'):
        fe, mesh = workloads.void_block_pressure(3, 2, 2, pressure=0.05, seed=1)
        fe.solver = solver.StaticImplicitSolver(precision='fp64', rel_tol=1e-12, load_steps=2, cg_tol=1e-26,
                                                preconditioner=pc)
        fe.initialize()
        J0, gX, gth = adjoint.compliance_gradient(fe)
        out.append((J0, gX.clone(), gth.clone()))
        if pc == 'amg':
            assert fe.solver.amg is not None and fe.solver.amg.n_levels >= 2
    (Jb, gXb, gthb), (Ja, gXa, gtha) = out
    assert abs(Ja - Jb) < 1e-12 * abs(Jb)
    assert float((gXa - gXb).abs().max()) < 1e-9 * float(gXb.abs().max())
    assert float((gtha - gthb).abs().max()) < 1e-9 * float(gthb.abs().max())
