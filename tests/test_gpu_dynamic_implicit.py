"""GPU parity of Newmark-beta implicit dynamics (reference: dynamic_implicit.py)
and the consistent mass (C3base.py:348-384, assembly.py:421-443): mass CSR
against the oracle (same pattern, 1e-13), Newmark steps against the oracle's
restatement with direct solves (1e-8), the initial acceleration M a = -R of a
loaded state, and energy conservation of the trapezoidal rule at small
amplitude."""
import numpy as np
import pytest
import torch

from fea_amd import meshgen, solver, workloads
from fea_amd.assemble import loads, materials
from oracle import assembly as oasm
from oracle import solvers as osol
from tests.helpers import nh, oracle_model, rel_err

pytestmark = pytest.mark.gpu


def model(etype='C3D8', load=None):
    if etype == 'C3D8':
        mesh = meshgen.hex_block(3, 2, 2, size=(1.5, 1.0, 1.0), jitter=0.1, seed=3)
    elif etype == 'C3D4':
        mesh = meshgen.kuhn_tet_block(3, 2, 2, size=(1.5, 1.0, 1.0), jitter=0.1, seed=3)
    else:
        mesh = meshgen.tet10_from_tet4(meshgen.kuhn_tet_block(2, 2, 2, jitter=0.1, seed=3))
    ld = None if load is None else (lambda nm: loads.Traction(nm, 'x1', load))
    fe = workloads._controller_from_mesh(mesh, materials.NeoHookean(1.0, 10.0), fixed_set='x0', load=ld, rho=1.3)
    om = oracle_model(mesh, nh(1.0, 10.0), 'x0', traction=load, load_set='x1' if load is not None else None, rho=1.3)
    return fe, om, mesh


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
def test_consistent_mass_matches_oracle(gpu, etype):
    fe, om, _ = model(etype)
    fe.assembly.initialize()
    p = fe.assembly.plan
    M = p.mass_matrix().cpu().numpy()
    M_o = osol.consistent_mass(om)
    M_o.sort_indices()
    assert np.array_equal(p.rowptr_np, M_o.indptr) and np.array_equal(p.col_np, M_o.indices)
    assert rel_err(M, M_o.data) < 1e-13


def _velocity(om, amp):
    X = om['nodes']
    V = np.zeros_like(X)
    V[:, 2] = amp * X[:, 0] ** 2                                     # bending-like, zero at x = 0
    V[:, 1] = 0.5 * amp * X[:, 0] * X[:, 2]
    return V.reshape(-1)[oasm.free_mask(om)]


@pytest.mark.parametrize('beta,gamma', [(0.25, 0.5), (0.3, 0.6)])
def test_newmark_matches_oracle(gpu, beta, gamma):
    fe, om, _ = model('C3D8')
    dt, n = 0.05, 6
    fe.solver = solver.DynamicImplicitSolver(deltaT=dt, time_end=n * dt, tol_error=1e-11, cg_tol=1e-26,
                                             cg_max_iter=20000, beta=beta, gamma=gamma)
    fe.initialize()
    GV0 = _velocity(om, 0.3)
    GC0 = np.zeros_like(GV0)
    assert fe.solver.solve(GC0=torch.as_tensor(GC0, device='cuda'), GV0=torch.as_tensor(GV0, device='cuda'))
    out = osol.newmark(om, GC0, GV0, dt, n, tol_error=1e-11, beta=beta, gamma=gamma)
    assert len(fe.solver._GC_list) == n + 1
    for k in (1, n):
        assert rel_err(fe.solver._GC_list[k].cpu().numpy(), out[k][0]) < 1e-8
        assert rel_err(fe.solver._GV_list[k].cpu().numpy(), out[k][1]) < 1e-7
    assert np.abs(out[n][0]).max() > 1e-2                            # the block moved
    # the equation of motion M GA + R(GC) = 0 holds at the last step
    p = fe.assembly.plan
    GC, GA = fe.solver._GC_list[n].contiguous(), fe.solver._GA_list[n].contiguous()
    MA = torch.empty_like(GA)
    p.spmv(p.mass_matrix(), GA, MA, 1)
    R = p.residual(p.gc_to_u(GC), 1.0, 1)
    assert float((MA + R).abs().max()) < 1e-6 * float(MA.abs().max())


def test_initial_acceleration_of_a_loaded_state(gpu):
    fe, om, _ = model('C3D8', load=(0.0, 0.0, -0.05))
    fe.solver = solver.DynamicImplicitSolver(deltaT=0.05, time_end=0.05)
    fe.initialize()
    GC0 = np.zeros(int(oasm.free_mask(om).sum()))
    a = fe.solver.get_current_acceleration(torch.as_tensor(GC0, device='cuda')).cpu().numpy()
    R0, _ = oasm.assemble_reduced(om, oasm.gc_to_u(om, GC0))
    import scipy.sparse.linalg as spla
    a_o = spla.spsolve(osol.consistent_mass(om).tocsc(), -R0)
    assert rel_err(a, a_o) < 1e-10
    # the load pulls the free end down: the mean z acceleration is negative
    assert a.reshape(-1, 3)[:, 2].mean() < 0


def test_energy_is_conserved_at_small_amplitude(gpu):
    """The trapezoidal rule conserves a linear system's energy exactly; at small
    amplitude the hyperelastic block is linear to O(amplitude)."""
    fe, om, _ = model('C3D8')
    dt, n = 0.05, 12
    # the incremental energy is ~1e-7 here and its float64 rounding ~1e-16 absolute: the
    # Armijo test gets a slack of that size, and |R| is converged to 1e-10 (energy error ~1e-20)
    fe.solver = solver.DynamicImplicitSolver(deltaT=dt, time_end=n * dt, tol_error=1e-10, cg_tol=1e-28,
                                             cg_max_iter=20000, armijo_slack=1e-8)
    fe.initialize()
    GV0 = _velocity(om, 1e-3)
    assert fe.solver.solve(GC0=torch.zeros(GV0.size, dtype=torch.float64, device='cuda'),
                           GV0=torch.as_tensor(GV0, device='cuda'))
    p = fe.assembly.plan
    M = p.mass_matrix()
    E = []
    for GC, GV in zip(fe.solver._GC_list, fe.solver._GV_list):
        MV = torch.empty_like(GV)
        p.spmv(M, GV.contiguous(), MV, 1)
        E.append(p.energy(p.gc_to_u(GC.contiguous())) + 0.5 * float(torch.dot(GV, MV)))
    E = np.array(E)
    assert np.abs(E - E[0]).max() < 1e-4 * E[0]
