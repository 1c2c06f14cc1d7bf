"""GPU solver parity: Jacobi-PCG vs a direct solve, static Newton (FP64 and
mixed FP32/FP64-IR) vs the oracle's Newton with a direct linear solver
(converged displacement within 1e-8 relative), explicit central difference
vs the oracle's restatement of dynamic_explicit.py."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from fea_amd import solver, workloads
from fea_amd._native import F32, F64
from oracle import assembly as oasm
from oracle import solvers as osol
from tests.helpers import mr, nh, oracle_model, rel_err

pytestmark = pytest.mark.gpu


def cube_models(n=3, p=0.2):
    fe, mesh = workloads.cube_pressure(n, pressure=p, jitter=0.05, seed=2)
    om = oracle_model(mesh, nh(1.0, 100.0), 'z0', pressure=p, load_set='z1')
    return fe, om


def test_pcg_matches_direct_solve(gpu):
    fe, om = cube_models()
    fe.assembly.initialize()
    asm = fe.assembly
    p = asm.plan
    U = p.gc_to_u(torch.as_tensor(0.01 * np.random.default_rng(0).standard_normal(p.n_free), device='cuda'))
    R, K = p.tangent(U, 1.0, F64)
    A = sp.csr_matrix((K.cpu().numpy(), p.col_np, p.rowptr_np), shape=(p.n_free, p.n_free))
    b = -R.cpu().numpy()
    x_ref = spla.spsolve(A.tocsc(), b)
    pcg = solver.PCG(p.n_free, p.nnz, p.rowptr, p.col, p.diag_slot, F64, p.device, grp=p.grp,
                     n_groups=p.spmv_groups, spmv_col=p.spmv_col)
    pcg.set_matrix(K)
    x = pcg.solve(-R, None, tol=1e-24, max_iter=5000).cpu().numpy()
    assert rel_err(x, x_ref) < 1e-8
    assert pcg.iterations < 5000
    # the same iterates as the oracle's reference CG on the Jacobi-scaled system (x = xs / D)
    # after a fixed number of iterations -- to a bound that grows with k: on this system
    # (some 400 iterations to 1e-10) CG amplifies the rounding-order differences of its
    # reductions; a different algorithm is O(1) off -- and the same iteration count to 1e-6
    # (deeper into convergence the count itself drifts with the rounding order)
    D = np.sqrt(np.abs(A.diagonal()))
    As = sp.diags(1 / D) @ A @ sp.diags(1 / D)
    for k, bound in ((3, 1e-9), (10, 1e-5), (40, 1e-3)):
        xk = pcg.solve(-R, None, tol=0.0, max_iter=k).cpu().numpy()
        xs, _ = osol.conjugate_gradient(As, b / D, np.zeros_like(b), tol=0.0, max_iter=k)
        assert pcg.iterations == k
        err = rel_err(xk, xs / D)
        print('PCG iterate %d vs the oracle CG: %.2e' % (k, err))
        assert err < bound
    _, its = osol.conjugate_gradient(As, b / D, np.zeros_like(b), tol=1e-6, max_iter=5000)
    pcg.solve(-R, None, tol=1e-6, max_iter=5000)
    print('PCG iterations to 1e-6: %d, oracle CG %d' % (pcg.iterations, its))
    assert abs(pcg.iterations - its) <= max(3, 0.05 * its)


def partial_models():
    """The cube with a face held in x only and a face held in x and y: node groups of
    1, 2 and 3 free rows (not triple-structured)."""
    from fea_amd import boundarys
    fe, om = cube_models()
    fe.assembly.add_boundary(boundarys.Boundary_Condition('block', 'x0', indexDoF=[0]))
    fe.assembly.add_boundary(boundarys.Boundary_Condition('block', 'y1', indexDoF=[0, 1]))
    fe.assembly.initialize()
    return fe


@pytest.mark.parametrize('mode', ['triple', 'node', 'partial'])
@pytest.mark.parametrize('prec', ['fp64', 'mixed'])
def test_block_jacobi_pcg_matches_oracle(gpu, prec, mode):
    """Node-block Jacobi PCG (fea_block_jacobi_inv + precond 1) against the oracle's
    restatement after fixed iteration counts, and converged against a direct solve; with the
    row-per-thread stage of triple-structured systems, the group-per-thread one, and a model
    with 1- and 2-row node groups (column-checked block extraction, generic SpMV)."""
    if mode == 'partial':
        fe = partial_models()
    else:
        fe, om = cube_models()
        fe.assembly.initialize()
    p = fe.assembly.plan
    mat = F64 if prec == 'fp64' else F32
    U = p.gc_to_u(torch.as_tensor(0.01 * np.random.default_rng(3).standard_normal(p.n_free), device='cuda'))
    R, K = p.tangent(U, 1.0, F64, val_dtype=mat)
    A = sp.csr_matrix((K.double().cpu().numpy(), p.col_np, p.rowptr_np), shape=(p.n_free, p.n_free))
    b = -R.cpu().numpy()
    grp = p.grp.cpu().numpy()
    if mode == 'partial':
        assert set(np.diff(grp).tolist()) == {1, 2, 3}
        assert p.spmv_groups > 0                                 # not triple-structured
        ng = p.spmv_groups
    else:
        assert np.diff(grp).max() == 3 and np.diff(grp).min() >= 1
        assert p.spmv_groups < 0                                 # triple-structured
        ng = p.spmv_groups if mode == 'triple' else -p.spmv_groups
    pcg = solver.PCG(p.n_free, p.nnz, p.rowptr, p.col, p.diag_slot, mat, p.device, vec_dtype=F64, grp=p.grp,
                     n_groups=ng, chunk=16, precond='block', spmv_col=p.spmv_col)
    pcg.set_matrix(K)
    for k in (1, 7, 25):
        x = pcg.solve(-R, None, tol=0.0, max_iter=k).cpu().numpy()
        y, _ = osol.pcg_block_jacobi(A, b, grp, tol=0.0, max_iter=k)
        assert pcg.iterations == k
        assert rel_err(x, y) < 1e-9
    x = pcg.solve(-R, None, tol=1e-24, max_iter=5000).cpu().numpy()
    assert rel_err(x, spla.spsolve(A.tocsc(), b)) < 1e-8
    # the iteration count to 1e-10 on the tangent at u = 0 (positive definite; at the random
    # state above, near indefinite, the count drifts with the rounding order of the
    # reductions, while the iterates above pin the algorithm)
    R0, K0 = p.tangent(torch.zeros_like(U), 1.0, F64, val_dtype=mat)
    A0 = sp.csr_matrix((K0.double().cpu().numpy(), p.col_np, p.rowptr_np), shape=(p.n_free, p.n_free))
    pcg.set_matrix(K0)
    pcg.solve(-R0, None, tol=1e-10, max_iter=5000)
    _, its = osol.pcg_block_jacobi(A0, -R0.cpu().numpy(), grp, tol=1e-10, max_iter=5000)
    print('block-Jacobi PCG iterations to 1e-10 at u = 0: %d, oracle %d' % (pcg.iterations, its))
    assert abs(pcg.iterations - its) <= max(3, 0.05 * its)


@pytest.mark.parametrize('model', ['cube', 'partial'])
@pytest.mark.parametrize('prec', ['fp64', 'mixed'])
def test_block_jacobi_inverse_and_fallback_match_oracle(gpu, prec, model):
    """fea_block_jacobi_inv against the oracle's restatement on every node block, with some
    blocks made indefinite (a negated diagonal entry): those fall back to 1 / |K_ii|. Also on
    a model with 1- and 2-row node groups (blocks confirmed against the CSR columns)."""
    from fea_amd import _native as nat
    from fea_amd._native import call
    if model == 'partial':
        fe = partial_models()
    else:
        fe, om = cube_models()
        fe.assembly.initialize()
    p = fe.assembly.plan
    mat = F64 if prec == 'fp64' else F32
    U = p.gc_to_u(torch.as_tensor(0.01 * np.random.default_rng(4).standard_normal(p.n_free), device='cuda'))
    _, K = p.tangent(U, 1.0, F64, val_dtype=mat)
    K = K.clone()
    grp = p.grp.cpu().numpy()
    flip = grp[:-1][::7]                                    # the first row of every 7th group
    ds = p.diag_slot.cpu().numpy()
    K[torch.as_tensor(ds[flip], device='cuda', dtype=torch.int64)] *= -1
    binv = torch.zeros(6 * p.n_groups, dtype=torch.float32, device='cuda')
    call('fea_block_jacobi_inv', mat, p.n_free, p.spmv_groups, p.grp, p.rowptr, p.col, p.diag_slot, K, binv,
         nat.stream_ptr())
    A = sp.csr_matrix((K.double().cpu().numpy(), p.col_np, p.rowptr_np), shape=(p.n_free, p.n_free))
    M = osol.block_jacobi_inverse(A, grp)
    iu = ([0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2])
    got = binv.cpu().numpy().reshape(-1, 6).astype(np.float64)
    want = M[:, iu[0], iu[1]]
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()
    # the flipped groups took the point-Jacobi fallback (diagonal only)
    fl = np.arange(0, p.n_groups, 7)
    assert np.all(got[fl][:, [1, 2, 4]] == 0.0)
    assert np.allclose(got[fl, 0], 1.0 / np.abs(A.diagonal()[grp[fl]]), rtol=1e-6)


@pytest.mark.parametrize('predictor', [True, 'auto'])
def test_static_block_jacobi_matches_oracle(gpu, predictor):
    """The bench's solver configuration (mixed precision, node-block Jacobi, Eisenstat-Walker
    forcing, the 'auto' predictor) lands on the oracle's equilibrium."""
    fe, mesh = workloads.beam_traction(8, 2, 2, traction=(0.0, 0.0, -0.01))
    fe.solver = solver.StaticImplicitSolver(precision='mixed', rel_tol=1e-10, load_steps=4, cg_tol=1e-12,
                                            armijo_slack=1e-12, preconditioner='block', predictor=predictor,
                                            forcing='ew' if predictor == 'auto' else 'fixed')
    GC = fe.solve()
    assert GC is not False
    om = oracle_model(mesh, mr(0.5, 0.1, 0.01), 'x0', traction=(0.0, 0.0, -0.01), load_set='x1')
    GC_o, _ = osol.static_solve(om, increments=4, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12)
    assert rel_err(GC.cpu().numpy(), GC_o) < 1e-8


def test_reference_cg_signature(gpu):
    A = sp.random(60, 60, density=0.1, random_state=1)
    A = (A @ A.T + 60 * sp.eye(60)).tocoo()
    b = np.random.default_rng(1).standard_normal(60)
    x = solver.conjugate_gradient(torch.as_tensor(np.stack([A.row, A.col])), torch.as_tensor(A.data),
                                  torch.as_tensor(b), tol=1e-24, max_iter=500)
    assert rel_err(x.cpu().numpy(), np.linalg.solve(A.toarray(), b)) < 1e-9


@pytest.mark.parametrize('precision', ['fp64', 'mixed', 'mixed-f32', 'mixed-ir'])
def test_static_cube_matches_oracle(gpu, precision):
    fe, om = cube_models(3, 0.2)
    fe.solver = solver.StaticImplicitSolver(precision=precision, rel_tol=1e-10, load_steps=5, cg_tol=1e-12,
                                            armijo_slack=1e-12)
    GC = fe.solve()
    assert GC is not False
    GC_o, _ = osol.static_solve(om, increments=5, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12)
    assert GC_o is not False
    assert rel_err(GC.cpu().numpy(), GC_o) < 1e-8


def test_static_beam_mixed_matches_fp64(gpu):
    res = {}
    for prec in ('fp64', 'mixed', 'mixed-f32', 'mixed-ir'):
        fe, mesh = workloads.beam_traction(8, 2, 2, traction=(0.0, 0.0, -0.01))
        fe.solver = solver.StaticImplicitSolver(precision=prec, rel_tol=1e-10, load_steps=4, cg_tol=1e-12,
                                                armijo_slack=1e-12)
        res[prec] = fe.solve().cpu().numpy()
    om = oracle_model(mesh, mr(0.5, 0.1, 0.01), 'x0', traction=(0.0, 0.0, -0.01), load_set='x1')
    GC_o, _ = osol.static_solve(om, increments=4, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12)
    assert rel_err(res['fp64'], GC_o) < 1e-8
    assert rel_err(res['mixed'], GC_o) < 1e-8
    assert rel_err(res['mixed-ir'], GC_o) < 1e-8
    assert rel_err(res['mixed-f32'], GC_o) < 1e-8


@pytest.mark.parametrize('predictor', ['quadratic', 'cubic', 5, 'auto', False])
def test_static_predictor_orders_match_oracle(gpu, predictor):
    """Higher-order load-increment predictors start each increment elsewhere but converge
    to the oracle's equilibrium (the oracle runs the same extrapolation and, for 'auto',
    picks the same order per increment)."""
    fe, mesh = workloads.beam_traction(8, 2, 2, traction=(0.0, 0.0, -0.01))
    fe.solver = solver.StaticImplicitSolver(precision='mixed', rel_tol=1e-10, load_steps=5, cg_tol=1e-12,
                                            armijo_slack=1e-12, predictor=predictor)
    GC = fe.solve()
    assert GC is not False
    om = oracle_model(mesh, mr(0.5, 0.1, 0.01), 'x0', traction=(0.0, 0.0, -0.01), load_set='x1')
    orders = []
    GC_o, its_o = osol.static_solve(om, increments=5, predictor=predictor, orders=orders, rel_tol=1e-10,
                                    linear='direct', armijo_slack_rel=1e-12)
    assert rel_err(GC.cpu().numpy(), GC_o) < 1e-8
    assert fe.solver.stats['predictor_orders'] == orders
    print('predictor %s: orders %s newton %d (oracle %d)' % (predictor, orders, fe.solver.stats['newton_iters'],
                                                             its_o))


def test_reference_convergence_rule(gpu):
    # default constructor: absolute max-norm test of static_implicit.py:290
    fe, om = cube_models(2, 0.05)
    fe.solver = solver.StaticImplicitSolver()
    GC = fe.solve()
    assert GC is not False
    R = fe.assembly.assemble_force(GC=GC)
    assert float(R.abs().max()) < 1e-5


def test_explicit_matches_oracle(gpu):
    fe, om = cube_models(2, 0.05)
    dt, n = 2e-3, 25
    fe.solver = solver.DynamicExplicitSolver(time_end=dt * n, time_per_storage=1.0, dt=dt)
    fe.initialize()
    GC0 = torch.zeros_like(fe.assembly.GC)
    fe.solver.solve(GC0=GC0)
    GC_o, GV_o, GA_o = osol.explicit_steps(om, np.zeros(GC0.shape[0]), np.zeros(GC0.shape[0]), dt, n)
    assert rel_err(fe.solver.GC.cpu().numpy(), GC_o) < 1e-10
    m = fe.assembly.plan.lumped_mass().cpu().numpy()
    assert rel_err(m, osol.lumped_mass(om)) < 1e-13


@pytest.mark.parametrize('pair', ['f64/f64', 'f32/f64', 'f32/f32'])
@pytest.mark.parametrize('grouped', [True, False])
def test_spmv_precision_pairs(gpu, pair, grouped):
    from fea_amd._native import F32, call
    from fea_amd import _native as nat
    fe, om = cube_models(3, 0.2)
    fe.assembly.initialize()
    p = fe.assembly.plan
    U = p.gc_to_u(torch.as_tensor(0.01 * np.random.default_rng(3).standard_normal(p.n_free), device='cuda'))
    _, K = p.tangent(U, 1.0, F64)
    A = sp.csr_matrix((K.cpu().numpy(), p.col_np, p.rowptr_np), shape=(p.n_free, p.n_free))
    x = np.random.default_rng(4).standard_normal(p.n_free)
    m, v = [F64 if s == 'f64' else F32 for s in pair.split('/')]
    tm = torch.float64 if m == F64 else torch.float32
    tv = torch.float64 if v == F64 else torch.float32
    val = K.to(tm)
    xd = torch.as_tensor(x, device='cuda', dtype=tv)
    y = torch.zeros_like(xd)
    call('fea_spmv', m, v, p.n_free, p.nnz, p.spmv_groups if grouped else 0, p.grp if grouped else None,
         p.rowptr, p.spmv_col if grouped else p.col, val, xd, y, nat.stream_ptr())
    ref = A @ x
    tol = 1e-13 if pair == 'f64/f64' else (1e-6 if pair == 'f32/f64' else 1e-5)
    assert rel_err(y.double().cpu().numpy(), ref) < tol


def test_spmv_node_modes_agree(gpu):
    """Generic node-grouped and triple-blocked SpMV give the row-CSR result (beam: triple)."""
    from fea_amd._native import F32, call
    from fea_amd import _native as nat
    fe, mesh = workloads.beam_traction(6, 3, 2)
    fe.assembly.initialize()
    p = fe.assembly.plan
    assert p.triple
    U = p.gc_to_u(torch.as_tensor(0.01 * np.random.default_rng(5).standard_normal(p.n_free), device='cuda'))
    _, K = p.tangent(U, 1.0, F64)
    x = torch.as_tensor(np.random.default_rng(6).standard_normal(p.n_free), device='cuda')
    outs = []
    # the triple-blocked mode reads the compact per-triple columns (a third of the CSR's)
    assert p.spmv_col.numel() * 9 == p.nnz
    for ng, grp, col in ((0, None, p.col), (p.n_groups, p.grp, p.col), (-p.n_groups, p.grp, p.spmv_col)):
        y = torch.zeros_like(x)
        call('fea_spmv', F64, F64, p.n_free, p.nnz, ng, grp, p.rowptr, col, K, x, y, nat.stream_ptr())
        outs.append(y.cpu().numpy())
    assert rel_err(outs[1], outs[0]) < 1e-14 and rel_err(outs[2], outs[0]) < 1e-14


def test_ew_forcing_same_solution(gpu):
    fe, om = cube_models(3, 0.2)
    fe.solver = solver.StaticImplicitSolver(rel_tol=1e-10, load_steps=3, cg_tol=1e-12, armijo_slack=1e-12,
                                            forcing='ew')
    GC = fe.solve()
    GC_o, _ = osol.static_solve(om, increments=3, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12)
    assert rel_err(GC.cpu().numpy(), GC_o) < 1e-8
