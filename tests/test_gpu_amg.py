"""GPU parity of the smoothed-aggregation multigrid preconditioner.

* The device V-cycle (float operators, binary16 prolongators, float64 vectors) against the
  oracle's V-cycle (oracle/amg.py, float64, the prolongators rounded as stored) on the same
  host-built hierarchy, with the device's lambda_max estimate, for float64 and
  float32 matrices, on a brick beam and on a tet cube with 1-, 2- and 3-row node groups.
* The device power iteration against a dense eigensolve of D^-1 A.
* AMG-PCG: converged against a direct solve, its iteration count against the oracle's
  PCG with the oracle V-cycle, and a stopped solve's iterate against the oracle's.
* Static Newton solves with preconditioner='amg' land on the oracle's direct-solve
  Newton equilibrium (1e-8), in far fewer CG iterations than node-block Jacobi.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from fea_amd import boundarys, solver, workloads
from fea_amd._native import F32, F64
from fea_amd.solver import amg as A
from oracle import amg as oamg
from oracle import solvers as osol
from tests.helpers import mr, oracle_model, rel_err

pytestmark = pytest.mark.gpu


def _model(kind, half=False):
    if kind == 'beam':
        fe, _ = workloads.beam_traction(16, 4, 4, traction=(0.0, 0.0, -0.01))
    else:
        fe, _ = workloads.cube_pressure(4, pressure=0.2, jitter=0.05, seed=3)
        fe.assembly.add_boundary(boundarys.Boundary_Condition('block', 'x0', indexDoF=[0]))
        fe.assembly.add_boundary(boundarys.Boundary_Condition('block', 'y1', indexDoF=[0, 1]))
    fe.assembly.initialize()
    p = fe.assembly.plan
    U0 = torch.zeros(3 * p.n_nodes, dtype=torch.float64, device='cuda')
    _, K0 = p.tangent(U0, 1.0, F64)
    amg = A.AMG(p, K0.clone(), degree=2, max_coarse=60, half_level0=half)
    return fe, p, amg, K0.clone()


def _oracle_levels(p, K0):
    lv = A.host_hierarchy(p.rowptr_np, p.col_np, K0.cpu().numpy(), p.grp.cpu().numpy(),
                          *A.dof_geometry(p.dof_map_np, p.X.cpu().numpy()), max_coarse=60)
    for l in range(1, len(lv) - 1):
        lv[l]['Dinv'] = sp.block_diag(list(lv[l]['dinv']), format='csr')
    for L in lv[:-1]:                    # the device stores the prolongators in binary16
        L['P'] = sp.csr_matrix((L['P'].data.astype(np.float16).astype(np.float64), L['P'].indices, L['P'].indptr),
                               shape=L['P'].shape)
    return lv


def _state(p, scale, seed=1):
    """A tangent away from the build state: a smooth displacement field."""
    X = p.X.cpu().numpy().reshape(-1, 3)
    U = np.zeros_like(X)
    U[:, 2] = -scale * X[:, 0] ** 2
    U[:, 0] = scale * 0.3 * X[:, 0] * X[:, 2]
    return torch.as_tensor(U.reshape(-1), device='cuda').contiguous()


def _half(p, K, amg):
    """The level-0 matrix the device's V-cycle smooths with: the binary16 copy of K on
    triple-structured matrices (amg.half_level0), K itself otherwise."""
    v = K.double().cpu().numpy()
    return sp.csr_matrix((oamg.half_values(v) if amg.half_level0 else v, p.col_np, p.rowptr_np),
                         shape=(p.n_free, p.n_free))


def _binv(p, K, mat):
    from fea_amd import _native as nat
    b = torch.zeros(6 * abs(p.spmv_groups), dtype=torch.float32, device='cuda')
    nat.ops().block_jacobi_inv(mat, p.spmv_groups, p.grp, p.rowptr, p.col, p.diag_slot, K, b)
    return b


@pytest.mark.parametrize('kind', ['beam', 'cube'])
@pytest.mark.parametrize('mat', [F64, F32])
@pytest.mark.parametrize('half', [False, True])
def test_vcycle_matches_oracle(gpu, kind, mat, half):
    """The device V-cycle against the oracle's on the same hierarchy; half: level 0 smooths
    with the binary16 copy of the tangent (triple-structured matrices only: the beam)."""
    fe, p, amg, K0 = _model(kind, half)
    if kind == 'cube':
        assert set(np.diff(p.grp.cpu().numpy()).tolist()) == {1, 2, 3}
    assert amg.half_level0 == (half and kind == 'beam')
    _, K = p.tangent(_state(p, 0.02), 1.0, F64, val_dtype=mat)
    binv = _binv(p, K, mat)
    rho = amg.estimate_rho(mat, K, binv)
    lv = _oracle_levels(p, K0)
    Kh = sp.csr_matrix((K.double().cpu().numpy(), p.col_np, p.rowptr_np), shape=(p.n_free, p.n_free))
    K16 = _half(p, K, amg)               # the level-0 smoothing's matrix
    D0 = oamg.group_block_inverse(Kh, p.grp.cpu().numpy())
    # the device's lambda_max against a dense eigensolve (power iteration: from below)
    lam = np.abs(np.linalg.eigvals((D0 @ Kh).toarray())).max()
    assert 0.9 * lam <= rho <= 1.001 * lam, (rho, lam)
    assert amg.n_levels == len(lv)
    rng = np.random.default_rng(5)
    for _ in range(2):
        b = rng.standard_normal(p.n_free)
        z = torch.zeros(p.n_free, dtype=torch.float64, device='cuda')
        amg.vcycle(mat, K, binv, torch.as_tensor(b, device='cuda'), z)
        z_o = oamg.vcycle(lv, K16, D0, rho, b, 2)
        assert rel_err(z.cpu().numpy(), z_o) < 1e-5


@pytest.mark.parametrize('kind', ['beam', 'cube'])
@pytest.mark.parametrize('mat', [F64, F32])
@pytest.mark.parametrize('half', [False, True])
def test_amg_pcg_matches_direct_and_oracle(gpu, kind, mat, half):
    fe, p, amg, K0 = _model(kind, half)
    R, K = p.tangent(_state(p, 0.02), 1.0, F64, val_dtype=mat)
    pcg = solver.PCG(p.n_free, p.nnz, p.rowptr, p.col, p.diag_slot, mat, p.device, vec_dtype=F64, grp=p.grp,
                     n_groups=p.spmv_groups, chunk=4, precond='amg', spmv_col=p.spmv_col, amg=amg)
    pcg.set_matrix(K)
    b = (-R).contiguous()
    x = pcg.solve(b, None, tol=1e-22, max_iter=400).cpu().numpy()
    Kh = sp.csr_matrix((K.double().cpu().numpy(), p.col_np, p.rowptr_np), shape=(p.n_free, p.n_free))
    bh = b.cpu().numpy()
    assert rel_err(x, spla.spsolve(Kh.tocsc(), bh)) < 1e-8
    assert pcg.iterations < 200
    # iteration count and an early iterate against the oracle PCG with the oracle V-cycle
    lv = _oracle_levels(p, K0)
    D0 = oamg.group_block_inverse(Kh, p.grp.cpu().numpy())
    K16 = _half(p, K, amg)
    Minv = lambda r: oamg.vcycle(lv, K16, D0, amg.rho0, r, 2)
    pcg.solve(b, None, tol=1e-12, max_iter=400)
    _, k_o = oamg.pcg(Kh, bh, Minv, tol=1e-12, max_iter=400)
    assert abs(pcg.iterations - k_o) <= max(2, 0.05 * k_o), (pcg.iterations, k_o)
    for k in (3, 10):
        xk = pcg.solve(b, None, tol=0.0, max_iter=k).cpu().numpy()
        xk_o, _ = oamg.pcg(Kh, bh, Minv, tol=0.0, max_iter=k)
        assert pcg.iterations == k
        assert rel_err(xk, xk_o) < 1e-4


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_static_amg_matches_oracle(gpu, precision):
    t = (0.0, 0.0, -0.01)
    fe, mesh = workloads.beam_traction(16, 4, 4, traction=t)
    fe.solver = solver.StaticImplicitSolver(precision=precision, rel_tol=1e-10, load_steps=4, cg_tol=1e-12,
                                            cg_max_iter=20000, armijo_slack=1e-12, preconditioner='amg',
                                            predictor='auto', forcing='ew')
    GC = fe.solve()
    assert GC is not False
    its_amg = fe.solver.stats['cg_iters']
    om = oracle_model(mesh, mr(0.5, 0.1, 0.01), 'x0', traction=t, load_set='x1')
    GC_o, _ = osol.static_solve(om, increments=4, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12)
    assert rel_err(GC.cpu().numpy(), GC_o) < 1e-8
    fe2, _ = workloads.beam_traction(16, 4, 4, traction=t)
    fe2.solver = solver.StaticImplicitSolver(precision=precision, rel_tol=1e-10, load_steps=4, cg_tol=1e-12,
                                             cg_max_iter=20000, armijo_slack=1e-12, preconditioner='block',
                                             predictor='auto', forcing='ew')
    assert fe2.solve() is not False
    print('static %s: CG iterations amg %d, block Jacobi %d' % (precision, its_amg, fe2.solver.stats['cg_iters']))
    assert its_amg * 4 < fe2.solver.stats['cg_iters']


@pytest.mark.parametrize('mat', [F64, F32])
def test_smoothing_step_matches_numpy(gpu, mat):
    """fea_amg_smooth0 (one level-0 Chebyshev step, the V-cycle's dominant kernel) against
    d' = c1 d + c2 M^-1 (b - A x), x' = x + d' with M the stored block-Jacobi inverses."""
    from fea_amd import _native as nat
    from fea_amd._native import call
    fe, p, amg, K0 = _model('beam')
    _, K = p.tangent(_state(p, 0.02), 1.0, F64, val_dtype=mat)
    binv = _binv(p, K, mat)
    rng = np.random.default_rng(9)
    n = p.n_free
    b, x, d = (torch.as_tensor(rng.standard_normal(n), device='cuda') for _ in range(3))
    xo = torch.zeros_like(x)
    dd = d.clone()
    call('fea_amg_smooth0', mat, n, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, K, binv, b, x, xo, dd, 0.7,
         0.3, nat.stream_ptr())
    Kh = sp.csr_matrix((K.double().cpu().numpy(), p.col_np, p.rowptr_np), shape=(n, n))
    grp = p.grp.cpu().numpy()
    B = binv.cpu().numpy().astype(np.float64).reshape(-1, 6)
    iu = ([0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2])
    Ms = np.zeros((len(grp) - 1, 3, 3))
    Ms[:, iu[0], iu[1]] = B
    Ms[:, iu[1], iu[0]] = B
    M = sp.block_diag([Ms[g, :grp[g + 1] - grp[g], :grp[g + 1] - grp[g]] for g in range(len(grp) - 1)], format='csr')
    bh, xh, dh = b.cpu().numpy(), x.cpu().numpy(), d.cpu().numpy()
    d_ref = 0.7 * dh + 0.3 * (M @ (bh - Kh @ xh))
    assert rel_err(dd.cpu().numpy(), d_ref) < 1e-12
    assert rel_err(xo.cpu().numpy(), xh + d_ref) < 1e-12
    # the first step from x = 0: d = xo = c2 M^-1 b
    call('fea_amg_smooth0', mat, n, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, K, binv, b, None, xo, dd, 0.0,
         0.3, nat.stream_ptr())
    assert rel_err(xo.cpu().numpy(), 0.3 * (M @ bh)) < 1e-12


@pytest.mark.parametrize('mat', [F64, F32])
def test_half_smoothing_step_matches_numpy(gpu, mat):
    """fea_amg_half_values + fea_amg_smooth0_half (the V-cycle's level-0 step on the binary16
    copy of the tangent) against the same step in numpy with oracle.amg.half_values(K): the
    copy's values bit for bit, the step to round-off."""
    from fea_amd import _native as nat
    from fea_amd._native import call
    fe, p, amg, K0 = _model('beam')
    _, K = p.tangent(_state(p, 0.02), 1.0, F64, val_dtype=mat)
    binv = _binv(p, K, mat)
    assert p.spmv_groups < 0                    # triple-structured (every free node has 3 DOFs)
    v16 = torch.empty(4 * p.nnz // 3, dtype=torch.int16, device='cuda')
    sc = torch.zeros(2, dtype=torch.float64, device='cuda')
    s = nat.stream_ptr()
    call('fea_amg_half_values', mat, p.nnz, K, v16, sc, s)
    torch.cuda.synchronize()
    k_o = oamg.half_values(K.double().cpu().numpy())
    quads = v16.cpu().numpy().view(np.float16).astype(np.float64).reshape(-1, 4)
    assert np.all(quads[:, 3] == 0.0)
    k_d = quads[:, :3].reshape(-1) * float(sc[1])
    assert np.array_equal(k_d, k_o)
    rng = np.random.default_rng(9)
    n = p.n_free
    b, x, d = (torch.as_tensor(rng.standard_normal(n), device='cuda') for _ in range(3))
    xo = torch.zeros_like(x)
    dd = d.clone()
    call('fea_amg_smooth0_half', n, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, v16, sc, binv, b, x, xo, dd,
         0.7, 0.3, s)
    Kh = sp.csr_matrix((k_o, p.col_np, p.rowptr_np), shape=(n, n))
    grp = p.grp.cpu().numpy()
    B = binv.cpu().numpy().astype(np.float64).reshape(-1, 6)
    iu = ([0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2])
    Ms = np.zeros((len(grp) - 1, 3, 3))
    Ms[:, iu[0], iu[1]] = B
    Ms[:, iu[1], iu[0]] = B
    M = sp.block_diag([Ms[g, :grp[g + 1] - grp[g], :grp[g + 1] - grp[g]] for g in range(len(grp) - 1)], format='csr')
    bh, xh, dh = b.cpu().numpy(), x.cpu().numpy(), d.cpu().numpy()
    d_ref = 0.7 * dh + 0.3 * (M @ (bh - Kh @ xh))
    assert rel_err(dd.cpu().numpy(), d_ref) < 1e-12
    assert rel_err(xo.cpu().numpy(), xh + d_ref) < 1e-12
