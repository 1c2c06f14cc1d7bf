"""CPU checks of the multigrid preconditioner: the host set-up (fea_amg_build, no GPU) and
the oracle's V-cycle (oracle/amg.py) it is compared with on the GPU.

Set-up, on tangents of a Mooney-Rivlin brick beam and of a Kuhn-tet cube with partially
fixed faces (node groups of 1, 2 and 3 rows):
* every coarse operator is the Galerkin product P^T A P of the level above (with the
  exported float32 prolongator), the coarsest inverse inverts it;
* the six rigid-body modes of an unconstrained body lie in the range of the smoothed
  prolongator (up to the tiny shift that makes the body's tangent definite): the
  near-null space is built right, independently of how it is coded;
* the smoothers' block inverses invert the diagonal blocks, and rho bounds lambda_max of
  D^-1 A from below within a few percent (power iteration);
* the aggregates cover every node group.
Oracle V-cycle: the smoothing polynomial is the shifted Chebyshev polynomial (closed
form), the V-cycle is symmetric positive definite, a two-level cycle whose prolongator is
the identity is the exact inverse, and AMG-PCG reaches the direct solution in far fewer
iterations than block-Jacobi PCG.
"""
import numpy as np
import numpy.polynomial.chebyshev as C
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from fea_amd import meshgen
from fea_amd.solver import amg as A
from oracle import amg as oamg
from oracle import assembly as oasm
from tests.helpers import mr, nh, oracle_model


def _system(kind, fix_all=True, shift=0.0):
    if kind == 'beam':
        mesh = meshgen.hex_block(16, 4, 4, size=(4.0, 1.0, 1.0))
        om = oracle_model(mesh, mr(0.5, 0.1, 0.01), 'x0', traction=(0, 0, -0.01), load_set='x1')
    else:
        mesh = meshgen.kuhn_tet_block(4, 4, 4, jitter=0.05, seed=3)
        om = oracle_model(mesh, nh(1.0, 20.0), 'z0')
        if fix_all:
            om['fixed'][mesh['node_sets']['x0'], 0] = True              # a face held in x only
            om['fixed'][mesh['node_sets']['y1'], :2] = True             # a face held in x and y
    if not fix_all:
        om['fixed'][:] = False
    N = mesh['nodes'].shape[0]
    _, K = oasm.assemble_reduced(om, np.zeros((N, 3)))
    K = sp.csr_matrix(K)
    if shift:
        K = sp.csr_matrix(K + shift * np.abs(K.diagonal()).mean() * sp.eye(K.shape[0]))
    K.sort_indices()
    free = oasm.free_mask(om)
    dof_map = -np.ones(3 * N, dtype=np.int64)
    dof_map[free] = np.arange(int(free.sum()))
    node = np.nonzero(free)[0] // 3
    grp = np.concatenate([[0], np.nonzero(np.diff(node))[0] + 1, [node.size]]).astype(np.int32)
    xyz, comp = A.dof_geometry(dof_map, mesh['nodes'])
    return K, grp, xyz, comp, mesh


@pytest.mark.parametrize('kind', ['beam', 'cube'])
def test_galerkin_levels_and_coarse_inverse(kind):
    K, grp, xyz, comp, _ = _system(kind)
    lv = A.host_hierarchy(K.indptr, K.indices, K.data, grp, xyz, comp, max_coarse=60)
    assert len(lv) >= (3 if kind == 'beam' else 2)
    assert all(lv[l + 1]['n'] < lv[l]['n'] for l in range(len(lv) - 1))
    if kind == 'cube':
        assert set(np.diff(grp).tolist()) == {1, 2, 3}
    for l in range(len(lv) - 1):
        P = lv[l]['P']
        Al = K if l == 0 else lv[l]['A']
        G = (P.T @ Al @ P).toarray()
        Ac = lv[l + 1]['A'].toarray()
        used = np.abs(P).sum(0).A1 > 0
        assert np.abs(G - Ac)[np.ix_(used, used)].max() < 1e-6 * np.abs(Ac).max()
        # a coarse DOF without a prolongator column is an identity row
        unused = np.nonzero(~used)[0]
        assert np.all(Ac[unused][:, unused] == np.eye(unused.size))
        # every fine row reaches its own aggregate, so every aggregate gets rows
        assert np.all(np.diff(lv[l]['raw']['prow']) >= 1)
        assert np.all(np.diff(lv[l]['raw']['pt_ptr'][:lv[l]['nagg'] + 1]) >= 1)
    c = lv[-1]
    assert np.abs(c['dense_inv'] @ c['A'].toarray() - np.eye(c['n'])).max() < 1e-3


@pytest.mark.parametrize('kind', ['beam', 'cube'])
def test_rigid_body_modes_in_prolongator_range(kind):
    K, grp, xyz, comp, mesh = _system(kind, fix_all=False, shift=1e-10)
    lv = A.host_hierarchy(K.indptr, K.indices, K.data, grp, xyz, comp, max_coarse=60)
    x = xyz - xyz.mean(0)
    B = np.zeros((xyz.shape[0], 6))
    B[np.arange(comp.size), comp] = 1.0
    for k, (i, j) in enumerate(((0, 1), (1, 2), (2, 0))):   # rotations: e_i x - e_j y style
        rot = np.zeros_like(x)
        rot[:, i] = -x[:, j]
        rot[:, j] = x[:, i]
        B[:, 3 + k] = rot[np.arange(comp.size), comp]
    P = lv[0]['P'].toarray()
    coef, *_ = np.linalg.lstsq(P, B, rcond=None)
    assert np.abs(P @ coef - B).max() < 1e-5 * np.abs(B).max()
    # and the body's tangent annihilates them (the check has teeth)
    assert np.abs(K @ B).max() < 1e-6 * np.abs(K).max()


def test_smoother_blocks_and_rho():
    K, grp, xyz, comp, _ = _system('beam')
    lv = A.host_hierarchy(K.indptr, K.indices, K.data, grp, xyz, comp, max_coarse=60)
    D0 = oamg.group_block_inverse(K, grp)
    lam0 = np.abs(np.linalg.eigvals((D0 @ K).toarray())).max()
    assert 0.95 * lam0 <= lv[0]['rho'] <= (1.0 + 1e-9) * lam0      # 1.1 rho then bounds lambda_max
    for l in range(1, len(lv) - 1):
        Al = lv[l]['A'].toarray()
        for g in range(lv[l]['dinv'].shape[0]):
            blk = Al[6 * g:6 * g + 6, 6 * g:6 * g + 6]
            assert np.abs(lv[l]['dinv'][g] @ blk - np.eye(6)).max() < 1e-4
        Dl = sp.block_diag(list(lv[l]['dinv']), format='csr')
        lam = np.abs(np.linalg.eigvals(Dl @ Al)).max()
        assert 0.95 * lam <= lv[l]['rho'] <= 1.0001 * lam


def test_chebyshev_smoother_is_the_shifted_chebyshev_polynomial():
    """On A = diag(lambda), D = I, b = A x*: after k steps from 0 the error is
    x* T_k((theta - lambda) / delta) / T_k(theta / delta) (Saad, Alg. 12.1)."""
    rho = 2.0
    lam = np.linspace(rho / 30.0, rho, 9)
    Ad = sp.diags(lam)
    xs = np.linspace(1.0, 2.0, 9)
    lmax, lmin = 1.1 * rho, rho / 30.0
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    for k in (1, 2, 3, 4):
        x = oamg.chebyshev(Ad, sp.eye(9), rho, Ad @ xs, None, k)
        Tk = C.Chebyshev.basis(k)
        want = xs * Tk((theta - lam) / delta) / Tk(theta / delta)
        assert np.abs((xs - x) - want).max() < 1e-12


def test_oracle_vcycle_spd_and_exact_two_level():
    K, grp, xyz, comp, _ = _system('beam')
    lv = A.host_hierarchy(K.indptr, K.indices, K.data, grp, xyz, comp, max_coarse=60)
    for l in range(1, len(lv) - 1):
        lv[l]['Dinv'] = sp.block_diag(list(lv[l]['dinv']), format='csr')
    D0 = oamg.group_block_inverse(K, grp)
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal((2, K.shape[0]))
    V = lambda v: oamg.vcycle(lv, K, D0, lv[0]['rho'], v, 2)
    assert abs(x @ V(y) - y @ V(x)) < 1e-10 * np.linalg.norm(x) * np.linalg.norm(V(y))
    for _ in range(5):
        v = rng.standard_normal(K.shape[0])
        assert v @ V(v) > 0
    # two levels with P = I and the exact coarse inverse: the cycle is K^-1
    Kd = K.toarray()
    two = [{'P': sp.eye(K.shape[0], format='csr')}, {'A': K, 'dense_inv': np.linalg.inv(Kd), 'n': K.shape[0]}]
    b = rng.standard_normal(K.shape[0])
    z = oamg.vcycle(two, K, D0, lv[0]['rho'], b, 2)
    assert np.abs(Kd @ z - b).max() < 1e-8 * np.abs(b).max()


def test_oracle_amg_pcg_converges_in_far_fewer_iterations():
    K, grp, xyz, comp, _ = _system('beam')
    lv = A.host_hierarchy(K.indptr, K.indices, K.data, grp, xyz, comp, max_coarse=60)
    for l in range(1, len(lv) - 1):
        lv[l]['Dinv'] = sp.block_diag(list(lv[l]['dinv']), format='csr')
    D0 = oamg.group_block_inverse(K, grp)
    b = np.random.default_rng(1).standard_normal(K.shape[0])
    x_d = spla.spsolve(K.tocsc(), b)
    x, k = oamg.pcg(K, b, lambda r: oamg.vcycle(lv, K, D0, lv[0]['rho'], r, 2), tol=1e-20, max_iter=500)
    _, kj = oamg.pcg(K, b, lambda r: D0 @ r, tol=1e-20, max_iter=20000)
    assert np.abs(x - x_d).max() < 1e-8 * np.abs(x_d).max()
    assert k * 5 < kj, (k, kj)
