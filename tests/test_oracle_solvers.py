"""Pin the oracle's solver restatements: CG vs numpy.linalg.solve, Newton
residual at the solution, lumped mass vs dense brute force and rho*V, explicit
central difference vs the closed form of an undamped 1-DOF oscillator."""
import numpy as np
import scipy.sparse as sp

from fea_amd import meshgen
from oracle import assembly as oasm
from oracle import elements as oel
from oracle import solvers as osol
from tests.helpers import nh, oracle_model


def test_cg_matches_dense_solve():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((40, 40))
    A = A @ A.T + 40 * np.eye(40)
    b = rng.standard_normal(40)
    x, its = osol.conjugate_gradient(sp.csr_matrix(A), b, tol=1e-28, max_iter=400)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_newton_converges_to_zero_residual():
    m = meshgen.kuhn_tet_block(2, 2, 2, jitter=0.05, seed=1)
    om = oracle_model(m, nh(1.0, 20.0), 'z0', pressure=0.1, load_set='z1')
    GC, its = osol.static_solve(om, increments=2, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12)
    R = oasm.assemble_force(om, oasm.gc_to_u(om, GC))
    r_ref = np.linalg.norm(oasm.assemble_force(om, np.zeros_like(om['nodes'])))
    assert np.linalg.norm(R) <= 1e-10 * r_ref
    # the CG-based Newton reaches the same point
    GC2, _ = osol.static_solve(om, increments=2, rel_tol=1e-10, linear='cg', cg_tol=1e-20,
                               armijo_slack_rel=1e-12)
    assert np.abs(GC2 - GC).max() <= 1e-8 * np.abs(GC).max()


def test_lumped_mass_bruteforce_and_total():
    m = meshgen.kuhn_tet_block(2, 2, 2, jitter=0.05, seed=2)
    om = oracle_model(m, nh(1.0, 20.0), 'z0', rho=2.5)
    oasm.prepare(om)
    g = om['groups'][0]
    vals = oel.mass_values(g['N'], g['wdet'], 2.5)
    r, c = oel.coo_indices(g['conn'])
    n = om['nodes'].size
    M = np.zeros((n, n))
    np.add.at(M, (r, c), vals)
    free = oasm.free_mask(om)
    Mr = M[np.ix_(free, free)]
    assert np.allclose(osol.lumped_mass(om), Mr.sum(1), rtol=1e-14)
    # without constraints every component carries rho * V
    om2 = dict(om, fixed=np.zeros_like(om['fixed']))
    assert abs(osol.lumped_mass(om2).sum() - 3 * 2.5 * 1.0) < 1e-12


def test_central_difference_oscillator():
    # undamped 1-DOF oscillator R(x) = w^2 x, m = 1, x0 = 1, v0 = 0: the central-difference
    # iterates are exactly x_n = cos(n theta) with cos(theta) = 1 - (w dt)^2 / 2
    w, dt, n = 2.0, 0.05, 200
    x, v, a = osol.central_difference(lambda g: w * w * g, np.ones(1), np.ones(1), np.zeros(1), dt, n)
    theta = np.arccos(1 - (w * dt) ** 2 / 2)
    assert abs(x[0] - np.cos(n * theta)) < 1e-12
