"""Pin the oracle's solver restatements: CG vs numpy.linalg.solve, Newton
residual at the solution, lumped mass vs dense brute force and rho*V, explicit
central difference vs the closed form of an undamped 1-DOF oscillator."""
import numpy as np
import pytest
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


def test_newmark_linear_oscillator_closed_form():
    """Newmark beta = 1/4, gamma = 1/2 (the trapezoidal rule) on m x'' + k x = 0:
    amplitude kept exactly and the discrete frequency 2/dt atan(w dt / 2)."""
    import scipy.sparse as sp
    m, k, dt, n = 2.0, 3.0, 0.3, 40
    out = osol.newmark_steps(lambda x: (k * x, sp.csr_matrix([[k]])), lambda x: 0.5 * k * float(x @ x),
                             sp.csr_matrix([[m]]), np.array([1.0]), np.array([0.0]), dt, n, tol_error=1e-14)
    w = np.sqrt(k / m)
    wd = 2.0 / dt * np.arctan(w * dt / 2.0)
    x = np.array([o[0][0] for o in out])
    assert np.abs(x - np.cos(wd * dt * np.arange(n + 1))).max() < 1e-12
    # the trapezoidal rule keeps the quadratic invariant m v^2 + k x^2 exactly
    v = np.array([o[1][0] for o in out])
    energy = 0.5 * m * v ** 2 + 0.5 * k * x ** 2
    assert np.abs(energy - energy[0]).max() < 1e-12 * energy[0]


@pytest.mark.parametrize('beta,gamma', [(0.25, 0.5), (0.3, 0.6), (0.3025, 0.6)])
def test_newmark_equation_of_motion_each_step(beta, gamma):
    """After every converged Newmark step the acceleration of :261-281 satisfies the
    equation of motion M GA + R(GC) = 0; with the predictor's dt^2 term written as
    (1/2 - beta) this holds for any beta (a literal 0.25 breaks it unless beta = 1/4).
    A 3-DOF nonlinear chain: R(x) = K x + c x^3 (elementwise), energy to match."""
    import scipy.sparse as sp
    K = np.array([[4.0, -1.5, 0.0], [-1.5, 3.0, -1.0], [0.0, -1.0, 2.5]])
    c = 0.7
    M = sp.csr_matrix(np.diag([1.3, 0.8, 2.0]) + 0.1 * np.ones((3, 3)))

    def rt(x):
        return K @ x + c * x ** 3, sp.csr_matrix(K + np.diag(3 * c * x ** 2))

    def en(x):
        return 0.5 * float(x @ K @ x) + 0.25 * c * float((x ** 4).sum())

    out = osol.newmark_steps(rt, en, M, np.array([0.4, -0.2, 0.1]), np.array([0.0, 0.5, -0.3]), 0.2, 15,
                             beta=beta, gamma=gamma, tol_error=1e-10)
    # M GA + R(GC) equals the incremental residual at Newton's last iterate (~1e-10
    # here, amplified 1/(beta dt^2) = 100 at most); the literal 0.25 leaves 0.46 at beta = 0.3
    for GC, _, GA in out:
        res = M @ GA + rt(GC)[0]
        assert np.abs(res).max() < 1e-7, res


def test_consistent_mass_closed_form():
    """Consistent mass of the reference's Gauss rules: C3D4 with its single point
    (tetrahedral.py:30-40) gives rho V / 16 for every node pair (rank one; rows sum
    to rho V / 4); C3D8's 2x2x2 rule integrates N_a N_b exactly, so on undistorted
    cells of size h the entry is rho h^3 prod over axes of 1/3 (same side) or 1/6."""
    from fea_amd import meshgen
    from tests.helpers import nh, oracle_model
    mesh = meshgen.kuhn_tet_block(2, 2, 2, jitter=0.1, seed=4)
    om = oracle_model(mesh, nh(1.0, 10.0), 'z0', rho=2.5)
    om['fixed'][:] = False
    M = osol.consistent_mass(om).toarray()
    X, conn = mesh['nodes'], mesh['conn']
    P = X[conn]
    V = np.abs(np.einsum('ei,ei->e', P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]))) / 6.0
    ref = np.zeros_like(M)
    for e in range(conn.shape[0]):
        for a in range(4):
            for b in range(4):
                for d in range(3):
                    ref[3 * conn[e, a] + d, 3 * conn[e, b] + d] += 2.5 * V[e] / 16.0
    assert np.abs(M - ref).max() < 1e-14 * np.abs(ref).max()
    assert abs(M.sum() - 3 * 2.5 * V.sum()) < 1e-12
    # C3D8, one undistorted cell of size h
    h = 0.5
    hm = meshgen.hex_block(1, 1, 1, size=(h, h, h))
    om8 = oracle_model(hm, nh(1.0, 10.0), 'z0', rho=2.0)
    om8['fixed'][:] = False
    M8 = osol.consistent_mass(om8).toarray()
    Xn = hm['nodes']
    ref8 = np.zeros_like(M8)
    for a in range(8):
        for b in range(8):
            w = np.prod([1.0 / 3.0 if abs(Xn[a, i] - Xn[b, i]) < 1e-12 else 1.0 / 6.0 for i in range(3)])
            for d in range(3):
                ref8[3 * a + d, 3 * b + d] = 2.0 * h ** 3 * w
    assert np.abs(M8 - ref8).max() < 1e-14 * np.abs(ref8).max()


def test_extrapolation_is_exact_on_polynomials():
    # degree-m Lagrange extrapolation reproduces any polynomial of degree <= m in the load level
    rng = np.random.default_rng(3)
    for order in (1, 2, 3):
        c = rng.standard_normal((order + 1, 5))
        f = lambda t: sum(c[d] * t ** d for d in range(order + 1))
        hist = [f(k) for k in (4.0, 3.0, 2.0, 1.0)]
        assert np.allclose(osol.extrapolate(hist, order), f(5.0), rtol=1e-12, atol=1e-12)
        # one degree higher is not reproduced
        g = lambda t: f(t) + t ** (order + 1)
        assert not np.allclose(osol.extrapolate([g(k) for k in (4.0, 3.0, 2.0, 1.0)], order), g(5.0))
    # with fewer states the degree drops: one state -> constant
    assert np.array_equal(osol.extrapolate([np.ones(3)], 3), np.ones(3))


def test_predictor_orders_reach_the_same_equilibrium():
    m = meshgen.kuhn_tet_block(2, 2, 2, jitter=0.05, seed=1)
    om = oracle_model(m, nh(1.0, 20.0), 'z0', pressure=0.1, load_set='z1')
    ref, _ = osol.static_solve(om, increments=4, predictor=False, rel_tol=1e-10, linear='direct',
                               armijo_slack_rel=1e-12)
    for pred in ('linear', 'quadratic', 'cubic'):
        GC, _ = osol.static_solve(om, increments=4, predictor=pred, rel_tol=1e-10, linear='direct',
                                  armijo_slack_rel=1e-12)
        assert np.abs(GC - ref).max() <= 1e-8 * np.abs(ref).max()


def test_auto_order_choice():
    # a smooth (analytic) path: every extra order misses less, so the choice climbs past the
    # highest testable order, up to the cap
    t = 0.1 * np.arange(9.0)[::-1]
    hist = [np.exp(t[j]) * np.ones(2) for j in range(9)]
    assert [osol.auto_order(hist[:h], 6) for h in (2, 3, 4, 5)] == [1, 2, 3, 4]
    assert osol.auto_order(hist, 6) == 6 and osol.auto_order(hist, 3) == 3
    # a zig-zag path: extrapolation amplifies the alternation, the constant guess misses least
    zig = [np.array([0.1 * j + (-1.0) ** j]) for j in range(6)][::-1]
    assert osol.auto_order(zig, 6) == 0


def test_auto_predictor_reaches_the_same_equilibrium():
    m = meshgen.kuhn_tet_block(2, 2, 2, jitter=0.05, seed=1)
    om = oracle_model(m, nh(1.0, 20.0), 'z0', pressure=0.1, load_set='z1')
    ref, its0 = osol.static_solve(om, increments=5, predictor=False, rel_tol=1e-10, linear='direct',
                                  armijo_slack_rel=1e-12)
    orders = []
    GC, its = osol.static_solve(om, increments=5, predictor='auto', orders=orders, rel_tol=1e-10,
                                linear='direct', armijo_slack_rel=1e-12)
    assert np.abs(GC - ref).max() <= 1e-8 * np.abs(ref).max()
    assert len(orders) == 4 and orders[0] == 1 and its < its0


def test_block_jacobi_pcg():
    # blocks of one row are the point Jacobi scaling: the same iterates as the reference's CG
    # on D^-1/2 A D^-1/2 (up to the float32 rounding of the stored inverse)
    rng = np.random.default_rng(5)
    Q = np.linalg.qr(rng.standard_normal((30, 30)))[0]
    A = Q @ np.diag(np.geomspace(1.0, 1e2, 30)) @ Q.T
    A = sp.csr_matrix(A + np.diag(rng.uniform(1, 5, 30)))
    b = rng.standard_normal(30)
    D = np.sqrt(np.abs(A.diagonal()))
    As = sp.diags(1 / D) @ A @ sp.diags(1 / D)
    for k in (1, 3, 10):
        x1, _ = osol.pcg_block_jacobi(A, b, np.arange(31), tol=0.0, max_iter=k)
        y, _ = osol.conjugate_gradient(As, b / D, tol=0.0, max_iter=k)
        assert np.abs(x1 - y / D).max() <= 1e-6 * np.abs(x1).max()
    # the stored 3x3 inverses invert the symmetric diagonal blocks (to float32), groups of 3, 3, 2, 1 ...
    grp = np.array([0, 3, 6, 8, 9, 12, 15, 18, 21, 24, 27, 30])
    M = osol.block_jacobi_inverse(A, grp)
    for g in range(len(grp) - 1):
        r0, nr = grp[g], grp[g + 1] - grp[g]
        B = A[r0:r0 + nr, r0:r0 + nr].toarray()
        assert np.allclose(M[g, :nr, :nr] @ B, np.eye(nr), atol=1e-5)
        assert np.all(M[g, nr:, :] == 0) and np.all(M[g, :, nr:] == 0)
    x, _ = osol.pcg_block_jacobi(A, b, grp, tol=1e-26, max_iter=500)
    assert np.allclose(x, np.linalg.solve(A.toarray(), b), rtol=1e-9, atol=1e-12)


def test_half_values_range_and_rounding():
    """oracle.amg.half_values: the largest scaled value lands in [2^14, 2^15) (inside the
    binary16 range), each value is within half a binary16 ulp (2^-11 relative) of the input
    where it stays normal, signs and zeros are kept, and a power-of-two scaling of the input
    scales the output exactly."""
    from oracle import amg as oamg
    rng = np.random.default_rng(4)
    for scale in (1e-3, 1.0, 7.5e4):
        v = scale * rng.standard_normal(5000)
        v[::97] = 0.0
        h = oamg.half_values(v)
        m = np.abs(v.astype(np.float32)).max()
        k = 15 - int(np.frexp(np.float32(m))[1])
        assert 2 ** 14 <= m * 2.0 ** k < 2 ** 15
        big = np.abs(v) * 2.0 ** k >= 2.0 ** -14                    # normal binary16
        assert np.all(np.abs(h[big] - v[big]) <= 2.0 ** -11 * np.abs(v[big]))
        assert np.all(h[v == 0.0] == 0.0) and np.all(np.sign(h[big]) == np.sign(v[big]))
        assert np.array_equal(oamg.half_values(8.0 * v), 8.0 * h)
