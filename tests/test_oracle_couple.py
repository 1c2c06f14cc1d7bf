"""Pins for the oracle's Couple / reference-point restatement (CPU).

The rotation and its derivatives are checked against the closed form of a rotation about
a coordinate axis and against central differences; the reduced residual and tangent
against differences of the total energy and of the residual, so a dropped term, a wrong
sign or a transposed index fails here.
"""
import numpy as np
import pytest

from fea_amd import meshgen
from oracle import couple as cp
from oracle import solvers as osol
from tests.helpers import nh, oracle_model, rel_err


def couple_model(etype='C3D8', moment=(0.0, 0.0, 0.02), force=(0.0, 0.01, 0.0), pressure=None, fixed_rp=None):
    if etype == 'C3D8':
        mesh = meshgen.hex_block(3, 2, 2, size=(1.5, 1.0, 1.0), jitter=0.1, seed=5)
    elif etype == 'C3D4':
        mesh = meshgen.kuhn_tet_block(3, 2, 2, size=(1.5, 1.0, 1.0), jitter=0.1, seed=5)
    else:
        mesh = meshgen.tet10_from_tet4(meshgen.kuhn_tet_block(2, 2, 2, size=(1.5, 1.0, 1.0), jitter=0.1, seed=5))
    om = oracle_model(mesh, nh(1.0, 10.0), 'x0', pressure=pressure, load_set='y1' if pressure else None)
    om['rps'] = [{'node': (1.9, 0.5, 0.45), 'force': force, 'moment': moment, 'fixed': fixed_rp}]
    om['couples'] = [{'nodes': np.asarray(mesh['node_sets']['x1']), 'rp': 0}]
    return om, mesh


def _rand_gc(om, scale, seed):
    rng = np.random.default_rng(seed)
    GC = scale * rng.standard_normal(cp.n_gc(om))
    GC[-3:] = rng.uniform(-0.4, 0.4, 3)                  # a finite rotation of the point
    return GC


def test_rotation_about_an_axis_is_the_textbook_matrix():
    rng = np.random.default_rng(0)
    y = rng.standard_normal((5, 3))
    for ang in (0.3, -1.2, 2.5):
        c, s = np.cos(ang), np.sin(ang)
        Rz = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
        assert rel_err(cp.rotate(np.array([0.0, 0.0, ang]), y), y @ Rz.T) < 1e-14
    z = rng.standard_normal(3)
    R = cp.rotate(z, np.eye(3)).T                        # columns R e_j
    assert np.abs(R.T @ R - np.eye(3)).max() < 1e-14 and abs(np.linalg.det(R) - 1.0) < 1e-14
    assert rel_err(cp.rotate(z, z[None]), z[None]) < 1e-14   # the axis is fixed


@pytest.mark.parametrize('zscale', [0.05, 0.7, 2.0])
def test_rotation_jacobians_match_central_differences(zscale):
    rng = np.random.default_rng(1)
    y = rng.standard_normal((4, 3))
    z = zscale * rng.standard_normal(3) / np.sqrt(3)
    Yd, Ydd = cp.rotation_jacobians(z, y)
    h = 1e-6
    fd1 = np.zeros_like(Yd)
    fd2 = np.zeros_like(Ydd)
    for p in range(3):
        e = np.zeros(3)
        e[p] = h
        fd1[:, :, p] = (cp.rotate(z + e, y) - cp.rotate(z - e, y)) / (2 * h)
        fd2[:, :, :, p] = (cp.rotation_jacobians(z + e, y)[0] - cp.rotation_jacobians(z - e, y)[0]) / (2 * h)
    assert rel_err(Yd, fd1) < 1e-8
    assert rel_err(Ydd, fd2) < 1e-7
    assert np.abs(Ydd - Ydd.transpose(0, 1, 3, 2)).max() < 1e-12      # a Hessian in (z_p, z_q)


def test_rotation_jacobian_at_zero_is_the_cross_product():
    """d(R(z) y)/dz at z = 0 is -[y]x (R(z) y = y + z x y + O(z^2))."""
    y = np.array([[0.3, -1.1, 0.7]])
    Yd, _ = cp.rotation_jacobians(np.zeros(3), y)
    cross = -np.array([[0.0, -y[0, 2], y[0, 1]], [y[0, 2], 0.0, -y[0, 0]], [-y[0, 1], y[0, 0], 0.0]])
    assert rel_err(Yd[0], cross) < 1e-15


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
def test_reduced_residual_is_the_energy_gradient(etype):
    om, _ = couple_model(etype, pressure=0.05)
    GC = _rand_gc(om, 0.02, 2)
    R = cp.assemble_force(om, GC)
    h = 1e-6
    fd = np.array([(cp.total_energy(om, GC + h * e) - cp.total_energy(om, GC - h * e)) / (2 * h)
                   for e in np.eye(GC.size)])
    assert rel_err(R, fd) < 1e-7
    R2, _ = cp.assemble_reduced(om, GC)
    assert rel_err(R2, R) < 1e-14


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8'])
def test_reduced_tangent_is_the_residual_derivative(etype):
    om, _ = couple_model(etype, pressure=0.05)
    GC = _rand_gc(om, 0.02, 3)
    _, K = cp.assemble_reduced(om, GC)
    h = 1e-6
    fd = np.stack([(cp.assemble_force(om, GC + h * e) - cp.assemble_force(om, GC - h * e)) / (2 * h)
                   for e in np.eye(GC.size)], axis=1)
    Kd = K.toarray()
    assert rel_err(Kd, fd) < 1e-6
    assert np.abs(Kd - Kd.T).max() < 1e-12 * np.abs(Kd).max()


def test_fixed_rp_dofs_leave_the_system():
    om, _ = couple_model(fixed_rp=[False, False, True, True, True, False])
    nf, rf = cp.layout(om)
    assert rf.tolist() == [True, True, False, False, False, True]
    assert cp.n_gc(om) == int(nf.sum()) + 3


def test_solve_moves_the_coupled_face_rigidly():
    om, mesh = couple_model('C3D8', moment=(0.0, 0.0, 0.05), force=(0.0, 0.02, 0.0))
    GC, its = osol.static_solve(om, increments=2, rel_tol=1e-11, linear='direct', max_iter=50,
                                armijo_slack_rel=1e-12)
    assert GC is not False
    R = cp.assemble_force(om, GC)
    assert np.linalg.norm(R) < 1e-10
    U, rp = cp.state(om, GC)
    nodes = np.asarray(mesh['node_sets']['x1'])
    x = mesh['nodes'][nodes] + U[nodes]
    d0 = np.linalg.norm(mesh['nodes'][nodes][:, None] - mesh['nodes'][nodes][None], axis=2)
    d1 = np.linalg.norm(x[:, None] - x[None], axis=2)
    assert np.abs(d1 - d0).max() < 1e-13                    # coupled nodes keep their distances
    assert abs(rp[0, 5]) > 1e-3 and abs(rp[0, 1]) > 1e-3    # the moment turned it, the force moved it
