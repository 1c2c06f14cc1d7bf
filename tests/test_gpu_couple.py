"""Couple constraints and reference points on the GPU against the oracle.

* kinematics: coupled displacements and their z-derivatives (fea_couple_kinematics)
  against the oracle's restatement of couple.py:186-292, and at z = 0 against the exact
  second derivative;
* the reduced residual, tangent and energy (fea_couple_condense through the Assembly
  API) against the oracle's T^T R, T^T K T + R d2u/dz2 on C3D4 / C3D8 / C3D10 blocks
  with a pressure on coupled nodes, point loads and fixed point DOFs;
* static solves against the oracle's Newton (FP64 and mixed precision);
* the reference's own Couple model (tests/pressure_test/C3D4Less.inp, 33 612 C3D4, the
  head coupled to a point at (0, 0, 80), pressure 0.06 on surface_1_All): the GPU
  solution is an equilibrium of the oracle's equations and moves the head rigidly.
"""
import os

import numpy as np
import pytest
import torch

from fea_amd import FEA_INP, ReferencePoint, boundarys, constraints, from_inp, loads, materials, solver, workloads
from oracle import assembly as oasm
from oracle import couple as cp
from oracle import solvers as osol
from tests.helpers import FACE_NODES, nh, rel_err
from tests.test_oracle_couple import _rand_gc, couple_model

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'inp')


def fea_couple(etype='C3D8', moment=(0.0, 0.0, 0.02), force=(0.0, 0.01, 0.0), pressure=None, fixed_rp=None):
    om, mesh = couple_model(etype, moment=moment, force=force, pressure=pressure, fixed_rp=fixed_rp)
    ld = None if pressure is None else (lambda nm: loads.Pressure(nm, 'y1', pressure))
    fe = workloads._controller_from_mesh(mesh, materials.NeoHookean(1.0, 10.0), fixed_set='x0', load=ld)
    a = fe.assembly
    rp = a.add_reference_point(ReferencePoint(list(om['rps'][0]['node'])))
    a.add_constraint(constraints.Couple(instance_name='block', set_nodes_name='x1', rp_name=rp))
    if force is not None:
        a.add_load(loads.Concentrate_Force(rp_name=rp, force=list(force)))
    if moment is not None:
        a.add_load(loads.Moment(rp_name=rp, moment=list(moment)))
    if fixed_rp is not None:
        a.add_boundary(boundarys.Boundary_Condition_RP(rp_name=rp, indexDoF=[i for i, f in enumerate(fixed_rp) if f]))
    return fe, om, mesh


@pytest.mark.parametrize('zmag', [0.0, 1e-7, 1e-3, 0.3, 1.4, 3.0])
def test_kinematics_match_oracle(gpu, zmag):
    fe, om, mesh = fea_couple('C3D8')
    fe.assembly.initialize()
    p = fe.assembly.plan
    rng = np.random.default_rng(7)
    GC = _rand_gc(om, 0.01, 7)
    z = rng.standard_normal(3)
    GC[-3:] = zmag * z / np.linalg.norm(z)
    U = p.gc_to_u(torch.as_tensor(GC, device='cuda'))
    p.kinematics(U, True)
    U_o, rp = cp.state(om, GC)
    assert rel_err(U[:3 * p.n_nodes].cpu().numpy().reshape(-1, 3), U_o) < 1e-14
    y = om['nodes'][p.cnode_np] - np.asarray(om['rps'][0]['node'])
    Yd = p._arr['ydot'].cpu().numpy()[:9 * y.shape[0]].reshape(-1, 3, 3)
    Ydd = p._arr['ydot2'].cpu().numpy()[:27 * y.shape[0]].reshape(-1, 3, 3, 3)
    Yd_o, Ydd_o = cp.rotation_jacobians(rp[0, 3:], y)
    # the reference's closed form carries ~eps / |z| of dY/dz in its (1 - cos) terms
    assert rel_err(Yd, Yd_o) < (max(1e-13, 4e-16 / zmag) if zmag > 0 else 1e-13)
    if zmag >= 1e-3:
        # the reference's closed form loses ~eps / |z|^2 of d2Y/dz2 to cancellation
        assert rel_err(Ydd, Ydd_o) < max(1e-13, 1e-15 / zmag ** 2)
    else:
        # exact value at z -> 0: d2(R(z) y)_i / dz_p dz_q = (d_ip y_q + d_iq y_p) / 2 - d_pq y_i
        I = np.eye(3)
        exact = 0.5 * (np.einsum('ip,nq->nipq', I, y) + np.einsum('iq,np->nipq', I, y)) - \
            np.einsum('pq,ni->nipq', I, y)
        assert rel_err(Ydd, exact) < 1e-6 if zmag > 0 else rel_err(Ydd, exact) < 1e-15


@pytest.mark.parametrize('etype,fixed_rp,prec', [('C3D8', None, 'fp64'), ('C3D4', [False, True, False, True, False, False], 'fp64'),
                                                 ('C3D10', None, 'fp64'), ('C3D8', None, 'mixed')])
def test_reduced_system_matches_oracle(gpu, etype, fixed_rp, prec):
    fe, om, _ = fea_couple(etype, pressure=0.05, fixed_rp=fixed_rp)
    a = fe.assembly
    a.initialize()
    GC = _rand_gc(om, 0.02, 11)
    assert a.GC.numel() == GC.size == cp.n_gc(om)
    R_o, K_o = cp.assemble_reduced(om, GC)
    GCt = torch.as_tensor(GC, device='cuda')
    if prec == 'fp64':
        R, Kidx, Kval = a.assemble_Stiffness_Matrix(GC=GCt)
        Kval = Kval.cpu().numpy()
        tol = 1e-12
    else:
        p = a.plan
        R, K32 = p.tangent(p.gc_to_u(GCt), 1.0, 1, val_dtype=0)
        Kidx, Kval = a.K_indices(), K32.cpu().numpy().astype(np.float64)
        tol = 1e-7
    Kidx = Kidx.cpu().numpy()
    K = np.zeros(K_o.shape)
    K[Kidx[0], Kidx[1]] = Kval
    assert rel_err(R.cpu().numpy(), R_o) < 1e-12
    assert np.abs(K - K_o.toarray()).max() < tol * np.abs(K_o.data).max()
    E = float(a._total_Potential_Energy(GC=GCt))
    assert abs(E - cp.total_energy(om, GC)) < 1e-12 * max(1.0, abs(E))
    assert rel_err(a.assemble_force(GC=GCt).cpu().numpy(), R_o) < 1e-12


def test_rgc_blocks_and_initial_rotation(gpu):
    fe, om, mesh = fea_couple('C3D8')
    a = fe.assembly
    a.initialize()
    rp = a._reference_points['rp-0']
    # the last reference point starts at rz = 1e-5 (assembly.py:113)
    assert float(a.GC[-1]) == 1e-5
    RGC = a._GC2RGC(a.GC)
    assert RGC[rp._RGC_index].shape == (6,) and float(RGC[rp._RGC_index][5]) == 1e-5
    nodes = np.asarray(mesh['node_sets']['x1'])
    U_o, _ = cp.state(om, a.GC.cpu().numpy())
    # |u| ~ 1e-5 |y| here: the oracle's y cos(t) - y carries eps |y| of cancellation
    assert rel_err(RGC[0].cpu().numpy()[nodes], U_o[nodes]) < 1e-10
    assert rel_err(a._RGC2GC(RGC).cpu().numpy(), a.GC.cpu().numpy()) == 0.0


@pytest.mark.parametrize('prec', ['fp64', 'mixed'])
def test_static_solve_matches_oracle(gpu, prec):
    moment, force = (0.0, 0.0, 0.05), (0.0, 0.02, 0.0)
    fe, om, mesh = fea_couple('C3D8', moment=moment, force=force, pressure=0.03)
    fe.solver = solver.StaticImplicitSolver(precision=prec, rel_tol=1e-11, load_steps=2, cg_tol=1e-24,
                                            cg_max_iter=20000, armijo_slack=1e-12)
    fe.initialize()
    GC = fe.solver.solve(GC0=torch.zeros_like(fe.assembly.GC))
    assert GC is not False
    GC_o, _ = osol.static_solve(om, increments=2, rel_tol=1e-11, linear='direct', max_iter=50,
                                armijo_slack_rel=1e-12)
    assert rel_err(GC.cpu().numpy(), GC_o) < 1e-8
    assert np.linalg.norm(cp.assemble_force(om, GC.cpu().numpy())) < 1e-9


def _c3d4less(pressure):
    fe, inp = workloads.inp_couple(os.path.join(GOLDEN, 'C3D4Less.inp.gz'), pressure=pressure)
    part = inp.part['final_model']
    X = part.nodes[:, 1:].numpy().astype(np.float64)
    conn = part.elems['C3D4'][:, 1:].astype(np.int64)
    mu, kappa = inp.material['Rubber'].mat_para
    fixed = np.zeros((X.shape[0], 3), dtype=bool)
    fixed[sorted(part.sets_nodes['surface_0_Bottom'])] = True
    pos = {eid: i for i, eid in enumerate(part.elems['C3D4'][:, 0])}
    facets = np.concatenate([conn[[pos[e] for e in ids]][:, FACE_NODES['C3D4'][f]]
                             for ids, f in part.surfaces['surface_1_All']])
    om = {'nodes': X, 'fixed': fixed, 'groups': [{'etype': 'C3D4', 'conn': conn, 'material': nh(mu, kappa)}],
          'pressures': [{'stype': 'T3', 'conn': facets, 'p': pressure}], 'tractions': [],
          'rps': [{'node': (0.0, 0.0, 80.0)}],
          'couples': [{'nodes': np.array(sorted(part.sets_nodes['surface_0_Head'])), 'rp': 0}]}
    return fe, om, np.array(sorted(part.sets_nodes['surface_0_Head']))


@pytest.mark.parametrize('predictor', [False, 'auto'])
def test_reference_couple_model_c3d4less(gpu, predictor, prec='mixed'):
    """tests/pressure_test/test_benchmark.py's model (C3D4Less.inp) at half its pressure (0.03;
    the full 0.06 curls the actuator through 2.4 rad and is the benchmark's workload); with
    'auto', the second increment starts from an extrapolation or, when its energy is higher,
    from the first increment's equilibrium."""
    fe, om, head = _c3d4less(0.03)
    fe.solver = solver.StaticImplicitSolver(precision=prec, rel_tol=1e-9, load_steps=2, cg_tol=1e-16,
                                            cg_max_iter=20000, armijo_slack=1e-12, forcing='ew', predictor=predictor)
    fe.initialize()
    a = fe.assembly
    assert a.GC.numel() == cp.n_gc(om)
    GC = fe.solver.solve(GC0=torch.zeros_like(a.GC))
    assert GC is not False
    g = GC.cpu().numpy()
    f_ext = np.linalg.norm(cp.assemble_force(om, np.zeros_like(g)))
    assert np.linalg.norm(cp.assemble_force(om, g)) < 1e-7 * f_ext        # oracle equilibrium
    U, rp = cp.state(om, g)
    x = om['nodes'][head] + U[head]
    d0 = np.linalg.norm(om['nodes'][head][:, None] - om['nodes'][head][None], axis=2)
    d1 = np.linalg.norm(x[:, None] - x[None], axis=2)
    assert np.abs(d1 - d0).max() < 1e-10 * d0.max()
    assert np.abs(rp[0]).max() > 1e-2                                          # the head moved
    print('C3D4Less: newton %d, cg %d, orders %s, rp (v, z) %s' % (
        fe.solver.stats['newton_iters'], fe.solver.stats['cg_iters'], fe.solver.stats['predictor_orders'],
        np.array2string(rp[0], precision=4)))
