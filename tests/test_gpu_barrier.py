"""The reference's Gauss-pair contact barrier (ContactSelf) on the GPU against the oracle.

Model: a Kuhn-tet block with a slot (an upper arm cantilevered over a lower one), the
slot's two faces one self-contact surface, the base fixed, and a downward dead traction
on the upper arm's top that closes the slot (at this mesh scale the reference's barrier
constants stop the arm with ~0.4 of overlap).  Checked: the candidate pair set, the barrier energy,
the assembled residual and tangent (through the Assembly API) and full static solves
against the oracle's Newton with direct solves, including a rebuild of the contact
pattern as the pairs change.
"""
import numpy as np
import pytest
import torch

from fea_amd import loads, materials, meshgen, solver, workloads
from oracle import assembly as oasm
from oracle import barrier as ob
from oracle import solvers as osol
from tests.helpers import FACE_NODES, facets, nh, oracle_model, rel_err

pytestmark = pytest.mark.gpu


def slotted_block(nx=6, ny=2, nz=5, size=(3.0, 1.0, 2.5), slot_x=1.0, slot_z=(1.0, 1.5)):
    """Kuhn-tet block minus the cells with x > slot_x and slot_z[0] < z < slot_z[1]; face
    sets 'slot' (both faces of the slot), 'arm_top' (top of the upper arm), node set 'z0'."""
    m = meshgen.kuhn_tet_block(nx, ny, nz, size=size)
    X, conn = m['nodes'], m['conn']
    c = X[conn].mean(axis=1)
    keep = ~((c[:, 0] > slot_x) & (c[:, 2] > slot_z[0]) & (c[:, 2] < slot_z[1]))
    conn = conn[keep]
    used = np.unique(conn)
    remap = np.full(X.shape[0], -1)
    remap[used] = np.arange(used.size)
    X, conn = X[used], remap[conn]
    tol = 1e-9

    def faces(pred):
        out = {}
        for f, loc in FACE_NODES['C3D4'].items():
            fx = X[conn[:, loc]]
            sel = np.nonzero(pred(fx))[0]
            if sel.size:
                out[f] = sel
        return [(ids, f) for f, ids in out.items()]
    on = lambda z: (lambda fx: np.all(np.abs(fx[:, :, 2] - z) < tol, axis=1) & np.all(fx[:, :, 0] > slot_x - tol, axis=1))
    lo, up = faces(on(slot_z[0])), faces(on(slot_z[1]))
    top = faces(on(size[2]))
    mesh = {'nodes': X, 'conn': conn, 'etype': 'C3D4',
            'node_sets': {'z0': np.nonzero(np.abs(X[:, 2]) < tol)[0]},
            'face_sets': {'slot': lo + up, 'slot_lo': lo, 'slot_up': up, 'arm_top': top}}
    return mesh


def model(q=0.05):
    mesh = slotted_block()
    fe = workloads._controller_from_mesh(mesh, materials.NeoHookean(1.0, 10.0), fixed_set='z0',
                                         load=lambda nm: loads.Traction(nm, 'arm_top', (0.0, 0.0, -q)))
    fe.assembly.add_load(loads.ContactSelf(instance_name='block', surface_name='slot'), 'contact')
    om = oracle_model(mesh, nh(1.0, 10.0), 'z0', traction=(0.0, 0.0, -q), load_set='arm_top')
    om['barriers'] = [{'conn1': facets(mesh, 'slot')[0], 'conn2': None}]
    return fe, om, mesh


def closed_state(om, mesh, seed=0):
    """A deformed state with the slot closed: the upper arm bent down by up to 0.9."""
    X = mesh['nodes']
    rng = np.random.default_rng(seed)
    U = np.zeros_like(X)
    arm = X[:, 2] > 1.25
    U[arm, 2] = -0.9 * np.clip((X[arm, 0] - 1.0) / 2.0, 0, 1) ** 1.5
    U += 0.01 * rng.standard_normal(X.shape)
    U[mesh['node_sets']['z0']] = 0.0
    return U


def test_pairs_and_energy_match_oracle(gpu):
    fe, om, mesh = model()
    fe.assembly.initialize()
    plan = fe.assembly.plan
    b = plan.barriers[0]
    conn = facets(mesh, 'slot')[0]
    assert b.F1 == conn.shape[0]
    U = closed_state(om, mesh)
    Ut = torch.as_tensor(U.reshape(-1), device='cuda')
    n = b.search(Ut, b.radius)
    got = np.sort(np.stack([b.pi[:n].cpu().numpy(), b.pj[:n].cpu().numpy()], 1), axis=1)
    got = {tuple(p) for p in got}
    # the product's facet order is its own; map pairs through the facets' node sets
    key = lambda f: tuple(sorted(f))
    prod = [key(f) for f in b.c1_np]
    orac = {key(f): k for k, f in enumerate(conn)}
    mapped = {tuple(sorted((orac[prod[i]], orac[prod[j]]))) for i, j in got}
    x = mesh['nodes'] + U
    want = {tuple(p) for p in ob.candidate_pairs(ob.centroids(x, conn), b.radius)}
    assert mapped == want and len(want) > 0
    E = float(b.energy(Ut).item())
    E_o = oasm.barrier_terms(om, om['barriers'][0], U, 0)[0].sum()
    assert E_o > 1e-8
    assert abs(E - E_o) < 1e-12 * E_o


def test_assembled_system_matches_oracle(gpu):
    fe, om, mesh = model()
    a = fe.assembly
    a.initialize()
    U = closed_state(om, mesh, seed=1)
    free = oasm.free_mask(om)
    GC = U.reshape(-1)[free]
    R, Kidx, Kval = a.assemble_Stiffness_Matrix(GC=torch.as_tensor(GC, device='cuda'))
    R_o, K_o = oasm.assemble_reduced(om, U)
    assert rel_err(R.cpu().numpy(), R_o) < 1e-11
    K = np.zeros(K_o.shape)
    Ki = Kidx.cpu().numpy()
    K[Ki[0], Ki[1]] = Kval.cpu().numpy()
    assert np.abs(K - K_o.toarray()).max() < 1e-11 * np.abs(K_o.data).max()
    E = float(a._total_Potential_Energy(GC=torch.as_tensor(GC, device='cuda')))
    assert abs(E - oasm.total_energy(om, U)) < 1e-12 * abs(E)


@pytest.mark.parametrize('prec', ['fp64', 'mixed'])
def test_static_solve_closes_the_slot_like_the_oracle(gpu, prec):
    fe, om, mesh = model(q=0.05)
    fe.solver = solver.StaticImplicitSolver(precision=prec, rel_tol=1e-10, load_steps=4, cg_tol=1e-24,
                                            cg_max_iter=20000, armijo_slack=1e-12, predictor=False)
    fe.initialize()
    GC = fe.solver.solve(GC0=torch.zeros_like(fe.assembly.GC))
    assert GC is not False
    GC_o, _ = osol.static_solve(om, increments=4, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12,
                                predictor=False, max_iter=100)
    assert GC_o is not False
    assert rel_err(GC.cpu().numpy(), GC_o) < 1e-7
    U = oasm.gc_to_u(om, GC_o)
    Eb = oasm.barrier_terms(om, om['barriers'][0], U, 0)[0].sum()
    assert Eb > 1e-6                                  # the slot closed onto the barrier
    print('barrier: newton %d cg %d rebuilds %d' % (fe.solver.stats['newton_iters'], fe.solver.stats['cg_iters'],
                                                    fe.assembly.plan.rebuilds))


def test_two_surface_contact_matches_oracle(gpu):
    """Contact (contact.py:714-1207) between the slot's two faces as two surfaces: ratio 1
    for every pair, pairs (upper facet, lower facet) within the candidate radius."""
    mesh = slotted_block()
    q = 0.05
    fe = workloads._controller_from_mesh(mesh, materials.NeoHookean(1.0, 10.0), fixed_set='z0',
                                         load=lambda nm: loads.Traction(nm, 'arm_top', (0.0, 0.0, -q)))
    fe.assembly.add_load(loads.Contact(instance_name1='block', instance_name2='block', surface_name1='slot_up',
                                       surface_name2='slot_lo'), 'contact')
    om = oracle_model(mesh, nh(1.0, 10.0), 'z0', traction=(0.0, 0.0, -q), load_set='arm_top')
    om['barriers'] = [{'conn1': facets(mesh, 'slot_up')[0], 'conn2': facets(mesh, 'slot_lo')[0]}]
    a = fe.assembly
    a.initialize()
    U = closed_state(om, mesh, seed=2)
    free = oasm.free_mask(om)
    GC = U.reshape(-1)[free]
    R, Kidx, Kval = a.assemble_Stiffness_Matrix(GC=torch.as_tensor(GC, device='cuda'))
    R_o, K_o = oasm.assemble_reduced(om, U)
    assert rel_err(R.cpu().numpy(), R_o) < 1e-11
    K = np.zeros(K_o.shape)
    Ki = Kidx.cpu().numpy()
    K[Ki[0], Ki[1]] = Kval.cpu().numpy()
    assert np.abs(K - K_o.toarray()).max() < 1e-11 * np.abs(K_o.data).max()
    assert abs(float(a._total_Potential_Energy(GC=torch.as_tensor(GC, device='cuda'))) -
               oasm.total_energy(om, U)) < 1e-12 * abs(oasm.total_energy(om, U))
    fe.solver = solver.StaticImplicitSolver(precision='fp64', rel_tol=1e-10, load_steps=4, cg_tol=1e-24,
                                            cg_max_iter=20000, armijo_slack=1e-12, predictor=False)
    fe.solver.initialize(a)
    GC = fe.solver.solve(GC0=torch.zeros_like(a.GC))
    assert GC is not False
    GC_o, _ = osol.static_solve(om, increments=4, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12,
                                predictor=False, max_iter=100)
    assert rel_err(GC.cpu().numpy(), GC_o) < 1e-7


@pytest.mark.parametrize('n', [1, 63, 1024, 1025, 70001])
def test_scan_i32_matches_cumsum(gpu, n):
    """fea_scan_i32 (the one-workgroup exclusive scan behind the pair search and the contact
    buckets): out[k] = sum in[:k], out[n] = the total."""
    from fea_amd import _native as nat
    from fea_amd._native import call
    a = np.random.default_rng(n).integers(0, 9, n).astype(np.int32)
    out = torch.full((n + 1,), -1, dtype=torch.int32, device='cuda')
    call('fea_scan_i32', n, torch.as_tensor(a, device='cuda'), out, nat.stream_ptr())
    assert np.array_equal(out.cpu().numpy(), np.concatenate([[0], np.cumsum(a)]).astype(np.int32))
