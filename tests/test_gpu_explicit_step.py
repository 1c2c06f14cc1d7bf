"""GPU: contact forces from the reused candidate lists (fea_contact_rebuild +
fea_contact_eval) equal those of the exact multi-launch search bit for bit, whether the
lists are fresh, older but within skin / 2, or stale (full scans), and the explicit step
with the per-tile force
reduction (fea_elem_residual_tiles + fea_explicit_update_tiles) stays within round-off
of the node-major staged step and repeats bit for bit; the oracle checks of the same
loop are in test_gpu_contact.py."""
import numpy as np
import pytest
import torch

from fea_amd import solver, workloads
from tests.test_gpu_contact import pressed

pytestmark = pytest.mark.gpu


def _contact_nodes_shift(st, U, shift):
    """U with every contact node moved by ``shift`` (a 3-vector) from U."""
    V = U.clone().view(-1, 3)
    V[st.cnode.long()] += torch.as_tensor(shift, dtype=torch.float64, device=U.device)
    return V.view(-1)


@pytest.mark.parametrize('n,seed,move', [(4, 0, 0.0), (7, 1, 0.0), (7, 1, 0.3), (7, 1, 0.7), (9, 2, 0.45)])
def test_reused_lists_equal_exact_search(gpu, n, seed, move):
    """fea_contact_rebuild at a state whose contact nodes sit move * skin away from U, then
    fea_contact_eval at U: the stored lists (move <= 1/2) or the full scans (move > 1/2)
    give the exact search's choice, forces, energy and contact-node forces bit for bit."""
    fe, info, U = pressed(n, 0.3, seed)
    print('reused contact lists n=%d seed=%d move=%.2f' % (n, seed, move), flush=True)
    load = fe.assembly.get_load('contact')
    st = load.state(fe.assembly.plan)
    fc_ref = st.forces(U, search=True).clone()
    e_ref = st.energy()
    chosen_ref, f_ref, w_ref = st.chosen.clone(), st.force.clone(), st.weight.clone()
    assert int((chosen_ref >= 0).sum()) > 0.3 * st.S
    for t in (st.fc, st.force, st.weight, st.chosen):
        t.fill_(-7)
    st.fc.zero_()
    d = move * load.skin / np.sqrt(3.0)
    st.rebuild(_contact_nodes_shift(st, U, (d, -d, d)))
    assert int(st.sup_max.item()) <= load.kmax
    fc = st.step_forces(U)
    torch.cuda.synchronize()
    from fea_amd.assemble.loads.contact import slots_max
    dmax = slots_max(st.dmax.cpu().numpy())
    assert abs(np.sqrt(dmax) - move * load.skin) < 1e-9
    assert torch.equal(st.chosen, chosen_ref)
    assert torch.equal(st.force, f_ref) and torch.equal(st.weight, w_ref)
    assert torch.equal(fc, fc_ref)                      # every entry: contact nodes and zeros
    assert st.energy() == e_ref
    # the stored facet -> slave lists are the inverse of the stored candidate lists
    S, kmax = st.S, load.kmax
    cnt = np.minimum(st.vcount.cpu().numpy(), kmax)
    cand = st.vcand.view(S, kmax).cpu().numpy()
    pairs = sorted((int(f), s) for s in range(S) for f in cand[s, :cnt[s]])
    istart, items = st.inv_start.cpu().numpy(), st.inv_items.cpu().numpy()
    inv = sorted((f, int(x)) for f in range(st.F) for x in items[istart[f]:istart[f + 1]])
    assert inv == pairs
    for f in range(st.F):
        assert np.all(np.diff(items[istart[f]:istart[f + 1]]) > 0)


def test_reused_lists_rebuild_on_motion(gpu):
    """The host rebuilds once the polled displacement passes skin / 4, and the forces
    stay those of the exact search at every step."""
    fe, info, U = pressed(5, 0.3, 3)
    load = fe.assembly.get_load('contact')
    st = load.state(fe.assembly.plan)
    step = 0.1 * load.skin
    for k in range(12):
        Uk = _contact_nodes_shift(st, U, (0.0, 0.0, k * step))
        fc = st.step_forces(Uk, poll_every=1).clone()
        torch.cuda.synchronize()
        fc_ref = st.forces(Uk, search=True).clone()
        assert torch.equal(fc, fc_ref), k
        st.fc.zero_()
    assert st.rebuilds >= 2


def _run_steps(precision, tiles, nsteps=60, n=4):
    fe, info = workloads.two_blocks_contact(n, velocity=1.0, gap=0.01, jitter=0.1, seed=5, upper_width=0.9,
                                            offset=(0.037, 0.021))
    fe.solver = solver.DynamicExplicitSolver(time_end=1.0, time_per_storage=1.0, dt=3e-3, store=False,
                                             precision=precision)
    fe.initialize()
    sol = fe.solver
    assert sol._tiles is not None
    p = fe.assembly.plan
    st = fe.assembly.get_load('contact').state(p)
    if not tiles:
        sol._tiles = None
        sol._contact_force = lambda u, *_: st.forces(u, search=True, check=False)     # the multi-launch contact
    U = p.gc_to_u(torch.zeros_like(fe.assembly.GC))
    vh = workloads.initial_velocity(fe, info).clone()
    acc = torch.zeros_like(vh)
    active = 0
    for k in range(nsteps):
        sol.step(U, vh, acc, sol._deltaT, first=(k == 0))
        if k % 10 == 9:
            active = max(active, int((st.chosen >= 0).sum()))
    torch.cuda.synchronize()
    return U.clone(), vh.clone(), acc.clone(), active


@pytest.mark.parametrize('n,precision,tol', [(4, 'fp32', 1e-9), (4, 'fp64', 1e-12), (9, 'fp32', 1e-9),
                                             (9, 'fp64', 1e-12)])
def test_tile_step_matches_node_major_and_repeats(gpu, n, precision, tol):
    """Per-tile partial sums change only the association of each node's float64 force sum:
    60 contact steps stay within tol of the node-major staged step, and two runs agree bit
    for bit (n = 4: one partial tile per block; n = 9: 4374 tets per block, 8 full tiles and
    a partial one)."""
    print('explicit tiles vs node-major, n=%d %s' % (n, precision), flush=True)
    U1, v1, a1, n1 = _run_steps(precision, True, n=n)
    U0, v0, a0, n0 = _run_steps(precision, False, n=n)
    U2, v2, a2, n2 = _run_steps(precision, True, n=n)
    assert n1 > 0 and n1 == n0                          # contact was active
    assert torch.equal(U1, U2) and torch.equal(v1, v2) and torch.equal(a1, a2)
    rel = lambda x, y: float((x - y).abs().max() / y.abs().max())
    assert rel(U1, U0) < tol and rel(v1, v0) < tol and rel(a1, a0) < 100 * tol


@pytest.mark.parametrize('overlap', [True, False])
def test_graph_chunks_equal_eager_steps(gpu, overlap):
    """run_steps replays captured 8-step chunks (contact on a side stream with overlap), with
    the host's polls and list rebuilds between replays: the same trajectory as stepping one
    call at a time, bit for bit, over 150 steps with contact active and lists rebuilt."""
    print('explicit graph chunks vs eager, overlap=%s' % overlap, flush=True)
    out = []
    for graph in (True, False):
        fe, info = workloads.two_blocks_contact(4, velocity=1.0, gap=0.01, jitter=0.1, seed=5)
        fe.solver = solver.DynamicExplicitSolver(time_end=1.0, time_per_storage=1.0, dt=1.5e-3, store=False,
                                                 precision='fp32', overlap=overlap, use_graph=graph)
        fe.initialize()
        sol = fe.solver
        p = fe.assembly.plan
        st = fe.assembly.get_load('contact').state(p)
        U = p.gc_to_u(torch.zeros_like(fe.assembly.GC))
        vh = workloads.initial_velocity(fe, info).clone()
        acc = torch.zeros_like(vh)
        active = 0
        sol.run_steps(U, vh, acc, sol._deltaT, 30, first=True)
        for _ in range(6):
            sol.run_steps(U, vh, acc, sol._deltaT, 20)
            torch.cuda.synchronize()
            active = max(active, int((st.chosen >= 0).sum()))
        sol._check_contacts()
        out.append((U.clone(), vh.clone(), acc.clone(), active, st.rebuilds))
    (U1, v1, a1, n1, r1), (U0, v0, a0, n0, r0) = out
    assert n1 > 0 and r1 >= 2 and r0 >= 2
    assert torch.equal(U1, U0) and torch.equal(v1, v0) and torch.equal(a1, a0)


@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_staged_tile_nodes_bit_identical(gpu, precision):
    """Element tiles reading their nodes from LDS (tile node lists, local connectivity) give
    the gathering kernel's partials, hence the same 40 steps, bit for bit."""
    print('explicit staged tile nodes, %s' % precision, flush=True)
    out = []
    for stage in (True, False):
        fe, info = workloads.two_blocks_contact(4, velocity=1.0, gap=0.01, jitter=0.1, seed=5)
        # the separate launches: the fused step's element launch has no node staging
        fe.solver = solver.DynamicExplicitSolver(time_end=1.0, time_per_storage=1.0, dt=3e-3, store=False,
                                                 precision=precision, fuse_contact=False)
        fe.solver.stage_nodes = stage
        fe.initialize()
        assert ('stage' in fe.solver._tiles) == stage
        p = fe.assembly.plan
        U = p.gc_to_u(torch.zeros_like(fe.assembly.GC))
        vh = workloads.initial_velocity(fe, info).clone()
        acc = torch.zeros_like(vh)
        fe.solver.run_steps(U, vh, acc, fe.solver._deltaT, 40, first=True)
        torch.cuda.synchronize()
        out.append((U.clone(), vh.clone(), acc.clone()))
    for a, b in zip(out[0], out[1]):
        assert torch.equal(a, b)


@pytest.mark.parametrize('precision,graph', [('fp32', False), ('fp32', True), ('fp64', False)])
def test_fused_contact_step_bit_identical(gpu, precision, graph):
    """The two-launch step (element pass + contact penalty; node update + contact-node forces
    + displacement tracking) against the separate launches (element pass, fea_contact_eval,
    update): the same trajectory and contact forces, bit for bit, over 150 steps with contact
    active and the lists rebuilt."""
    print('explicit fused contact step, %s graph=%s' % (precision, graph), flush=True)
    out = []
    for fuse in (True, False):
        fe, info = workloads.two_blocks_contact(4, velocity=1.0, gap=0.01, jitter=0.1, seed=5)
        fe.solver = solver.DynamicExplicitSolver(time_end=1.0, time_per_storage=1.0, dt=1.5e-3, store=False,
                                                 precision=precision, use_graph=graph, fuse_contact=fuse)
        fe.initialize()
        sol = fe.solver
        p = fe.assembly.plan
        st = fe.assembly.get_load('contact').state(p)
        assert (sol._fused_contact() is not None) == fuse
        U = p.gc_to_u(torch.zeros_like(fe.assembly.GC))
        vh = workloads.initial_velocity(fe, info).clone()
        acc = torch.zeros_like(vh)
        active = 0
        sol.run_steps(U, vh, acc, sol._deltaT, 30, first=True)
        for _ in range(6):
            sol.run_steps(U, vh, acc, sol._deltaT, 20)
            torch.cuda.synchronize()
            active = max(active, int((st.chosen >= 0).sum()))
        sol._check_contacts()
        out.append((U.clone(), vh.clone(), acc.clone(), st.fc.clone(), active, st.rebuilds))
    (U1, v1, a1, f1, n1, r1), (U0, v0, a0, f0, n0, r0) = out
    assert n1 > 0 and r1 >= 2 and r0 >= 2
    assert torch.equal(U1, U0) and torch.equal(v1, v0) and torch.equal(a1, a0)
    assert torch.equal(f1, f0)
