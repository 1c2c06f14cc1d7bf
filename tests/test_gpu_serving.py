"""GPU tests of the batched PCG and the continuous-batching solve server:
batched systems against direct solves and against themselves alone (bit for
bit), HIP-graph replays against eager launches, server results against the
one-at-a-time solver and the oracle's direct-solve Newton (1e-8 relative),
alone vs in the mix (bit for bit), admission under an HBM budget."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from fea_amd import serving, workloads
from fea_amd._native import F32, F64
from fea_amd.solver._linear_solver import BatchedPCG, PcgVectors
from oracle import solvers as osol
from tests.helpers import mr, nh, oracle_model, rel_err

pytestmark = pytest.mark.gpu


def _systems(n_sys, dtype):
    """Tangents of small jittered cubes / beams at a random displacement."""
    out = []
    for k in range(n_sys):
        if k % 2 == 0:
            fe, _ = workloads.cube_pressure(2 + k % 3, pressure=0.2, jitter=0.05, seed=k)
        else:
            fe, _ = workloads.beam_traction(6 + k, 2, 2)
        fe.assembly.initialize()
        p = fe.assembly.plan
        rng = np.random.default_rng(10 + k)
        U = p.gc_to_u(torch.as_tensor(0.01 * rng.standard_normal(p.n_free), device='cuda'))
        R, K = p.tangent(U, 1.0, F64, val_dtype=dtype)
        out.append((fe, p, R.clone(), K.clone()))
    return out


def _solve(bpcg, systems, tol=1e-26, max_iter=6000):
    entries, vecs = [], []
    for fe, p, R, K in systems:
        v = PcgVectors(p.n_free, F64, 'cuda', n_groups=p.n_groups)
        v.b.copy_(R).neg_()
        v.x.zero_()
        entries.append((p, K, v, tol, max_iter))
        vecs.append(v)
    iters = bpcg.solve(entries)
    return [v.x.clone() for v in vecs], iters


@pytest.mark.parametrize('pc', ['jacobi', 'block'])
@pytest.mark.parametrize('dtype', [F64, F32])
def test_batched_pcg_matches_direct_and_alone(gpu, dtype, pc):
    systems = _systems(5, dtype)
    bp = BatchedPCG(dtype, F64, 'cuda', chunk=16, precond=pc)
    xs, iters = _solve(bp, systems)
    for (fe, p, R, K), x, it in zip(systems, xs, iters):
        A = sp.csr_matrix((K.double().cpu().numpy(), p.col_np, p.rowptr_np), shape=(p.n_free, p.n_free))
        b = -R.cpu().numpy()
        ref = spla.spsolve(A.tocsc(), b)
        xh = x.cpu().numpy()
        assert np.linalg.norm(A @ xh - b) <= 1e-9 * np.linalg.norm(b), it
        assert rel_err(xh, ref) < 1e-7
        assert 0 < it < 6000
    # every system alone: identical iterates, so identical bits and counts
    for k, s in enumerate(systems):
        xa, ia = _solve(BatchedPCG(dtype, F64, 'cuda', chunk=16, precond=pc), [s])
        assert torch.equal(xa[0], xs[k]) and ia[0] == iters[k]
    # a different mix and order
    sub = [systems[3], systems[0], systems[4]]
    xm, im = _solve(BatchedPCG(dtype, F64, 'cuda', chunk=7, precond=pc), sub)
    for x, k in zip(xm, (3, 0, 4)):
        assert torch.equal(x, xs[k])


@pytest.mark.parametrize('pc', ['jacobi', 'block'])
def test_batched_pcg_node_groups_of_one_and_two_rows(gpu, pc):
    """Systems whose row groups are not all 3 rows (a face held in x only: its nodes have 2
    free rows; a face held in x and y: 1 free row; the generic node-grouped SpMV, and for node-block Jacobi the blocks confirmed
    against the CSR columns) beside a triple-structured one: the batched solve matches a
    direct solve and each system's iterates do not depend on the batch."""
    from fea_amd import boundarys
    systems = []
    sizes = set()
    for k, (face, dofs) in enumerate((('x0', [0]), ('y1', [0]), ('y1', [0, 1]))):
        fe, _ = workloads.cube_pressure(3, pressure=0.2, jitter=0.05, seed=20 + k)
        fe.assembly.add_boundary(boundarys.Boundary_Condition('block', face, indexDoF=dofs))
        fe.assembly.initialize()
        p = fe.assembly.plan
        g = np.diff(p.grp.cpu().numpy())
        sizes |= set(g.tolist())
        assert p.spmv_groups > 0 and g.min() < 3                 # not triple-structured
        U = p.gc_to_u(torch.as_tensor(0.01 * np.random.default_rng(30 + k).standard_normal(p.n_free),
                                      device='cuda'))
        R, K = p.tangent(U, 1.0, F64, val_dtype=F64)
        systems.append((fe, p, R.clone(), K.clone()))
    assert sizes == {1, 2, 3}                                    # 1-, 2- and 3-row node groups all occur
    systems += _systems(1, F64)
    xs, iters = _solve(BatchedPCG(F64, F64, 'cuda', chunk=16, precond=pc), systems)
    for (fe, p, R, K), x, it in zip(systems, xs, iters):
        A = sp.csr_matrix((K.double().cpu().numpy(), p.col_np, p.rowptr_np), shape=(p.n_free, p.n_free))
        b = -R.cpu().numpy()
        assert rel_err(x.cpu().numpy(), spla.spsolve(A.tocsc(), b)) < 1e-7, it
    for k, s in enumerate(systems):
        xa, ia = _solve(BatchedPCG(F64, F64, 'cuda', chunk=16, precond=pc), [s])
        assert torch.equal(xa[0], xs[k]) and ia[0] == iters[k]


def test_batched_pcg_graph_replays_match_eager(gpu):
    systems = _systems(4, F32)
    eager, it_e = _solve(BatchedPCG(F32, F64, 'cuda', chunk=8), systems, tol=1e-18)
    bg = BatchedPCG(F32, F64, 'cuda', chunk=8, use_graph=True)
    for rep in range(3):
        # eager work and synchronisation between graph-driven solves
        torch.randn(1000, device='cuda').sum().item()
        xg, it_g = _solve(bg, systems, tol=1e-18)
        torch.cuda.synchronize()
        assert it_g == it_e
        for a, b in zip(xg, eager):
            assert torch.equal(a, b)
    assert bg._graph is not None


def test_batched_pcg_one_graph_serves_changing_batches(gpu):
    """The captured chunk reads the batch's counts on the device: one graph, replayed for
    batches of different members and sizes (with eager work between replays), gives each
    system's eager iterates bit for bit."""
    systems = _systems(5, F32)
    bg = BatchedPCG(F32, F64, 'cuda', chunk=8, use_graph=True)
    for subset in ([0, 1, 2, 3, 4], [3], [4, 1], [0, 2, 3], [1, 2, 3, 4], [2]):
        sel = [systems[i] for i in subset]
        eager, it_e = _solve(BatchedPCG(F32, F64, 'cuda', chunk=8), sel, tol=1e-18)
        torch.randn(1000, device='cuda').sum().item()
        xg, it_g = _solve(bg, sel, tol=1e-18)
        torch.cuda.synchronize()
        assert it_g == it_e, subset
        for a, b in zip(xg, eager):
            assert torch.equal(a, b), subset
    assert bg._graph is not None and bg._graph_key[0] == 'dev'


def _small_requests(n=6, seed=3, e_range=(300, 2500), qps=0.0):
    return serving.random_requests(n, qps, seed=seed, e_range=e_range, aspect_range=(1.0, 2.5))


def test_server_matches_one_at_a_time_solver(gpu):
    """Batched mixed-precision server against the one-at-a-time FP64 solver, both converged
    to a relative residual of 1e-12 (at the default 1e-10 two converged solves can differ by
    ~1e-8: the stopping points, not the solvers)."""
    reqs = _small_requests()
    srv = serving.SolveServer(precision='mixed', max_batch=16, rel_tol=1e-12)
    srv.run(reqs, realtime=False)
    assert all(r.status == 'done' for r in reqs), [(r.rid, r.status, r.error) for r in reqs]
    ref, _, _ = serving.solve_one_at_a_time(reqs, precision='fp64', rel_tol=1e-12)
    for r in reqs:
        assert ref[r.rid] is not None
        assert rel_err(r.GC.numpy(), ref[r.rid].numpy()) < 1e-8, r.rid
    assert max(srv.batch_sizes) == len(reqs)


def test_server_matches_oracle_newton(gpu):
    """Two requests against the oracle restatement (direct linear solves), both
    converged to a relative residual of 1e-12."""
    reqs = [r for r in _small_requests(n=8, seed=5, e_range=(200, 800))][:2]
    srv = serving.SolveServer(precision='fp64', rel_tol=1e-12)
    srv.run(reqs, realtime=False)
    from tests.helpers import facets
    for r in reqs:
        assert r.status == 'done'
        mesh = serving.request_mesh(r.spec)
        m = r.spec['material']
        mat = nh(m[1], m[2]) if m[0] == 'NH' else mr(m[1], m[2], m[3])
        kind, val = r.spec['load']
        om = oracle_model(mesh, mat, r.spec['fixed'], pressure=val if kind == 'pressure' else None,
                          traction=val if kind == 'traction' else None, load_set=serving.LOADED_SET)
        GC_o, _ = osol.static_solve(om, increments=r.spec['load_steps'], rel_tol=1e-12, linear='direct',
                                    armijo_slack_rel=1e-12)
        assert rel_err(r.GC.numpy(), GC_o) < 1e-8


def test_server_alone_and_in_mix_identical(gpu):
    reqs = _small_requests(n=5, seed=7)
    serving.SolveServer(precision='mixed').run(reqs, realtime=False)
    for r in reqs:
        alone = serving.SolveRequest(rid=r.rid, spec=r.spec)
        serving.SolveServer(precision='mixed').run([alone], realtime=False)
        assert alone.status == 'done'
        assert torch.equal(alone.GC, r.GC), r.rid
        assert alone.cg_iters == r.cg_iters and alone.newton_iters == r.newton_iters


def test_admission_respects_hbm_budget(gpu):
    reqs = _small_requests(n=6, seed=11)
    ests = [serving.estimate_footprint(r.spec, 'mixed') for r in reqs]
    budget = int(1.05 * max(ests) + min(ests))
    srv = serving.SolveServer(precision='mixed', hbm_budget=budget)
    srv.run(reqs, realtime=False)
    assert all(r.status == 'done' for r in reqs)
    assert srv.peak_used <= budget
    assert max(srv.batch_sizes) < len(reqs)
    for r in reqs:
        assert r.footprint <= r.footprint_est, (r.footprint, r.footprint_est)
    # a request larger than the whole budget is rejected, the others still run
    big = serving.SolveRequest(rid=99, spec=dict(reqs[0].spec, cells=(40, 40, 40)))
    srv2 = serving.SolveServer(precision='mixed', hbm_budget=budget)
    srv2.run([big] + _small_requests(n=2, seed=12), realtime=False)
    assert big.status == 'rejected'


def test_requests_arriving_during_work_join_the_batch(gpu):
    reqs = _small_requests(n=6, seed=13, e_range=(1500, 4000), qps=40.0)
    srv = serving.SolveServer(precision='mixed')
    srv.run(reqs, realtime=True)
    assert all(r.status == 'done' for r in reqs)
    # some request was admitted while an earlier one was still running
    overlap = any(b.t_admit > a.t_admit and b.t_admit < a.t_done for a in reqs for b in reqs)
    assert overlap
    assert len(set(srv.batch_sizes)) > 1
    for r in reqs:
        assert r.t_admit >= r.arrival - 1e-9 and r.latency >= 0


def test_server_graph_chunks_identical_to_eager(gpu):
    """The tick path replaying the captured PCG chunk (use_graph) gives every request's
    solution bit for bit as the eager launches, with requests arriving while others run
    (the batch layout changes between replays)."""
    out = []
    for graph in (True, False):
        reqs = _small_requests(n=8, seed=11, qps=40.0)
        srv = serving.SolveServer(precision='mixed', max_batch=16, use_graph=graph)
        srv.run(reqs, realtime=True)
        assert all(r.status == 'done' for r in reqs), [(r.rid, r.status, r.error) for r in reqs]
        out.append({r.rid: r.GC.clone() for r in reqs})
        if graph:
            assert srv.pcg._graph is not None and srv.pcg.chunks > 3
    for rid in out[0]:
        assert torch.equal(out[0][rid], out[1][rid]), rid


def test_server_multigrid_requests_match_one_at_a_time(gpu):
    """Requests above amg_min_dof free DOF solve their linear systems with the
    multigrid-preconditioned PCG, chunk by chunk, while the others share the batched PCG;
    every result matches the one-at-a-time FP64 solve (1e-8, both converged to a relative
    residual of 1e-12), and the multigrid path ran."""
    reqs = _small_requests(n=6, seed=21, e_range=(800, 4000), qps=30.0)
    srv = serving.SolveServer(precision='mixed', max_batch=16, amg_min_dof=2500, amg_chunk=4, rel_tol=1e-12)
    srv.run(reqs, realtime=True)
    assert all(r.status == 'done' for r in reqs), [(r.rid, r.status, r.error) for r in reqs]
    assert srv.amg_setup_s > 0.0
    ref, _, _ = serving.solve_one_at_a_time(reqs, precision='fp64', rel_tol=1e-12)
    n_amg = 0
    for r in reqs:
        fe, _ = serving.request_model(r.spec)
        fe.assembly.initialize()
        n_amg += fe.assembly.plan.n_free >= 2500
        assert rel_err(r.GC.numpy(), ref[r.rid].numpy()) < 1e-8, r.rid
    assert 0 < n_amg < len(reqs)
