"""Parity at the workloads' full sizes (not toy meshes).

* The 10x10x10 Kuhn-tet cube (6 000 C3D4, 3 630 free DOF, NeoHookean mu=1 kappa=100,
  follower pressure 0.2, 5 increments): the GPU Newton solve in FP64 and in mixed
  precision against the oracle's Newton with direct linear solves, <= 1e-8 relative.
* The 128x32x32 C3D8 Mooney-Rivlin beam (418 176 free DOF) solved as the benchmark solves
  it (mixed precision, node-block Jacobi, Eisenstat-Walker forcing, 'auto' predictor):
  the oracle's FP64 residual at the GPU's converged displacement is <= 1e-9 of the
  external load (2-norms).  The displacement is then within kappa(K) times that of the
  exact equilibrium; the solution's agreement with an FP64 GPU solve is the benchmark's
  rel_err_vs_fp64 field.
* A 14 112-element C3D10 void block: the compliance shape gradient from the adjoint
  against central differences of full Newton solves on 8 seeded directions, and the
  element VJP at the converged state against the torch-autograd oracle.
* Two 10^3-cell Kuhn-tet blocks (6 000 C3D4 each) in penalty contact: 60 explicit
  central-difference steps at 0.9 of the stability limit against the oracle's loop.
* The largest pressure request of the 16-request serving stream (89 488 C3D8,
  NeoHookean, clamped on y0, follower pressure on the five other faces): the solve server
  (mixed) and the one-at-a-time multigrid FP64 solve converge to within 1e-7 of each
  other, and the oracle's FP64 residual at the server's displacement is <= 1e-9 of the
  pressure load (2-norms).
"""
import numpy as np
import pytest
import torch

from fea_amd import adjoint, serving, solver, workloads
from oracle import assembly as oasm
from oracle import autodiff as ad
from oracle import contact as oc
from oracle import solvers as osol
from oracle import surfaces as osurf
from tests.helpers import facets, mr, nh, oracle_model, rel_err
from tests.test_gpu_contact import oracle_two_blocks

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_cube_full_size_matches_oracle_newton(gpu, precision):
    fe, mesh = workloads.cube_pressure(10, pressure=0.2, jitter=0.05, seed=0)
    fe.solver = solver.StaticImplicitSolver(precision=precision, rel_tol=1e-10, load_steps=5, cg_tol=1e-12,
                                            cg_max_iter=20000, armijo_slack=1e-12, predictor='auto',
                                            preconditioner='block')
    GC = fe.solve()
    assert GC is not False
    assert fe.assembly.plan.n_free == 3630
    om = oracle_model(mesh, nh(1.0, 100.0), 'z0', pressure=0.2, load_set='z1')
    GC_o, _ = osol.static_solve(om, increments=5, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12)
    assert GC_o is not False
    err = rel_err(GC.cpu().numpy(), GC_o)
    print('cube 10x10x10 %s: rel err %.2e vs oracle, newton %d cg %d' % (
        precision, err, fe.solver.stats['newton_iters'], fe.solver.stats['cg_iters']))
    assert err < 1e-8


def test_beam_full_size_oracle_residual(gpu):
    t = (0.0, 0.0, -0.01)
    fe, mesh = workloads.beam_traction(128, 32, 32, traction=t)
    fe.solver = solver.StaticImplicitSolver(maximum_iteration=400, precision='mixed', rel_tol=1e-10, load_steps=10,
                                            cg_tol=1e-8, cg_max_iter=20000, armijo_slack=1e-12, forcing='ew',
                                            predictor='auto', preconditioner='block')
    GC = fe.solve()
    assert GC is not False
    assert fe.assembly.plan.n_free == 418176
    GCh = GC.cpu().numpy()
    del fe
    torch.cuda.empty_cache()
    om = oracle_model(mesh, mr(0.5, 0.1, 0.01), 'x0', traction=t, load_set='x1')
    free = oasm.free_mask(om)
    conn, stype = facets(mesh, 'x1')
    f_ext = osurf.traction_force(stype, mesh['nodes'], conn, t).reshape(-1)[free]
    R = oasm.assemble_force(om, oasm.gc_to_u(om, GCh))
    ratio = np.linalg.norm(R) / np.linalg.norm(f_ext)
    print('beam 128x32x32 mixed: |R_oracle(u_gpu)| / |f_ext| = %.2e, tip deflection %.4f' % (
        ratio, -GCh.reshape(-1, 3)[:, 2].min()))
    assert ratio <= 1e-9
    assert -GCh.reshape(-1, 3)[:, 2].min() > 0.1                      # the beam bent


def test_void_c3d10_gradient_at_size(gpu):
    fe, mesh = workloads.void_block_pressure(7, 3, 3)
    assert mesh['conn'].shape[0] >= 12000 and mesh['etype'] == 'C3D10'
    fe.solver = solver.StaticImplicitSolver(precision='fp64', rel_tol=1e-11, load_steps=2, cg_tol=1e-22,
                                            cg_max_iter=20000, armijo_slack=1e-12, preconditioner='block')
    fe.initialize()
    J0, gX, _ = adjoint.compliance_gradient(fe)
    assert J0 > 0
    p = fe.assembly.plan
    X0 = p.X.reshape(-1).clone()
    GC_base = fe.assembly.GC.clone()
    rng = np.random.default_rng(21)
    h = 1e-6
    worst = 0.0
    for _ in range(8):
        d = torch.as_tensor(rng.uniform(-1, 1, X0.numel()), device='cuda')
        Jp = adjoint.compliance(fe, X=X0 + h * d, GC0=GC_base)[0].item()
        Jm = adjoint.compliance(fe, X=X0 - h * d, GC0=GC_base)[0].item()
        fd = (Jp - Jm) / (2 * h)
        an = float(torch.dot(gX, d))
        worst = max(worst, abs(fd - an) / (float(gX.norm()) * float(d.norm())))
    p.set_coordinates(X0)
    print('void C3D10 %d elements: max |fd - grad.d| / (|grad| |d|) = %.2e' % (mesh['conn'].shape[0], worst))
    assert worst < 1e-8
    # the element VJP at the converged state against the torch-autograd oracle
    U = p.gc_to_u(fe.assembly.GC.contiguous())
    lam = torch.as_tensor(rng.standard_normal(U.numel()), device='cuda')
    gx, gu, _ = adjoint.elem_vjp(p, U, lam, want_x=True, want_u=True, want_theta=False)
    ox, ou, _ = ad.internal_force_vjp('C3D10', mesh['conn'], mesh['nodes'], U.cpu().numpy().reshape(-1, 3),
                                      (0.5, 0.0, 100.0), lam.cpu().numpy().reshape(-1, 3))
    assert np.abs(gx.cpu().numpy() - ox.reshape(-1)).max() < 1e-11 * np.abs(ox).max()
    assert np.abs(gu.cpu().numpy() - ou.reshape(-1)).max() < 1e-11 * np.abs(ou).max()


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-10), ('fp32', 1e-5)])
def test_explicit_contact_ten_cubed_blocks(gpu, precision, tol):
    # a narrower, shifted upper block keeps the slave nodes off the master edges: the trajectory
    # is then a smooth function of round-off (aligned blocks leave facet choices to it when
    # they first meet)
    fe, info = workloads.two_blocks_contact(10, velocity=1.0, gap=0.01, jitter=0.1, seed=5, upper_width=0.9,
                                            offset=(0.037, 0.021))
    fe.solver = solver.DynamicExplicitSolver(time_end=1.0, time_per_storage=1.0, cfl_factor=0.9,
                                             precision=precision)
    fe.initialize()
    dt = fe.solver._deltaT
    nsteps = 60
    fe.solver = solver.DynamicExplicitSolver(time_end=dt * nsteps, time_per_storage=1.0, dt=dt, precision=precision)
    fe.initialize()
    assert fe.assembly.plan.n_nodes == 2 * 11 ** 3
    GV0 = workloads.initial_velocity(fe, info)
    GC0 = torch.zeros_like(fe.assembly.GC)
    fe.solver.solve(GC0=GC0, GV0=GV0)
    load = fe.assembly.get_load('contact')
    key = (dt, nsteps)
    if key not in _ORACLE_EXPLICIT:
        model, slave, master, area = oracle_two_blocks(info)
        active = []

        def fc(u):
            f = oc.contact_force(slave, master, model['nodes'], u, load.radius, load.penalty, load.depth_max, area)
            active.append(np.count_nonzero(f))
            return f

        GC_o, _, _ = osol.explicit_steps(model, np.zeros(GC0.shape[0]), GV0.cpu().numpy(), dt, nsteps,
                                         extra_force=fc)
        _ORACLE_EXPLICIT[key] = (GC_o, sum(a > 0 for a in active))
    GC_o, n_active = _ORACLE_EXPLICIT[key]
    err = rel_err(fe.solver.GC.cpu().numpy(), GC_o)
    print('explicit two 10^3 blocks %s: %d steps dt %.3e, rel err %.2e, contact active in %d steps' % (
        precision, nsteps, dt, err, n_active))
    assert n_active > 20                                              # contact was active
    assert err < tol


_ORACLE_EXPLICIT = {}


def test_serving_pressure_request_full_size_oracle_residual(gpu):
    req = serving.random_requests(16, 2.0, seed=0, e_range=(2e3, 2e5))[8]
    spec = req.spec
    assert spec['etype'] == 'C3D8' and spec['cells'] == (68, 47, 28) and spec['load'][0] == 'pressure'
    serving.SolveServer(precision='mixed').run([req], realtime=False)
    assert req.status == 'done', req.error
    ref, _, _ = serving.solve_one_at_a_time([req], precision='fp64', preconditioner='amg')
    assert ref[req.rid] is not None
    torch.cuda.empty_cache()
    mesh = serving.request_mesh(spec)
    _, mu, kappa = spec['material']
    p = spec['load'][1]
    om = oracle_model(mesh, nh(mu, kappa), spec['fixed'], pressure=p, load_set=serving.LOADED_SET)
    free = oasm.free_mask(om)
    conn, stype = facets(mesh, serving.LOADED_SET)
    u = oasm.gc_to_u(om, req.GC.numpy())
    F, _ = osurf.pressure_force_stiffness(stype, mesh['nodes'] + u, conn, p, onlyforce=True)
    f = np.zeros(mesh['nodes'].size)
    np.add.at(f, osurf.pressure_force_idx(conn), F)
    ratio = np.linalg.norm(oasm.assemble_force(om, u)) / np.linalg.norm(f[free])
    err = rel_err(req.GC.numpy(), ref[req.rid].numpy())
    print('serving request 8 (%d C3D8): server |R_oracle(u)| / |f_p| = %.2e in %d Newton iterations; '
          'rel diff vs the one-at-a-time FP64 multigrid solve %.2e' % (req.n_elements(), ratio, req.newton_iters, err))
    assert ratio <= 1e-9
    assert err < 1e-7
