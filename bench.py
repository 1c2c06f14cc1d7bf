#!/usr/bin/env python
"""Benchmark of the solid-mechanics hot path on one MI355X.

Default workload (``--workload beam``): cantilever beam 4x1x1 meshed as
128x32x32 C3D8 (131072 elements, 8 Gauss points each, 418176 free DOF),
Mooney-Rivlin C10=0.5 C01=0.1 D1=0.01, clamped at x=0, uniform dead traction
(0,0,-0.01) on x=4 ramped over 10 increments; Newton to ||R|| <= 1e-10
||f_ext|| with the tangent stored in FP32 and Jacobi-PCG with FP64 vectors
(FP64 element arithmetic, residual, energy and convergence test).  One "step"
= one full static solve (10 increments) from u = 0 with the model resident in
HBM.  The line also carries a bounded sample of the continuous-batching solve
server against the one-at-a-time FP64 path.

Other workloads (one JSON line each):
  cube      6000 C3D4 unit cube, NeoHookean mu=1 kappa=100, follower pressure 0.2,
            5 increments, FP64 (solves/s, p50/p99 over the timed solves);
  explicit  two 40^3-cell Kuhn-tet blocks (768000 C3D4), penalty contact, dt = 0.9 CFL,
            FP32 element arithmetic with an FP64 energy balance (steps/s);
  adjoint   48000 C3D10 block with a perturbed spherical void under pressure:
            compliance and its gradient w.r.t. every nodal coordinate
            (gradient evaluations/s) with a central-difference check;
  serving   a Poisson stream of static solve requests (C3D4/C3D8, 2e3..2e5
            elements) through the continuous-batching server vs one at a time.

Every line carries the roofline of the workload's dominant kernel (timed live
with HIP events on the stream the kernel is launched on) and the CPU baseline
(the FP64 oracle timed on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--workload', default='beam',
                    choices=['beam', 'cube', 'explicit', 'adjoint', 'serving', 'kernels', 'couple', 'contact'])
    ap.add_argument('--steps', type=int, default=None)
    ap.add_argument('--warmup', type=int, default=None)
    ap.add_argument('--nx', type=int, default=128)
    ap.add_argument('--ny', type=int, default=32)
    ap.add_argument('--nz', type=int, default=32)
    ap.add_argument('--increments', type=int, default=10)
    ap.add_argument('--precision', default='mixed', choices=['mixed', 'mixed-f32', 'mixed-ir', 'fp64'])
    ap.add_argument('--rel-tol', type=float, default=1e-10)
    ap.add_argument('--cg-tol', type=float, default=1e-8)
    ap.add_argument('--no-fp64-compare', action='store_true')
    ap.add_argument('--forcing', default='ew', choices=['ew', 'fixed'])
    ap.add_argument('--predictor', default='auto',
                    help="beam: load-increment predictor: none, linear, quadratic, cubic, auto or an order 0..6")
    ap.add_argument('--preconditioner', default='amg', choices=['amg', 'block', 'jacobi'],
                    help='beam: PCG preconditioner (smoothed-aggregation multigrid, node-block or the '
                         'reference\'s point Jacobi)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-plain', action='store_true', help='beam: skip the plain FP64 reference-algorithm solve')
    ap.add_argument('--spmv-reps', type=int, default=200)
    ap.add_argument('--max-newton', type=int, default=400, help='Newton iterations per solve (all increments) before failing')
    ap.add_argument('--serving-requests', type=int, default=None)
    ap.add_argument('--serving-qps', type=float, default=2.0)
    ap.add_argument('--serving-emax', type=float, default=None)
    ap.add_argument('--serving-emin', type=float, default=2e3)
    ap.add_argument('--serving-amg-builders', type=int, default=1,
                    help='serving: host threads building multigrid hierarchies concurrently')
    ap.add_argument('--serving-eager', action='store_true',
                    help='serving: launch each PCG chunk eagerly instead of replaying the captured HIP graph')
    ap.add_argument('--cells', type=int, default=40, help='explicit: cells per block edge')
    ap.add_argument('--explicit-overlap', action='store_true',
                    help='explicit: contact forces on a side stream, concurrent with the element pass')
    ap.add_argument('--explicit-graph', action='store_true',
                    help='explicit: steps replayed in 8-step chunks from a captured HIP graph')
    ap.add_argument('--fd-directions', type=int, default=64, help='adjoint: central-difference check directions')
    ap.add_argument('--couple-pressure', type=float, default=0.06, help='couple: pressure on surface_1_All')
    ap.add_argument('--couple-increments', type=int, default=4, help='couple: load increments')
    ap.add_argument('--couple-predictor', default='auto', help='cube / couple / contact: load-increment predictor')
    ap.add_argument('--cube-preconditioner', default='amg', choices=['block', 'jacobi', 'amg'],
                    help='cube: PCG preconditioner')
    ap.add_argument('--adjoint-preconditioner', default='amg', choices=['block', 'jacobi', 'amg'],
                    help='adjoint: preconditioner of the forward and the adjoint PCG')
    ap.add_argument('--couple-preconditioner', default='block', choices=['block', 'jacobi', 'amg'],
                    help='cube / couple / contact: PCG preconditioner')
    return ap.parse_args()


def log(*a):
    print('#', *a, file=sys.stderr, flush=True)


def event_time(torch, fn, reps, warm=3):
    """Average seconds per call of fn() on torch's current stream (where the
    library launches), bracketed by HIP events."""
    for _ in range(warm):
        fn()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def bound_call(name, *args):
    """A C-ABI call with its arguments converted once (device addresses, scalars), so a
    timing loop measures the kernels and not Python's per-call argument conversion."""
    from fea_amd import _native as nat
    fn = getattr(nat.lib(), name)
    conv = tuple(nat._arg(a) for a in args)

    def call():
        rc = fn(*conv)
        if rc != 0:
            nat.check(rc, name)
    call.keep = args
    return call


def spmv_bytes(plan, val_bytes):
    """Compulsory bytes of one node-grouped SpMV: values, the column stream (triple-blocked:
    one int32 per neighbour triple of each group; otherwise one per entry of each node's
    shared column list), the input vector read and the output written once (float64), the
    group heads' row pointers and the group table."""
    if plan.triple:
        cols = int(plan.spmv_col.numel())
    else:
        lens = np.diff(plan.rowptr_np)
        cols = int(lens[plan.grp.cpu().numpy()[:-1]].sum())
    return plan.nnz * val_bytes + cols * 4 + plan.n_free * 16 + (plan.n_groups + 1) * 8


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE (one counter
# per run) on `bench.py --workload kernels`, each scaled by the same counters' reading for a
# 256 MiB device copy (read bytes = 2 x FETCH_SIZE KiB x 1024, written = 1 x WRITE_SIZE KiB x 1024);
# the summary and the commands are in profiles/traffic_beam128x32x32_kernels.md.  A kernel
# whose workload or template differs from the measured one reports null.
MEASURED_TRAFFIC = {
    ('beam 128x32x32 C3D8', 'k_pcg_spmv<float,double,16,2>'): 156.13e6 + 3.59e6,
    ('beam 128x32x32 C3D8', 'k_amg_cheb0<float,16,2,0,0>'): 166.38e6 + 6.74e6,
    # the same two counters on `bench.py --workload explicit --steps 100 --warmup 20 --no-cpu`
    # (profiles/explicit_element_pass_counters.md)
    ('two 40^3-cell blocks C3D4', 'k_elem_r_tiles<float,4>'): 28.51e6 + 7.19e6,
}


def roofline(kernel, achieved, peak, unit, bound, traffic_key=None, **extra):
    d = {'bound': bound, 'kernel': kernel, 'achieved': achieved, 'peak': peak, 'unit': unit,
         'frac': achieved / peak, 'traffic': MEASURED_TRAFFIC.get(traffic_key)}
    d.update(extra)
    return d


def base_line(args, metric, value, unit, steps, warmup, elapsed, dtype, config, higher=True):
    return {'metric': metric, 'value': value, 'unit': unit, 'n_gpus': 1, 'steps': steps, 'warmup': warmup,
            'ms_per_step': 1e3 * elapsed / max(1, steps), 'higher_is_better': higher, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': dtype, 'data': 'synthetic (structured mesh generated in-process)',
            'config': config}


# ============================================================================ beam (C2)
_T0 = [time.perf_counter()]


def _progress(k, st):
    print('#   t=%.2fs increment %d newton %d cg %d alpha %.3g |R| %.3e' % (
        time.perf_counter() - _T0[0], k, st['newton_iters'], st['cg_iters'], st['alpha'], st['residual']),
        file=sys.stderr, flush=True)


def _pred(args, name=None):
    name = args.predictor if name is None else name
    return int(name) if name.isdigit() else name


def make_solver(precision, args, load_steps=None, progress=True, predictor=True, preconditioner='jacobi'):
    from fea_amd import solver
    return solver.StaticImplicitSolver(maximum_iteration=args.max_newton, precision=precision,
                                       rel_tol=args.rel_tol, load_steps=load_steps or args.increments,
                                       cg_tol=args.cg_tol, cg_max_iter=20000, armijo_slack=1e-12,
                                       progress=_progress if progress else None, forcing=args.forcing,
                                       predictor=predictor, preconditioner=preconditioner)


def run_solve(fe, torch):
    GC0 = torch.zeros_like(fe.assembly.GC)
    t0 = time.perf_counter()
    _T0[0] = t0
    GC = fe.solver.solve(GC0=GC0)
    torch.cuda.synchronize()
    if GC is False:
        raise RuntimeError('Newton solve failed to converge')
    return GC, time.perf_counter() - t0


def _triple_sw(plan):
    """Sub-wave width the library picks for a triple-blocked CSR (neighbour triples per node)."""
    work = plan.nnz / max(1, plan.n_free) / 3.0
    return 32 if work > 32 else (16 if work > 8 else 8)


def spmv_roofline(fe, torch, reps, mat_dtype):
    from fea_amd import _native as nat
    from fea_amd._native import F32, F64, call
    p = fe.assembly.plan
    val = p.csr_values(mat_dtype)
    x = torch.rand(p.n_free, dtype=torch.float64, device='cuda')
    q = torch.empty_like(x)
    ws = torch.zeros(int(nat.lib().fea_pcg_ws_size(p.n_free)), dtype=torch.float64, device='cuda')
    s = nat.stream_ptr()
    t = event_time(torch, bound_call('fea_pcg_spmv', mat_dtype, F64, p.n_free, p.nnz, p.spmv_groups, p.grp,
                                     p.rowptr, p.spmv_col, val, x, q, ws, s), reps)
    return t, spmv_bytes(p, 4 if mat_dtype == F32 else 8)


def smooth_roofline(fe, torch, reps, mat_dtype, half=False):
    """The multigrid V-cycle's dominant kernel: one level-0 Chebyshev step (r = b - A x with
    the tangent, fea_amg_smooth0, as the solver's V-cycle runs it, or with its binary16 copy,
    fea_amg_smooth0_half, the AMG option half_level0; node-block Jacobi, d and x updates) timed with
    HIP events over back-to-back launches (rocprof gives these dispatches the same duration;
    inside the solve, between the V-cycle's other kernels, the same kernel runs faster).
    Algorithmic bytes: the SpMV's (values, column stream, x read, group heads) plus b and d
    read, d and the new x written (float64) and the 6-float block inverse per group."""
    from fea_amd import _native as nat
    from fea_amd._native import F32, call
    p = fe.assembly.plan
    val = p.csr_values(mat_dtype)
    n = p.n_free
    binv = torch.zeros(6 * abs(p.spmv_groups), dtype=torch.float32, device='cuda')
    nat.ops().block_jacobi_inv(mat_dtype, p.spmv_groups, p.grp, p.rowptr, p.col, p.diag_slot, val, binv)
    b, x, d = (torch.rand(n, dtype=torch.float64, device='cuda') for _ in range(3))
    xo = torch.empty_like(x)
    s = nat.stream_ptr()
    half = half and p.spmv_groups < 0
    if half:
        v16 = torch.empty(4 * p.nnz // 3, dtype=torch.int16, device='cuda')
        sc = torch.zeros(2, dtype=torch.float64, device='cuda')
        call('fea_amg_half_values', mat_dtype, p.nnz, val, v16, sc, s)
        t = event_time(torch, bound_call('fea_amg_smooth0_half', n, p.nnz, p.spmv_groups, p.grp, p.rowptr,
                                         p.spmv_col, v16, sc, binv, b, x, xo, d, 0.5, 0.25, s), reps)
        vb = 2
    else:
        t = event_time(torch, bound_call('fea_amg_smooth0', mat_dtype, n, p.nnz, p.spmv_groups, p.grp, p.rowptr,
                                         p.spmv_col, val, binv, b, x, xo, d, 0.5, 0.25, s), reps)
        vb = 4 if mat_dtype == F32 else 8
    alg = spmv_bytes(p, vb) + 24 * n + 24 * p.n_groups
    return t, alg


def element_rate(fe, torch, reps=10, kernel_dtype=None):
    """Tangent assembly as the mixed solver runs it (FP64 element kernel + residual,
    CSR gather rounded to FP32), timed with events on the solver's stream."""
    from fea_amd._native import F32, F64
    p = fe.assembly.plan
    kd = F64 if kernel_dtype is None else kernel_dtype
    U = torch.zeros(3 * p.n_nodes, dtype=torch.float64, device='cuda')
    t = event_time(torch, lambda: p.tangent(U, 1.0, kd, val_dtype=F32), reps, warm=1)
    n_elem = sum(g.n for g in p.groups if g.kind == 'elem')
    return n_elem / t, t


def cpu_baseline_beam(args, plan, stats):
    """FP64 numpy oracle on a bounded sample, extrapolated to one full solve."""
    import scipy.sparse as sp
    from fea_amd import meshgen
    from oracle import assembly as oasm
    from oracle import materials as omat
    from oracle import solvers as osol
    threads = len(os.sched_getaffinity(0))
    nzs = max(1, args.nz // 16)
    mesh = meshgen.hex_block(args.nx, args.ny, nzs, size=(4.0, 1.0, nzs / args.nz))
    N = mesh['nodes'].shape[0]
    fixed = np.zeros((N, 3), dtype=bool)
    fixed[mesh['node_sets']['x0']] = True
    om = {'nodes': mesh['nodes'], 'fixed': fixed, 'pressures': [], 'tractions': [],
          'groups': [{'etype': 'C3D8', 'conn': mesh['conn'], 'material': omat.MooneyRivlin(0.5, 0.1, 0.01)}]}
    U = 1e-3 * np.random.default_rng(0).standard_normal(mesh['nodes'].shape)
    oasm.assemble_reduced(om, U)
    t0 = time.perf_counter()
    oasm.assemble_reduced(om, U)
    t_asm = (time.perf_counter() - t0) * (args.nz / nzs)
    t0 = time.perf_counter()
    oasm.total_energy(om, U)
    t_energy = (time.perf_counter() - t0) * (args.nz / nzs)
    n = plan.n_free
    A = sp.csr_matrix((np.ones(plan.nnz), plan.col_np, plan.rowptr_np), shape=(n, n))
    A = A + sp.diags(np.diff(plan.rowptr_np).astype(float))
    b = np.ones(n)
    t0 = time.perf_counter()
    osol.conjugate_gradient(A, b, np.zeros(n), tol=0.0, max_iter=20)
    t_cg = (time.perf_counter() - t0) / 20
    n_newton = stats['newton_iters']
    t_solve = n_newton * t_asm + stats['cg_iters'] * t_cg + stats['energy_evals'] * t_energy
    return {'value': 1.0 / t_solve, 'unit': 'solves/s', 'cores': threads, 'kind': 'port',
            'sample': ('FP64 numpy oracle: tangent assembly + free-DOF reduction of a %dx%dx%d C3D8 slice '
                       '(scaled x%d), energy of the slice scaled, 20 oracle-CG iterations on the full '
                       '418k-DOF CSR structure; extrapolated with the GPU solve\'s counts (%d Newton, %d CG, '
                       '%d energy evals): %.1f s per solve'
                       % (args.nx, args.ny, nzs, args.nz // nzs, n_newton, stats['cg_iters'],
                          stats['energy_evals'], t_solve)),
            't_assembly_s': t_asm, 't_cg_iter_s': t_cg, 't_energy_s': t_energy}


def cpu_direct_route_beam(args, plan, stats, t_asm, t_energy):
    """The reference's CPU route: every Newton iteration assembles the tangent and solves
    it directly (static_implicit.py:329 -> pypardiso; scipy's spsolve here).  The direct
    solve of the full 418k-DOF beam does not fit a bounded sample, so it is timed on the
    same 4:1:1 beam at three coarser resolutions and extrapolated with the fitted power of
    the DOF count; the assembly and energy times come from the slice sample."""
    import scipy.sparse.linalg as spla
    from fea_amd import meshgen
    from oracle import assembly as oasm
    from oracle import materials as omat
    sizes, times = [], []
    for m in (4, 6, 8):
        mesh = meshgen.hex_block(4 * m, m, m, size=(4.0, 1.0, 1.0))
        N = mesh['nodes'].shape[0]
        fixed = np.zeros((N, 3), dtype=bool)
        fixed[mesh['node_sets']['x0']] = True
        om = {'nodes': mesh['nodes'], 'fixed': fixed, 'pressures': [], 'tractions': [],
              'groups': [{'etype': 'C3D8', 'conn': mesh['conn'], 'material': omat.MooneyRivlin(0.5, 0.1, 0.01)}]}
        U = 1e-3 * np.random.default_rng(0).standard_normal(mesh['nodes'].shape)
        R, K = oasm.assemble_reduced(om, U)
        K = K.tocsc()
        t0 = time.perf_counter()
        spla.spsolve(K, -R)
        times.append(time.perf_counter() - t0)
        sizes.append(K.shape[0])
    slope = float(np.polyfit(np.log(sizes), np.log(times), 1)[0])
    t_direct = times[-1] * (plan.n_free / sizes[-1]) ** slope
    n_newton = stats['newton_iters']
    t_solve = n_newton * (t_asm + t_direct) + stats['energy_evals'] * t_energy
    return {'value': 1.0 / t_solve, 'unit': 'solves/s', 'cores': 1, 'kind': 'port',
            'sample': ('reference CPU route (Newton with a direct linear solve per iteration, '
                       'static_implicit.py:329): FP64 numpy-oracle tangent assembly of a 1/%d slice scaled to the '
                       'full beam (%.2f s) and scipy spsolve (SuperLU, single thread) timed on the 4:1:1 beam at '
                       '%s DOF (%s s), extrapolated as DOF^%.2f to %d DOF (%.0f s); x %d Newton iterations of the '
                       'GPU solve + %d energy evaluations: %.0f s per solve; single-threaded numpy/scipy'
                       % (args.nz // max(1, args.nz // 16), t_asm, sizes, ['%.3f' % t for t in times], slope,
                          plan.n_free, t_direct, n_newton, stats['energy_evals'], t_solve)),
            't_direct_solve_s': t_direct, 'direct_solve_dof_power': slope}


def bench_beam(args, torch):
    from fea_amd import workloads
    from fea_amd._native import F32, F64
    steps = 2 if args.steps is None else args.steps
    warmup = 1 if args.warmup is None else args.warmup
    fe, mesh = workloads.beam_traction(args.nx, args.ny, args.nz)
    fe.solver = make_solver(args.precision, args, predictor=_pred(args), preconditioner=args.preconditioner)
    fe.initialize()
    p = fe.assembly.plan
    log('model: %d elements, %d nodes, %d free DOF, nnz %d' % (mesh['conn'].shape[0], mesh['nodes'].shape[0],
                                                              p.n_free, p.nnz))
    for i in range(warmup):
        _, t = run_solve(fe, torch)
        log('warmup solve %d: %.2f s, stats %s' % (i, t, {k: v for k, v in fe.solver.stats.items()
                                                         if k != 'residual_history'}))
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    GC = None
    for i in range(steps):
        GC, t = run_solve(fe, torch)
        lat.append(t)
        log('timed solve %d: %.3f s' % (i, t))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    stats = dict(fe.solver.stats)
    value = steps / elapsed

    fp64 = {}
    if args.precision != 'fp64' and not args.no_fp64_compare:
        fe64, _ = workloads.beam_traction(args.nx, args.ny, args.nz)
        fe64.solver = make_solver('fp64', args, predictor=_pred(args), preconditioner=args.preconditioner)
        fe64.initialize()
        GC64, t64 = run_solve(fe64, torch)
        err = float((GC - GC64).abs().max() / GC64.abs().max())
        fp64 = {'fp64_solves_per_s': 1.0 / t64, 'fp64_solve_s': t64, 'rel_err_vs_fp64': err,
                'fp64_cg_iters': fe64.solver.stats['cg_iters'], 'fp64_newton_iters': fe64.solver.stats['newton_iters'],
                'speedup_vs_fp64': t64 / (elapsed / steps)}
        log('fp64 solve %.3f s, rel err %.2e' % (t64, err))
        del fe64

    plain = {}
    if args.precision != 'fp64' and not args.no_plain:
        # the reference's algorithm on the GPU in FP64: point-Jacobi (the reference's
        # scaling) CG at a fixed tolerance, each increment started from the last converged
        # state (no predictor), no inexact-Newton forcing
        fep, _ = workloads.beam_traction(args.nx, args.ny, args.nz)
        fep.solver = make_solver('fp64', args, predictor=False, preconditioner='jacobi')
        fep.solver.forcing = 'fixed'
        fep.initialize()
        GCp, tp = run_solve(fep, torch)
        plain = {'plain_fp64_solves_per_s': 1.0 / tp, 'plain_fp64_solve_s': tp,
                 'plain_fp64_rel_err': float((GC - GCp).abs().max() / GCp.abs().max()),
                 'plain_fp64_newton_iters': fep.solver.stats['newton_iters'],
                 'plain_fp64_cg_iters': fep.solver.stats['cg_iters'],
                 'speedup_vs_plain_fp64': tp / (elapsed / steps),
                 'plain_fp64_config': 'fp64, point Jacobi, no predictor, fixed CG tolerance %g on the squared '
                                      'scaled residual ratio, Newton rel tol %g' % (args.cg_tol, args.rel_tol)}
        log('plain fp64 solve %.3f s (%d Newton, %d CG)' % (tp, plain['plain_fp64_newton_iters'],
                                                            plain['plain_fp64_cg_iters']))
        del fep

    mat = F64 if args.precision == 'fp64' else F32
    kname_t = 'float' if mat == F32 else 'double'
    t_spmv, alg_bytes = spmv_roofline(fe, torch, args.spmv_reps, mat)
    achieved = alg_bytes / t_spmv / 1e9
    spmv_line = roofline('k_pcg_spmv<%s,double,%d,2> (fea_pcg_spmv)' % (kname_t, _triple_sw(p)),
                         achieved, HBM_PEAK_GBS, 'GB/s', 'hbm', alg_bytes_per_launch=alg_bytes,
                         avg_launch_us=t_spmv * 1e6,
                         traffic_key=('beam %dx%dx%d C3D8' % (args.nx, args.ny, args.nz),
                                      'k_pcg_spmv<%s,double,%d,2>' % (kname_t, _triple_sw(p))))
    if args.preconditioner == 'amg':
        t_sm, alg_sm = smooth_roofline(fe, torch, args.spmv_reps, mat)
        dominant = roofline('k_amg_cheb0<%s,%d,2,0,0> (fea_amg_smooth0: one level-0 Chebyshev step of the '
                            'multigrid V-cycle)' % (kname_t, _triple_sw(p)), alg_sm / t_sm / 1e9, HBM_PEAK_GBS,
                            'GB/s', 'hbm', alg_bytes_per_launch=alg_sm, avg_launch_us=t_sm * 1e6,
                            traffic_key=('beam %dx%dx%d C3D8' % (args.nx, args.ny, args.nz),
                                         'k_amg_cheb0<%s,%d,2,0,0>' % (kname_t, _triple_sw(p))))
    else:
        dominant = spmv_line
    elem_rate, t_tan = element_rate(fe, torch)
    elem_rate32, t_tan32 = element_rate(fe, torch, kernel_dtype=F32)
    serving = serving_sample(args, torch)
    cpu = None
    cg_route = None
    if not args.no_cpu:
        cg_route = cpu_baseline_beam(args, p, stats)
        cpu = cpu_direct_route_beam(args, p, stats, cg_route['t_assembly_s'], cg_route['t_energy_s'])

    line = base_line(args, 'static Newton solves/s (mixed FP32/FP64) at 1e-8 rel err', value, 'solves/s', steps,
                     warmup, elapsed, 'f64' if args.precision == 'fp64' else 'f32+f64',
                     {'workload': 'cantilever beam 4x1x1, %dx%dx%d C3D8, Mooney-Rivlin C10=0.5 C01=0.1 D1=0.01, '
                                  'clamped x=0, dead traction (0,0,-0.01) on x=4, %d increments, Newton rel tol %g'
                                  % (args.nx, args.ny, args.nz, args.increments, args.rel_tol),
                      'precision': args.precision, 'predictor': args.predictor,
                      'preconditioner': args.preconditioner, 'free_dof': p.n_free, 'nnz': p.nnz})
    line.update({
        'latency_p50_s': float(np.percentile(lat, 50)), 'latency_p99_s': float(np.percentile(lat, 99)),
        'newton_iters_per_solve': stats.get('newton_iters'), 'cg_iters_per_solve': stats.get('cg_iters'),
        'predictor_orders': stats.get('predictor_orders'),
        'energy_evals_per_solve': stats.get('energy_evals'),
        'amg': None if getattr(fe.solver, 'amg', None) is None else {
            'setup_s': fe.solver.amg.setup_time, 'levels': [l['n'] for l in fe.solver.amg.levels],
            'operator_complexity': fe.solver.amg.operator_complexity, 'smoother_degree': fe.solver.amg.degree,
            'note': 'hierarchy built once per model in initialize() from the tangent at u = 0 (like the CSR '
                    'pattern); every timed solve smooths with the current tangent'},
        'elem_tangent_evals_per_s': elem_rate, 'tangent_assembly_ms': 1e3 * t_tan,
        'elem_tangent_evals_per_s_f32_contraction': elem_rate32,
        'roofline': dominant,
        'roofline_pcg_spmv': spmv_line,
        'cpu_baseline': cpu,
        'cpu_cg_route_estimate': None if cg_route is None else {
            'value': cg_route['value'], 'unit': 'solves/s', 'sample': cg_route['sample']},
        'serving_sample': serving,
    })
    line.update(fp64)
    line.update(plain)
    return line


# ============================================================================ serving (C5)
_SERVING_EAGER = False      # --serving-eager: the tick path launches the PCG chunk without the graph
_SERVING_BUILDERS = 1       # --serving-amg-builders: concurrent multigrid host set-ups


def _serving_run(torch, reqs):
    """The stream through the server twice -- saturated (every request queued at
    t = 0: batching throughput) and at its Poisson arrival times (latency) --
    and the plain path: each request solved alone by the FP64 solver, timed, with
    a FIFO queue of those times under the same arrivals for its latencies."""
    import copy
    from fea_amd import serving
    warm = serving.random_requests(1, 0.0, seed=99, e_range=(500, 600))
    serving.SolveServer(precision='mixed').run(warm, realtime=False)

    def serve(rs, realtime):
        t0 = time.perf_counter()

        def on_tick(s):
            if s.rounds % 500 == 0:
                log('serving (%s) tick %d t=%.2f in flight %d' % ('poisson' if realtime else 'saturated', s.rounds,
                                                                  time.perf_counter() - t0, len(s.active)))
        srv = serving.SolveServer(precision='mixed', max_batch=64, use_graph=not _SERVING_EAGER,
                                  amg_builders=_SERVING_BUILDERS)
        srv.run(rs, realtime=realtime, on_round=on_tick)
        return srv, time.perf_counter() - t0

    sat = [serving.SolveRequest(rid=r.rid, spec=r.spec, arrival=0.0) for r in reqs]
    srv_sat, wall_sat = serve(sat, False)
    srv, wall = serve(reqs, True)
    done = [r for r in reqs if r.status == 'done']
    lat = [r.latency for r in done]
    secs, secs_amg, errs = [], [], []
    for r in sorted(reqs, key=lambda r: r.arrival):
        ref, sec, _ = serving.solve_one_at_a_time([r], precision='fp64')
        secs.append(sec[0])
        if r.status == 'done' and ref[r.rid] is not None:
            errs.append(float((r.GC - ref[r.rid]).abs().max() / ref[r.rid].abs().max()))
        ref_a, sec_a, _ = serving.solve_one_at_a_time([r], precision='fp64', preconditioner='amg')
        secs_amg.append(sec_a[0])
        if r.status == 'done' and ref_a[r.rid] is not None:
            errs.append(float((r.GC - ref_a[r.rid]).abs().max() / ref_a[r.rid].abs().max()))
        log('one-at-a-time fp64 request %d (%d elements): %.2f s point Jacobi, %.2f s multigrid'
            % (r.rid, r.n_elements(), sec[0], sec_a[0]))
    end, lat1 = 0.0, []
    for r, t in zip(sorted(reqs, key=lambda r: r.arrival), secs):
        end = max(end, r.arrival) + t
        lat1.append(end - r.arrival)
    end_a, lat1a = 0.0, []
    for r, t in zip(sorted(reqs, key=lambda r: r.arrival), secs_amg):
        end_a = max(end_a, r.arrival) + t
        lat1a.append(end_a - r.arrival)
    n_sat = sum(1 for r in sat if r.status == 'done')
    return {
        'requests': len(reqs), 'completed': len(done), 'qps_offered': None,
        'elements_range': [int(min(r.n_elements() for r in reqs)), int(max(r.n_elements() for r in reqs))],
        'saturated_solves_per_s': n_sat / wall_sat,
        'one_at_a_time_fp64_solves_per_s': len(reqs) / sum(secs),
        'saturated_throughput_ratio_vs_fp64_one_at_a_time': (n_sat / wall_sat) / (len(reqs) / sum(secs)),
        'one_at_a_time_fp64_amg_solves_per_s': len(reqs) / sum(secs_amg),
        'saturated_throughput_ratio_vs_fp64_amg_one_at_a_time': (n_sat / wall_sat) / (len(reqs) / sum(secs_amg)),
        'saturated_max_in_flight': int(max(srv_sat.batch_sizes)) if srv_sat.batch_sizes else 0,
        'poisson_solves_per_s': len(done) / wall, 'poisson_wall_s': wall,
        'latency_p50_s': float(np.percentile(lat, 50)) if lat else None,
        'latency_p99_s': float(np.percentile(lat, 99)) if lat else None,
        'one_at_a_time_fp64_latency_p50_s': float(np.percentile(lat1, 50)),
        'one_at_a_time_fp64_latency_p99_s': float(np.percentile(lat1, 99)),
        'one_at_a_time_fp64_makespan_s': end,
        'one_at_a_time_fp64_amg_latency_p50_s': float(np.percentile(lat1a, 50)),
        'one_at_a_time_fp64_amg_latency_p99_s': float(np.percentile(lat1a, 99)),
        'max_in_flight': int(max(srv.batch_sizes)) if srv.batch_sizes else 0, 'ticks': srv.rounds,
        'peak_admitted_gb': srv.peak_used / 1e9,
        'max_rel_err_vs_fp64': max(errs) if errs else None,
        'not_completed': [{'rid': r.rid, 'elements': r.n_elements(), 'status': r.status, 'error': r.error,
                           'newton_iters': r.newton_iters, 'cg_iters': r.cg_iters}
                          for r in reqs if r.status != 'done'],
        'saturated_not_completed': [{'rid': r.rid, 'status': r.status, 'error': r.error} for r in sat
                                    if r.status != 'done'],
    }, srv


def serving_sample(args, torch):
    """Bounded sample of the solve server inside the default line."""
    from fea_amd import serving
    n = 8 if args.serving_requests is None else args.serving_requests
    if n <= 0:
        return None
    emax = 3e4 if args.serving_emax is None else args.serving_emax
    reqs = serving.random_requests(n, args.serving_qps, seed=1, e_range=(args.serving_emin, emax))
    out, _ = _serving_run(torch, reqs)
    out['qps_offered'] = args.serving_qps
    out['sample'] = '%d requests, elements log-uniform in [%.0e, %.0e], Poisson arrivals at %.1f/s' % (
        n, args.serving_emin, emax, args.serving_qps)
    return out


def bench_serving(args, torch):
    from fea_amd import serving
    from fea_amd._native import F32, F64, call
    from fea_amd import _native as nat
    from fea_amd.solver._linear_solver import BatchedPCG, PcgVectors
    n = 24 if args.serving_requests is None else args.serving_requests
    emax = 2e5 if args.serving_emax is None else args.serving_emax
    reqs = serving.random_requests(n, args.serving_qps, seed=0, e_range=(args.serving_emin, emax))
    out, srv = _serving_run(torch, reqs)
    out['qps_offered'] = args.serving_qps
    # roofline: the batched SpMV stage over the stream's systems at once (tangents at u = 0)
    sub = sorted(reqs, key=lambda r: r.n_elements())
    entries, keep, total_bytes = [], [], 0
    for r in sub:
        fe, _ = serving.request_model(r.spec)
        fe.assembly.initialize()
        p = fe.assembly.plan
        U = torch.zeros(3 * p.n_nodes, dtype=torch.float64, device='cuda')
        _, K = p.tangent(U, 1.0, F64, val_dtype=F32)
        v = PcgVectors(p.n_free, F64, 'cuda')
        v.b.copy_(torch.rand_like(v.b))
        entries.append((p, K.clone(), v, 1e-30, 10))
        keep.append(fe)
        total_bytes += spmv_bytes(p, 4)
    bp = BatchedPCG(F32, F64, 'cuda')
    bp.start(entries)
    nsys, st, vt = bp._layout(entries)
    s = nat.stream_ptr()
    t = event_time(torch, lambda: call('fea_pcg_batch_spmv', F32, F64, nsys, bp._sys_dev, bp._ss_dev, st, s), 100)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_requests(sub[:3])
    line = base_line(args, 'static solves/s through the continuous-batching server (mixed FP32/FP64), saturated',
                     out['saturated_solves_per_s'], 'solves/s', n, 1, n / out['saturated_solves_per_s'], 'f32+f64',
                     {'workload': 'Poisson stream of %d static solve requests at %.1f/s: C3D4/C3D8 blocks, '
                                  'elements log-uniform in [%.0e, %.0e], aspect ratios U[1,4], NH/MR +-50%%, '
                                  'one face clamped, follower pressure on the five others or a traction on the '
                                  'opposite face, 2 increments' % (n, args.serving_qps, args.serving_emin, emax)})
    line.update(out)
    line['roofline'] = roofline('k_batch_spmv<float,double> (fea_pcg_batch_spmv), all %d systems in one launch' % nsys,
                                total_bytes / t / 1e9, HBM_PEAK_GBS, 'GB/s', 'hbm', alg_bytes_per_launch=total_bytes,
                                avg_launch_us=t * 1e6)
    line['cpu_baseline'] = cpu
    return line


def cpu_baseline_requests(reqs):
    """Oracle direct-solve Newton (FP64, CPU) of the given requests, one at a time."""
    from fea_amd import serving
    from oracle import solvers as osol
    from tests.helpers import mr, nh, oracle_model
    t0 = time.perf_counter()
    for r in reqs:
        mesh = serving.request_mesh(r.spec)
        m = r.spec['material']
        mat = nh(m[1], m[2]) if m[0] == 'NH' else mr(m[1], m[2], m[3])
        kind, val = r.spec['load']
        om = oracle_model(mesh, mat, r.spec['fixed'], pressure=val if kind == 'pressure' else None,
                          traction=val if kind == 'traction' else None, load_set=serving.LOADED_SET)
        osol.static_solve(om, increments=r.spec['load_steps'], rel_tol=1e-10, linear='direct',
                          armijo_slack_rel=1e-12)
    t = time.perf_counter() - t0
    return {'value': len(reqs) / t, 'unit': 'solves/s', 'cores': len(os.sched_getaffinity(0)), 'kind': 'port',
            'sample': 'FP64 numpy oracle Newton with direct solves on the %d smallest requests (%s elements), '
                      'one at a time: %.1f s' % (len(reqs), ', '.join(str(r.n_elements()) for r in reqs), t)}


# ============================================================================ cube (C1)
def bench_cube(args, torch):
    from fea_amd import workloads, solver
    from fea_amd._native import F64
    from oracle import solvers as osol
    from tests.helpers import nh, oracle_model
    steps = 20 if args.steps is None else args.steps
    warmup = 2 if args.warmup is None else args.warmup
    fe, mesh = workloads.cube_pressure(10, pressure=0.2, jitter=0.05, seed=0)
    fe.solver = solver.StaticImplicitSolver(precision='fp64', rel_tol=1e-10, load_steps=5, cg_tol=1e-12,
                                            cg_max_iter=20000, armijo_slack=1e-12,
                                            predictor=_pred(args, args.couple_predictor),
                                            preconditioner=args.cube_preconditioner)
    fe.initialize()
    p = fe.assembly.plan
    for _ in range(warmup):
        run_solve(fe, torch)
    lat = []
    t_start = time.perf_counter()
    for _ in range(steps):
        GC, t = run_solve(fe, torch)
        lat.append(t)
    elapsed = time.perf_counter() - t_start
    stats = dict(fe.solver.stats)
    t_spmv, alg = spmv_roofline(fe, torch, 500, F64)
    cpu = None
    err = None
    if not args.no_cpu:
        om = oracle_model(mesh, nh(1.0, 100.0), 'z0', pressure=0.2, load_set='z1')
        t0 = time.perf_counter()
        GC_o, _ = osol.static_solve(om, increments=5, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12)
        t_cpu = time.perf_counter() - t0
        err = float(np.abs(GC.cpu().numpy() - GC_o).max() / np.abs(GC_o).max())
        cpu = {'value': 1.0 / t_cpu, 'unit': 'solves/s', 'cores': len(os.sched_getaffinity(0)), 'kind': 'port',
               'sample': 'one full FP64 oracle solve of the same model (Newton with direct linear solves): %.2f s'
                         % t_cpu}
    line = base_line(args, 'static Newton solves/s (FP64) at 1e-8 rel err', steps / elapsed, 'solves/s', steps,
                     warmup, elapsed, 'f64',
                     {'workload': 'unit cube 10x10x10 Kuhn-split C3D4 (6000 elements, 3630 free DOF), NeoHookean '
                                  'mu=1 kappa=100, z=0 fixed, follower pressure 0.2 on z=1, 5 increments, Newton rel '
                                  'tol 1e-10, 5% interior jitter', 'precision': 'fp64',
                      'predictor': args.couple_predictor, 'preconditioner': args.cube_preconditioner})
    line.update({'latency_p50_s': float(np.percentile(lat, 50)), 'latency_p99_s': float(np.percentile(lat, 99)),
                 'newton_iters_per_solve': stats['newton_iters'], 'cg_iters_per_solve': stats['cg_iters'],
                 'predictor_orders': stats.get('predictor_orders'),
                 'rel_err_vs_oracle': err,
                 'roofline': roofline('k_pcg_spmv<double,double,%d,2> (fea_pcg_spmv)' % _triple_sw(p),
                                      alg / t_spmv / 1e9,
                                      HBM_PEAK_GBS, 'GB/s', 'hbm', alg_bytes_per_launch=alg,
                                      avg_launch_us=t_spmv * 1e6,
                                      note='1.9 MB working set: launch-latency bound, not bandwidth bound'),
                 'cpu_baseline': cpu})
    return line


# ============================================================================ Couple model
def bench_couple(args, torch):
    """The reference's Couple benchmark model (tests/pressure_test/C3D4Less.inp: 33 612 C3D4
    of a soft actuator, NeoHookean from its INP, follower pressure on surface_1_All, the base
    fixed, the head coupled to a reference point at (0, 0, 80)), solved from rest."""
    from fea_amd import workloads
    from fea_amd._native import F32, F64
    steps = 1 if args.steps is None else args.steps
    warmup = 0 if args.warmup is None else args.warmup
    path = os.path.join(ROOT, 'tests', 'golden', 'inp', 'C3D4Less.inp.gz')
    fe, inp = workloads.inp_couple(path, pressure=args.couple_pressure)
    # a plain linear extrapolation of the actuator's large rotation starts far from the next
    # equilibrium; 'auto' falls back to the previous one when the extrapolation's energy is higher
    fe.solver = make_solver(args.precision, args, load_steps=args.couple_increments,
                            predictor=_pred(args, args.couple_predictor),
                            preconditioner=args.couple_preconditioner)
    fe.initialize()
    p = fe.assembly.plan
    for _ in range(warmup):
        run_solve(fe, torch)
    lat = []
    t_start = time.perf_counter()
    for _ in range(steps):
        GC, t = run_solve(fe, torch)
        lat.append(t)
    elapsed = time.perf_counter() - t_start
    stats = dict(fe.solver.stats)
    rp = GC[-6:].cpu().numpy()
    mat = F64 if args.precision == 'fp64' else F32
    t_spmv, alg = spmv_roofline(fe, torch, 500, mat)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_couple(inp, GC.cpu().numpy(), args.couple_pressure, stats['newton_iters'])
    line = base_line(args, 'static Newton solves/s of the reference Couple model', steps / elapsed, 'solves/s',
                     steps, warmup, elapsed, 'f64' if args.precision == 'fp64' else 'f32+f64',
                     {'workload': 'C3D4Less.inp (reference pressure test): %d C3D4, NeoHookean mu=%.4g kappa=%.4g, '
                                  'pressure %g on surface_1_All, surface_0_Bottom fixed, surface_0_Head coupled to a '
                                  'reference point at (0,0,80), %d increments, Newton rel tol %g'
                                  % (inp.part['final_model'].n_elements(), *inp.material['Rubber'].mat_para,
                                     args.couple_pressure, args.couple_increments, args.rel_tol),
                      'precision': args.precision, 'predictor': args.couple_predictor,
                      'preconditioner': args.couple_preconditioner, 'free_dof': p.n_free,
                      'nnz': p.nnz})
    line['data'] = 'the reference\'s INP mesh (tests/pressure_test/C3D4Less.inp, committed gzipped as a fixture)'
    line.update({'latency_p50_s': float(np.percentile(lat, 50)),
                 'newton_iters_per_solve': stats['newton_iters'], 'cg_iters_per_solve': stats['cg_iters'],
                 'predictor_orders': stats.get('predictor_orders'),
                 'reference_point_v_z': rp.tolist(),
                 'roofline': roofline('k_pcg_spmv (fea_pcg_spmv) on the reduced CSR', alg / t_spmv / 1e9,
                                      HBM_PEAK_GBS, 'GB/s', 'hbm', alg_bytes_per_launch=alg,
                                      avg_launch_us=t_spmv * 1e6,
                                      note='%.1f MB working set: launch-latency bound' % (alg / 1e6)),
                 'cpu_baseline': cpu})
    return line


def cpu_baseline_couple(inp, GC, pressure, newton_iters):
    """One FP64 oracle Newton iteration of the same model (Couple-reduced assembly + direct
    solve of the Jacobi-scaled system), timed at the GPU's solution; solves/s extrapolated
    with the GPU run's Newton count."""
    from oracle import couple as cp
    from oracle import solvers as osol
    from tests.helpers import FACE_NODES, nh
    part = inp.part['final_model']
    X = part.nodes[:, 1:].numpy().astype(np.float64)
    conn = part.elems['C3D4'][:, 1:].astype(np.int64)
    fixed = np.zeros((X.shape[0], 3), dtype=bool)
    fixed[sorted(part.sets_nodes['surface_0_Bottom'])] = True
    pos = {eid: i for i, eid in enumerate(part.elems['C3D4'][:, 0])}
    facets = np.concatenate([conn[[pos[e] for e in ids]][:, FACE_NODES['C3D4'][f]]
                             for ids, f in part.surfaces['surface_1_All']])
    om = {'nodes': X, 'fixed': fixed,
          'groups': [{'etype': 'C3D4', 'conn': conn, 'material': nh(*inp.material['Rubber'].mat_para)}],
          'pressures': [{'stype': 'T3', 'conn': facets, 'p': pressure}], 'tractions': [],
          'rps': [{'node': (0.0, 0.0, 80.0)}],
          'couples': [{'nodes': np.array(sorted(part.sets_nodes['surface_0_Head'])), 'rp': 0}]}
    cp.assemble_reduced(om, GC)                              # warm-up (Gauss precompute)
    t0 = time.perf_counter()
    R, K = cp.assemble_reduced(om, GC)
    osol.jacobi_scaled_solve(K, -R, np.zeros_like(R), linear='direct')
    t_it = time.perf_counter() - t0
    return {'value': 1.0 / (t_it * newton_iters), 'unit': 'solves/s', 'cores': len(os.sched_getaffinity(0)),
            'kind': 'port', 'residual_at_gpu_solution_rel': float(np.linalg.norm(R) / np.linalg.norm(
                cp.assemble_force(om, np.zeros_like(GC)))),
            'sample': 'one FP64 oracle Newton iteration (reduced assembly + direct solve): %.2f s, x %d Newton '
                      'iterations of the GPU solve (extrapolated)' % (t_it, newton_iters)}


# ============================================================================ contact barrier
# FP64 operations of one candidate pair's gradient + Hessian in fea_barrier_assemble: 324
# Hessian entries x 9 Gauss pairs x 34 (six outer-product terms and the d2D, d2L terms),
# plus the d2n and projection terms (~9 000)
BARRIER_FLOP_PER_PAIR = 324 * 9 * 34 + 9000


def bench_contact(args, torch):
    """The reference's self-contact model (tests/gradient_test/C3D4.inp with test_grad.py's
    loads: pressure, four ContactSelf surfaces, Couple), solved from rest; the line also
    carries the distance to the reference's golden vector C3D4_results.npy."""
    from fea_amd import workloads
    from fea_amd._native import F32, F64
    steps = 1 if args.steps is None else args.steps
    warmup = 0 if args.warmup is None else args.warmup
    path = os.path.join(ROOT, 'tests', 'golden', 'inp', 'gradient_C3D4.inp.gz')
    surfaces = ['surface_%d_All' % i for i in range(4)]
    fe, inp = workloads.inp_couple(path, pressure=0.06, rp=(0.0, 0.0, 70.0), contact_surfaces=surfaces)
    fe.solver = make_solver(args.precision, args, load_steps=args.couple_increments,
                            predictor=_pred(args, args.couple_predictor),
                            preconditioner=args.couple_preconditioner)
    fe.initialize()
    p = fe.assembly.plan
    for _ in range(warmup):
        run_solve(fe, torch)
    lat = []
    t_start = time.perf_counter()
    for _ in range(steps):
        GC, t = run_solve(fe, torch)
        lat.append(t)
    elapsed = time.perf_counter() - t_start
    stats = dict(fe.solver.stats)
    gold = np.load(os.path.join(ROOT, 'tests', 'golden', 'gradient_C3D4_results.npy'), allow_pickle=False)
    g = GC.cpu().numpy()
    # roofline: the barrier gradient/Hessian kernel over the staged pairs at the solution
    U = p.gc_to_u(GC)
    Un = U[:3 * p.n_nodes]
    inner = p.inner if hasattr(p, 'inner') else p._plan
    groups = [gr for gr in inner.groups if gr.kind == 'barrier']
    ks = inner.kstage(F64)
    n_staged = sum(b.spi.size for b in p.barriers)
    # staged pairs that are candidates at the solution: the ones the kernel evaluates
    x = (inner.X.reshape(-1) + Un).reshape(-1, 3).cpu().numpy()
    n_cand = 0
    for b in p.barriers:
        c1 = x[b.c1_np].sum(axis=1) / 3.0
        c2 = x[b.c2_np].sum(axis=1) / 3.0
        d2 = ((c1[b.spi] - c2[b.spj]) ** 2).sum(axis=1)
        n_cand += int((d2 <= b.radius ** 2).sum())

    def assemble():
        for gr in groups:
            gr.load.assemble(Un, gr.dst_k, ks, F64, gr.dst_r, inner.rstage)
    t_asm = event_time(torch, assemble, 20)
    flops = BARRIER_FLOP_PER_PAIR * n_cand
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_contact(fe, inp, g, stats['newton_iters'])
    line = base_line(args, 'static Newton solves/s of the reference self-contact model', steps / elapsed,
                     'solves/s', steps, warmup, elapsed, 'f64' if args.precision == 'fp64' else 'f32+f64',
                     {'workload': 'gradient_test/C3D4.inp (reference): %d C3D4, pressure 0.06 on surface_1_All, '
                                  'ContactSelf on surface_0..3_All, surface_0_Bottom fixed, surface_0_Head coupled to '
                                  'a reference point at (0,0,70), %d increments, Newton rel tol %g'
                                  % (inp.part['final_model'].n_elements(), args.couple_increments, args.rel_tol),
                      'precision': args.precision, 'predictor': args.couple_predictor,
                      'preconditioner': args.couple_preconditioner, 'free_dof': p.n_free})
    line['data'] = 'the reference\'s INP mesh (tests/gradient_test/C3D4.inp, committed gzipped as a fixture)'
    line.update({'latency_p50_s': float(np.percentile(lat, 50)),
                 'newton_iters_per_solve': stats['newton_iters'], 'cg_iters_per_solve': stats['cg_iters'],
                 'predictor_orders': stats.get('predictor_orders'),
                 'pattern_rebuilds': p.rebuilds, 'staged_pairs': int(n_staged), 'candidate_pairs_at_solution': n_cand,
                 'rel_diff_vs_reference_golden': float(np.abs(g - gold).max() / np.abs(gold).max()),
                 'reference_point_v_z': g[-6:].tolist(),
                 'roofline': roofline('k_bar_assemble_wg (fea_barrier_assemble), all staged pairs of the 4 surfaces',
                                      flops / t_asm / 1e12, FP64_PEAK_TFLOPS, 'TFLOP/s', 'mfma',
                                      alg_flop_per_launch=flops, avg_launch_us=t_asm * 1e6,
                                      note='FP64 vector arithmetic (no FP64 MFMA form); %d of the %d staged '
                                           'pairs are candidates at the solution' % (n_cand, n_staged)),
                 'cpu_baseline': cpu})
    return line


def cpu_baseline_contact(fe, inp, GC, newton_iters):
    """One FP64 oracle Newton iteration of the same model, bounded: the Couple-reduced
    element + pressure assembly and a direct solve timed in full, the barrier terms timed on
    a sample of 2 000 candidate pairs and scaled to all; solves/s extrapolated with the GPU
    run's Newton count."""
    from oracle import barrier as ob
    from oracle import couple as cp
    from oracle import solvers as osol
    from tests.helpers import FACE_NODES, nh
    part = inp.part['final_model']
    X = part.nodes[:, 1:].numpy().astype(np.float64)
    conn = part.elems['C3D4'][:, 1:].astype(np.int64)
    fixed = np.zeros((X.shape[0], 3), dtype=bool)
    fixed[sorted(part.sets_nodes['surface_0_Bottom'])] = True
    pos = {eid: i for i, eid in enumerate(part.elems['C3D4'][:, 0])}
    fac = lambda s: np.concatenate([conn[[pos[e] for e in ids]][:, FACE_NODES['C3D4'][f]]
                                    for ids, f in part.surfaces[s]])
    om = {'nodes': X, 'fixed': fixed,
          'groups': [{'etype': 'C3D4', 'conn': conn, 'material': nh(*inp.material['Rubber'].mat_para)}],
          'pressures': [{'stype': 'T3', 'conn': fac('surface_1_All'), 'p': 0.06}], 'tractions': [],
          'rps': [{'node': (0.0, 0.0, 70.0)}],
          'couples': [{'nodes': np.array(sorted(part.sets_nodes['surface_0_Head'])), 'rp': 0}]}
    cp.assemble_reduced(om, GC)
    t0 = time.perf_counter()
    R, K = cp.assemble_reduced(om, GC)
    osol.jacobi_scaled_solve(K, -R, np.zeros_like(R), linear='direct')
    t_base = time.perf_counter() - t0
    U, _ = cp.state(om, GC)
    x = X + U
    n_pairs, t_bar = 0, 0.0
    for s in ['surface_%d_All' % i for i in range(4)]:
        f = fac(s)
        pairs = ob.candidate_pairs(ob.centroids(x, f), 3.0)
        ratio = ob.self_ratio(X, f, pairs)
        keep = ratio.sum((1, 2)) > 0
        pairs, ratio = pairs[keep], ratio[keep]
        n_pairs += pairs.shape[0]
        m = min(2000, pairs.shape[0])
        if m:
            t1 = time.perf_counter()
            ob.barrier((X, X), x, f, x, f, pairs[:m], ratio[:m], order=2)
            t_bar += (time.perf_counter() - t1) * pairs.shape[0] / m
    t_it = t_base + t_bar
    return {'value': 1.0 / (t_it * newton_iters), 'unit': 'solves/s', 'cores': len(os.sched_getaffinity(0)),
            'kind': 'port',
            'sample': 'one FP64 oracle Newton iteration: reduced element/pressure/Couple assembly + direct solve '
                      '%.2f s, barrier terms of %d active pairs %.2f s (timed on 2000 per surface, scaled), x %d '
                      'Newton iterations of the GPU solve (extrapolated)' % (t_base, n_pairs, t_bar, newton_iters)}


# ============================================================================ explicit (C3)
def bench_explicit(args, torch):
    from fea_amd import workloads, solver
    from fea_amd._native import F32, F64
    from fea_amd import _native as nat
    steps = 20000 if args.steps is None else args.steps
    warmup = 200 if args.warmup is None else args.warmup
    fe, info = workloads.two_blocks_contact(args.cells)
    fe.solver = solver.DynamicExplicitSolver(time_end=1.0, time_per_storage=1e30, cfl_factor=0.9, store=False,
                                             precision='fp32', overlap=args.explicit_overlap,
                                             use_graph=args.explicit_graph)
    fe.initialize()
    sol = fe.solver
    p = fe.assembly.plan
    dt = sol._deltaT
    GV0 = workloads.initial_velocity(fe, info)
    U = p.gc_to_u(torch.zeros_like(fe.assembly.GC))
    vh = GV0.clone()
    acc = torch.zeros_like(vh)
    e0 = sol.energy_balance(U, vh, None, dt)
    sol.run_steps(U, vh, acc, dt, warmup, first=True)
    torch.cuda.synchronize()
    hist = [e0]
    t_start = time.perf_counter()
    chunk = max(1, steps // 10)
    done = 0
    while done < steps:
        m = min(chunk, steps - done)
        sol.run_steps(U, vh, acc, dt, m)
        done += m
        torch.cuda.synchronize()
        log('explicit step %d t=%.2fs' % (warmup + done, time.perf_counter() - t_start))
    elapsed = time.perf_counter() - t_start
    sol._check_contacts()
    hist.append(sol.energy_balance(U, vh, acc, dt))
    drift = abs(hist[-1]['total'] - hist[0]['total']) / hist[0]['kinetic']
    # roofline: the FP32 residual kernel of one block pair (one step's element pass)
    g = [g for g in p.groups if g.kind == 'elem']
    E = sum(x.n for x in g)

    tl = sol._tiles
    fn = nat.raw('fea_elem_residual_tiles')

    def residual():
        s = nat.stream_ptr()
        for a in tl['elem_args']:
            nat.check_rc(fn(a[0], a[1], a[2], a[3], a[4], U.data_ptr(), a[5], a[6], a[7], a[8], a[9], a[10],
                            a[11], a[12], a[13], s), 'fea_elem_residual_tiles')
    t_res = event_time(torch, residual, 200)
    alg = E * 4 * 4 + p.n_nodes * 3 * 8 * 3                 # conn + X + U read, nodal force written
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_explicit(fe, info, dt)
    line = base_line(args, 'explicit central-difference steps/s with penalty contact', steps / elapsed, 'steps/s',
                     steps, warmup, elapsed, 'f32+f64',
                     {'workload': 'two %d^3-cell Kuhn-tet blocks (%d C3D4), NeoHookean mu=1 kappa=50 rho=1, upper '
                                  'block at 0.1 onto the lower (base fixed), node-to-surface penalty 1e3, dt = 0.9 CFL '
                                  '(%.3e)' % (args.cells, E, dt), 'precision': 'fp32 element forces, fp64 state',
                      'contact_stream': 'side' if args.explicit_overlap else 'same',
                      'graph_chunks': bool(args.explicit_graph)})
    line.update({'dt': dt, 'simulated_time': dt * (steps + warmup),
                 'energy': {'initial': hist[0], 'final': hist[-1], 'drift_rel_to_initial_kinetic': drift},
                 'roofline': roofline('k_elem_r_tiles<float,4> (fea_elem_residual_tiles, one step\'s element pass: '
                                      'both blocks)',
                                      alg / t_res / 1e9, HBM_PEAK_GBS, 'GB/s', 'hbm', alg_bytes_per_launch=alg,
                                      avg_launch_us=t_res * 1e6,
                                      traffic_key=('two %d^3-cell blocks C3D4' % args.cells, 'k_elem_r_tiles<float,4>')),
                 'cpu_baseline': cpu})
    return line


def cpu_baseline_explicit(fe, info, dt):
    """One oracle residual evaluation of the full two-block model = one step's element work."""
    from oracle import assembly as oasm
    from oracle import materials as omat
    lower, upper = info['meshes']
    nl = lower['nodes'].shape[0]
    nodes = np.concatenate([lower['nodes'], upper['nodes']])
    fixed = np.zeros((nodes.shape[0], 3), dtype=bool)
    fixed[lower['node_sets']['z0']] = True
    om = {'nodes': nodes, 'fixed': fixed, 'pressures': [], 'tractions': [],
          'groups': [{'etype': 'C3D4', 'conn': lower['conn'], 'material': omat.NeoHookean(1.0, 50.0), 'rho': 1.0},
                     {'etype': 'C3D4', 'conn': upper['conn'] + nl, 'material': omat.NeoHookean(1.0, 50.0),
                      'rho': 1.0}]}
    U = 1e-4 * np.random.default_rng(0).standard_normal(nodes.shape)
    oasm.assemble_force(om, U)
    t0 = time.perf_counter()
    oasm.assemble_force(om, U)
    t = time.perf_counter() - t0
    return {'value': 1.0 / t, 'unit': 'steps/s', 'cores': len(os.sched_getaffinity(0)), 'kind': 'port',
            'sample': 'one FP64 numpy-oracle residual evaluation of the full two-block mesh (the element work of '
                      'one central-difference step; update and contact omitted): %.2f s' % t}


# ============================================================================ adjoint (C4)
def bench_adjoint(args, torch):
    from fea_amd import adjoint, workloads, solver
    from fea_amd._native import F64
    from fea_amd import _native as nat
    steps = 2 if args.steps is None else args.steps
    warmup = 1 if args.warmup is None else args.warmup
    fe, mesh = workloads.void_block_pressure()
    fe.solver = solver.StaticImplicitSolver(precision='fp64', rel_tol=1e-11, load_steps=2, cg_tol=1e-22,
                                            cg_max_iter=20000, armijo_slack=1e-12,
                                            preconditioner=args.adjoint_preconditioner)
    fe.initialize()
    p = fe.assembly.plan
    E = sum(g.n for g in p.groups if g.kind == 'elem')
    for _ in range(warmup):
        adjoint.compliance_gradient(fe)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(steps):
        J0, gX, gth = adjoint.compliance_gradient(fe)
        torch.cuda.synchronize()
        log('gradient %d: %.2f s' % (i, time.perf_counter() - t_start))
    elapsed = time.perf_counter() - t_start
    # central-difference check on seeded random directions: two static solves per direction,
    # each started from the unperturbed equilibrium at the full load
    X0 = p.X.reshape(-1).clone()
    GC_base = fe.assembly.GC.clone()
    rng = np.random.default_rng(13)
    h = 1e-6
    errs, nerrs = [], []
    t_fd = time.perf_counter()
    for k in range(args.fd_directions):
        d = torch.as_tensor(rng.uniform(-1, 1, X0.numel()), device='cuda')
        Jp = adjoint.compliance(fe, X=X0 + h * d, GC0=GC_base)[0].item()
        Jm = adjoint.compliance(fe, X=X0 - h * d, GC0=GC_base)[0].item()
        if k % 16 == 15:
            log('finite differences: %d directions, %.1f s' % (k + 1, time.perf_counter() - t_fd))
        fd = (Jp - Jm) / (2 * h)
        an = float(torch.dot(gX, d))
        errs.append(abs(fd - an) / max(abs(an), 1e-300))
        nerrs.append(abs(fd - an) / (float(gX.norm()) * float(d.norm())))
    p.set_coordinates(X0)
    # roofline: the adjoint element kernel (d(lam.f_int)/dX, /dU, /dtheta of the C3D10 group)
    U = torch.rand(3 * p.n_nodes, dtype=torch.float64, device='cuda') * 1e-3
    lam = torch.rand_like(U)
    t_vjp = event_time(torch, lambda: adjoint.elem_vjp(p, U, lam, want_x=True, want_u=True, want_theta=True), 20)
    flops = E * 4 * 2 * (99 * 10 + 27 * 10 * 10)            # the C3D10 tangent contraction count per pass
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_adjoint(mesh, p, E)
    line = base_line(args, 'compliance shape-gradient evaluations/s (static solve + adjoint)', steps / elapsed,
                     'gradients/s', steps, warmup, elapsed, 'f64',
                     {'workload': '2x1x1 block with a perturbed spherical void (r=0.3, Fourier amplitude 0.02), '
                                  '%d C3D10, NeoHookean mu=1 kappa=100, pressure 0.05 in the void, x=0 fixed, '
                                  'gradient w.r.t. all %d nodal coordinates' % (E, 3 * p.n_nodes),
                      'preconditioner': args.adjoint_preconditioner})
    line.update({'compliance': J0, 'fd_check_rel_err_max': max(errs), 'fd_check_rel_err_median': float(np.median(errs)),
                 'fd_check_err_over_grad_norm_max': max(nerrs), 'fd_directions': len(errs), 'fd_h': h,
                 'roofline': roofline('k_elem_adjoint<10> (fea_elem_adjoint + nodal gather)', flops / t_vjp / 1e12,
                                      FP64_PEAK_TFLOPS, 'TFLOP/s', 'mfma', alg_flops_per_launch=flops,
                                      avg_launch_us=t_vjp * 1e6,
                                      note='FP64 vector arithmetic; peak = MI355X FP64 78.6 TFLOP/s; algorithmic '
                                           'FLOPs = the C3D10 tangent contraction 2 g (99 nen + 27 nen^2) per element'),
                 'cpu_baseline': cpu})
    return line


def cpu_baseline_adjoint(mesh, p, E, frac=8):
    import torch
    from oracle import autodiff as ad
    conn = mesh['conn'][: E // frac]
    X = mesh['nodes']
    U = 1e-3 * np.random.default_rng(0).standard_normal(X.shape)
    lam = np.random.default_rng(1).standard_normal(X.shape)
    t0 = time.perf_counter()
    ad.internal_force_vjp('C3D10', conn, X, U, (0.5, 0.0, 100.0), lam)
    t = (time.perf_counter() - t0) * frac
    return {'value': 1.0 / t, 'unit': 'element-VJP passes/s', 'cores': len(os.sched_getaffinity(0)),
            'kind': 'port', 'sample': 'torch-autograd oracle VJP of the internal force (CPU, FP64) over 1/%d of '
                                      'the C3D10 elements, scaled: %.2f s per full pass' % (frac, t)}


# ============================================================================ kernels (counter runs)
def bench_kernels(args, torch):
    """A short fixed sequence for hardware-counter runs: the beam's tangent
    assembly (FP64 element kernel + CSR gather to FP32), 20 launches of the PCG
    SpMV stage on the FP32 tangent, 20 level-0 Chebyshev steps of the multigrid
    V-cycle, and 20 plain device copies of 256 MiB (the known-size calibration for
    FETCH_SIZE / WRITE_SIZE)."""
    from fea_amd import workloads
    from fea_amd._native import F32, F64, call
    from fea_amd import _native as nat
    fe, mesh = workloads.beam_traction(args.nx, args.ny, args.nz)
    fe.assembly.initialize()
    p = fe.assembly.plan
    U = torch.zeros(3 * p.n_nodes, dtype=torch.float64, device='cuda')
    for _ in range(3):
        p.tangent(U, 1.0, F64, val_dtype=F32)
    val = p.csr_values(F32)
    x = torch.rand(p.n_free, dtype=torch.float64, device='cuda')
    q = torch.empty_like(x)
    ws = torch.zeros(int(nat.lib().fea_pcg_ws_size(p.n_free)), dtype=torch.float64, device='cuda')
    s = nat.stream_ptr()
    for _ in range(20):
        call('fea_pcg_spmv', F32, F64, p.n_free, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, val, x, q, ws,
             s)
    binv = torch.zeros(6 * abs(p.spmv_groups), dtype=torch.float32, device='cuda')
    nat.ops().block_jacobi_inv(F32, p.spmv_groups, p.grp, p.rowptr, p.col, p.diag_slot, val, binv)
    xb, xd = torch.rand_like(x), torch.rand_like(x)
    for _ in range(20):
        call('fea_amg_smooth0', F32, p.n_free, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, val, binv, x, xb, q,
             xd, 0.5, 0.25, s)
    v16 = torch.empty(4 * p.nnz // 3, dtype=torch.int16, device='cuda')
    sc = torch.zeros(2, dtype=torch.float64, device='cuda')
    call('fea_amg_half_values', F32, p.nnz, val, v16, sc, s)
    for _ in range(20):
        call('fea_amg_smooth0_half', p.n_free, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, v16, sc, binv, x,
             xb, q, xd, 0.5, 0.25, s)
    a = torch.empty(256 << 20, dtype=torch.uint8, device='cuda').fill_(1)
    b = torch.empty_like(a)
    for _ in range(20):
        b.copy_(a)
    torch.cuda.synchronize()
    return {'workload': 'kernels', 'spmv_alg_bytes': spmv_bytes(p, 4), 'copy_bytes': a.numel(),
            'n_elem': int(mesh['conn'].shape[0]), 'nnz': p.nnz, 'n_free': p.n_free, 'n_kstage': p.n_kstage}


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on exactly one GPU: this build has no multi-GPU path '
              '(requested --gpus %d)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    torch.set_num_threads(len(os.sched_getaffinity(0)))
    global _SERVING_EAGER, _SERVING_BUILDERS
    _SERVING_EAGER = bool(args.serving_eager)
    _SERVING_BUILDERS = int(args.serving_amg_builders)
    fn = {'beam': bench_beam, 'cube': bench_cube, 'explicit': bench_explicit, 'adjoint': bench_adjoint,
          'serving': bench_serving, 'kernels': bench_kernels, 'couple': bench_couple,
          'contact': bench_contact}[args.workload]
    line = fn(args, torch)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
