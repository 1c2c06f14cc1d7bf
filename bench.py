#!/usr/bin/env python
"""Benchmark: mixed-precision static Newton solves/s on one MI355X.

Workload default: cantilever beam 4x1x1 meshed as 128x32x32 C3D8 (131072
elements, 8 Gauss points each, 418176 free DOF), Mooney-Rivlin C10=0.5 C01=0.1
D1=0.01, clamped at x=0, uniform dead traction (0,0,-0.01) on x=4 ramped over 10
increments; Newton to ||R|| <= 1e-10 ||f_ext|| with FP32 element tangents and
FP32 Jacobi-PCG inside an FP64 refinement loop (FP64 residual / energy / J terms).

One "step" = one full static solve (10 increments) from u = 0 with the model
resident in HBM.  The line carries the roofline of the dominant kernel (the
PCG SpMV stage, timed live with HIP events on the solver's stream) and the CPU
baseline (the FP64 numpy oracle timed on a bounded sample of the same workload).
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--nx', type=int, default=128)
    ap.add_argument('--ny', type=int, default=32)
    ap.add_argument('--nz', type=int, default=32)
    ap.add_argument('--increments', type=int, default=10)
    ap.add_argument('--precision', default='mixed', choices=['mixed', 'mixed-f32', 'mixed-ir', 'fp64'])
    ap.add_argument('--rel-tol', type=float, default=1e-10)
    ap.add_argument('--cg-tol', type=float, default=1e-8)
    ap.add_argument('--no-fp64-compare', action='store_true')
    ap.add_argument('--forcing', default='ew', choices=['ew', 'fixed'])
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--spmv-reps', type=int, default=200)
    ap.add_argument('--max-newton', type=int, default=400, help='Newton iterations per solve (all increments) before failing')
    return ap.parse_args()


_T0 = [time.perf_counter()]


def _progress(k, st):
    print('#   t=%.2fs increment %d newton %d cg %d alpha %.3g |R| %.3e' % (
        time.perf_counter() - _T0[0], k, st['newton_iters'], st['cg_iters'], st['alpha'], st['residual']),
        file=sys.stderr, flush=True)


def make_solver(precision, args):
    from fea_amd import solver
    return solver.StaticImplicitSolver(maximum_iteration=args.max_newton, precision=precision,
                                       rel_tol=args.rel_tol, load_steps=args.increments,
                                       cg_tol=args.cg_tol, cg_max_iter=20000, armijo_slack=1e-12,
                                       progress=_progress, forcing=args.forcing)


def run_solve(fe, torch):
    GC0 = torch.zeros_like(fe.assembly.GC)
    t0 = time.perf_counter()
    _T0[0] = t0
    GC = fe.solver.solve(GC0=GC0)
    torch.cuda.synchronize()
    if GC is False:
        raise RuntimeError('Newton solve failed to converge')
    return GC, time.perf_counter() - t0


def spmv_roofline(fe, torch, reps):
    """Average duration of the PCG SpMV stage (fea_pcg_spmv: FP32 matrix, FP64 vectors) via HIP events.

    Compulsory bytes of one launch: FP32 values (4 B per nonzero), one int32
    column index per entry of each node's shared column list (the rows of a
    node reuse it), the FP64 input vector read once and the FP64 output written
    once (8 B per row each), and the row pointers of the group heads plus the
    group table (4 B each)."""
    from fea_amd import _native as nat
    from fea_amd._native import F32, F64, call
    p = fe.assembly.plan
    val = p.csr_values(F32)
    x = torch.rand(p.n_free, dtype=torch.float64, device='cuda')
    q = torch.empty_like(x)
    ws = torch.zeros(int(nat.lib().fea_pcg_ws_size(p.n_free)), dtype=torch.float64, device='cuda')
    s = nat.stream_ptr()
    for _ in range(5):
        call('fea_pcg_spmv', F32, F64, p.n_free, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.col, val, x, q, ws, s)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        call('fea_pcg_spmv', F32, F64, p.n_free, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.col, val, x, q, ws, s)
    e1.record(stream)
    e1.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    n, nnz = p.n_free, p.nnz
    lens = np.diff(p.rowptr_np)
    grp = p.grp.cpu().numpy()
    shared_cols = int(lens[grp[:-1]].sum())          # one column list per node group
    alg_bytes = nnz * 4 + shared_cols * 4 + n * (8 + 8) + (p.n_groups + 1) * 4 * 2
    return t, alg_bytes


def element_rate(fe, torch, reps=10, kernel_dtype=None):
    """Tangent assembly as the mixed solver runs it (FP64 element kernel + residual,
    CSR gather rounded to FP32), timed with events on the solver's stream."""
    from fea_amd._native import F32, F64
    p = fe.assembly.plan
    kd = F64 if kernel_dtype is None else kernel_dtype
    U = torch.zeros(3 * p.n_nodes, dtype=torch.float64, device='cuda')
    p.tangent(U, 1.0, kd, val_dtype=F32)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        p.tangent(U, 1.0, kd, val_dtype=F32)
    e1.record(stream)
    e1.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    n_elem = sum(g.n for g in p.groups if g.kind == 'elem')
    return n_elem / t, t


def cpu_baseline(args, plan, stats):
    """FP64 numpy oracle on a bounded sample, extrapolated to one full solve."""
    import scipy.sparse as sp
    from fea_amd import meshgen
    from oracle import assembly as oasm
    from oracle import materials as omat
    from oracle import solvers as osol
    threads = len(os.sched_getaffinity(0))
    # (1) tangent assembly of a 1/16 slice (same element size / material)
    nzs = max(1, args.nz // 16)
    mesh = meshgen.hex_block(args.nx, args.ny, nzs, size=(4.0, 1.0, nzs / args.nz))
    N = mesh['nodes'].shape[0]
    fixed = np.zeros((N, 3), dtype=bool)
    fixed[mesh['node_sets']['x0']] = True
    om = {'nodes': mesh['nodes'], 'fixed': fixed, 'pressures': [], 'tractions': [],
          'groups': [{'etype': 'C3D8', 'conn': mesh['conn'], 'material': omat.MooneyRivlin(0.5, 0.1, 0.01)}]}
    U = 1e-3 * np.random.default_rng(0).standard_normal(mesh['nodes'].shape)
    oasm.assemble_reduced(om, U)                           # warm (Gauss precompute)
    t0 = time.perf_counter()
    oasm.assemble_reduced(om, U)
    t_asm = (time.perf_counter() - t0) * (args.nz / nzs)
    t0 = time.perf_counter()
    oasm.total_energy(om, U)
    t_energy = (time.perf_counter() - t0) * (args.nz / nzs)
    # (2) 20 CG iterations of the oracle CG on a matrix with the full CSR structure
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


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on exactly one GPU: this build has no multi-GPU path '
              '(requested --gpus %d)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from fea_amd import workloads
    from fea_amd._native import F64

    fe, mesh = workloads.beam_traction(args.nx, args.ny, args.nz)
    fe.solver = make_solver(args.precision, args)
    fe.initialize()
    print('# model: %d elements, %d nodes, %d free DOF, nnz %d' % (
        mesh['conn'].shape[0], mesh['nodes'].shape[0], fe.assembly.plan.n_free, fe.assembly.plan.nnz),
        file=sys.stderr, flush=True)
    for i in range(args.warmup):
        _, t = run_solve(fe, torch)
        print('# warmup solve %d: %.2f s, stats %s' % (i, t, {k: v for k, v in fe.solver.stats.items()
                                                            if k != 'residual_history'}),
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    GC = None
    for i in range(args.steps):
        GC, t = run_solve(fe, torch)
        lat.append(t)
        print('# timed solve %d: %.3f s' % (i, t), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    stats = dict(fe.solver.stats)
    value = args.steps / elapsed

    fp64 = {}
    if args.precision != 'fp64' and not args.no_fp64_compare:
        fe64, _ = workloads.beam_traction(args.nx, args.ny, args.nz)
        fe64.solver = make_solver('fp64', args)
        fe64.initialize()
        GC64, t64 = run_solve(fe64, torch)
        err = float((GC - GC64).abs().max() / GC64.abs().max())
        fp64 = {'fp64_solves_per_s': 1.0 / t64, 'fp64_solve_s': t64, 'rel_err_vs_fp64': err,
                'fp64_cg_iters': fe64.solver.stats['cg_iters'], 'fp64_newton_iters': fe64.solver.stats['newton_iters']}
        print('# fp64 solve %.3f s, rel err %.2e' % (t64, err), file=sys.stderr, flush=True)
        del fe64

    t_spmv, alg_bytes = spmv_roofline(fe, torch, args.spmv_reps)
    achieved = alg_bytes / t_spmv / 1e9
    elem_rate, t_tan = element_rate(fe, torch)
    from fea_amd._native import F32 as _F32
    elem_rate32, t_tan32 = element_rate(fe, torch, kernel_dtype=_F32)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, fe.assembly.plan, stats)

    line = {
        'metric': 'static Newton solves/s (mixed FP32/FP64-IR) at 1e-8 rel err',
        'value': value, 'unit': 'solves/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * elapsed / args.steps, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64' if args.precision == 'fp64' else 'f32+f64',
        'data': 'synthetic (structured mesh generated in-process)',
        'config': {'workload': 'cantilever beam 4x1x1, %dx%dx%d C3D8, Mooney-Rivlin C10=0.5 C01=0.1 D1=0.01, '
                               'clamped x=0, dead traction (0,0,-0.01) on x=4, %d increments, Newton rel tol %g'
                               % (args.nx, args.ny, args.nz, args.increments, args.rel_tol),
                   'precision': args.precision, 'free_dof': fe.assembly.plan.n_free, 'nnz': fe.assembly.plan.nnz},
        'latency_p50_s': float(np.percentile(lat, 50)), 'latency_p99_s': float(np.percentile(lat, 99)),
        'newton_iters_per_solve': stats.get('newton_iters'), 'cg_iters_per_solve': stats.get('cg_iters'),
        'energy_evals_per_solve': stats.get('energy_evals'),
        'elem_tangent_evals_per_s': elem_rate, 'tangent_assembly_ms': 1e3 * t_tan,
        'elem_tangent_evals_per_s_f32_contraction': elem_rate32,
        'roofline': {'bound': 'hbm', 'kernel': 'k_pcg_spmv<float,double,32,true> (fea_pcg_spmv)',
                     'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': None,
                     'alg_bytes_per_launch': alg_bytes, 'avg_launch_us': t_spmv * 1e6},
        'cpu_baseline': cpu,
    }
    line.update(fp64)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
