"""The PyTorch-ROCm operators (torch.ops.fea, TORCH_LIBRARY over the C ABI) against the
bare C ABI called through ctypes: the element tangent/residual of structural_Force and
its energy, the CSR and residual gathers of the assembly, and the PCG give bit-identical
results; and the operators refuse tensors of the wrong device, dtype, layout or size
before anything is launched."""
import numpy as np
import pytest
import torch

from fea_amd import _native as nat
from fea_amd import workloads
from fea_amd._native import F32, F64, call

pytestmark = pytest.mark.gpu


def _plan(etype='C3D8'):
    if etype == 'C3D8':
        fe, _ = workloads.beam_traction(8, 3, 3, traction=(0.0, 0.0, -0.01))
    else:
        fe, _ = workloads.cube_pressure(3, pressure=0.2, jitter=0.05, seed=2, quadratic=(etype == 'C3D10'))
    fe.assembly.initialize()
    p = fe.assembly.plan
    U = p.gc_to_u(torch.as_tensor(0.02 * np.random.default_rng(1).standard_normal(p.n_free), device='cuda'))
    return p, U


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
@pytest.mark.parametrize('dtype,val_dtype', [(F64, F64), (F64, F32), (F32, F32)])
def test_tangent_ops_match_ctypes(gpu, etype, dtype, val_dtype):
    p, U = _plan(etype)
    R_op, K_op = p.tangent(U, 1.0, dtype, val_dtype=val_dtype)
    R_op, K_op = R_op.clone(), K_op.clone()
    # the same entry points through ctypes, into fresh buffers
    s = nat.stream_ptr()
    ks = torch.zeros_like(p.kstage(dtype))
    rs = torch.zeros_like(p.rstage)
    for g in p.groups:                      # the calls DevicePlan.tangent makes, through ctypes
        if g.kind == 'elem':
            call('fea_elem_residual_tangent', g.ntype, dtype, g.n, g.conn, p.X, U, g.params, g.dst_k, ks, g.dst_r, rs, s)
        elif g.kind == 'pressure':
            if dtype == F64:
                call('fea_surface_pressure', g.ntype, dtype, g.n, g.conn, p.X, U, float(g.load.pressure), g.dst_k,
                     ks, g.dst_r, rs, s)
            else:                           # float32 stiffness, float64 residual
                call('fea_surface_pressure', g.ntype, dtype, g.n, g.conn, p.X, U, float(g.load.pressure), g.dst_k,
                     ks, None, None, s)
                call('fea_surface_pressure', g.ntype, F64, g.n, g.conn, p.X, U, float(g.load.pressure), None, None,
                     g.dst_r, rs, s)
        elif g.kind == 'traction':
            call('fea_surface_traction', g.ntype, g.n, g.conn, p.X, U, np.asarray(g.load.traction, dtype=np.float64),
                 1.0, g.dst_r, rs, None, s)
    K_c = torch.zeros_like(K_op)
    call('fea_gather_csr', dtype, val_dtype, p.n_blocks, p.blk_a, p.blk_b, p.blk_colpos, p.kseg, p.dof_map, p.rowptr,
         ks, K_c, s)
    R_c = torch.zeros_like(R_op)
    call('fea_gather_residual', p.n_nodes, p.rseg, p.dof_map, rs, R_c, s)
    torch.cuda.synchronize()
    assert torch.equal(K_op, K_c)
    assert torch.equal(R_op, R_c)
    # residual-only and energy operators
    R2 = p.residual(U, 1.0, dtype).clone()
    rs2 = torch.zeros_like(p.rstage)
    for g in p.groups:
        if g.kind == 'elem':
            call('fea_elem_residual', g.ntype, dtype, g.n, g.conn, p.X, U, g.params, g.dst_r, rs2, s)
        elif g.kind == 'pressure':
            call('fea_surface_pressure', g.ntype, dtype, g.n, g.conn, p.X, U, float(g.load.pressure), None, None,
                 g.dst_r, rs2, s)
        elif g.kind == 'traction':
            call('fea_surface_traction', g.ntype, g.n, g.conn, p.X, U, np.asarray(g.load.traction, dtype=np.float64),
                 1.0, g.dst_r, rs2, None, s)
    R2c = torch.zeros_like(R2)
    call('fea_gather_residual', p.n_nodes, p.rseg, p.dof_map, rs2, R2c, s)
    torch.cuda.synchronize()
    assert torch.equal(R2, R2c)
    g = [g for g in p.groups if g.kind == 'elem'][0]
    nb = nat.lib().fea_elem_energy_blocks(g.n)
    e_op = torch.zeros(nb, dtype=torch.float64, device='cuda')
    e_c = torch.zeros_like(e_op)
    nat.ops().elem_energy(g.ntype, g.conn, p.X, U, g.params.tolist(), e_op)
    call('fea_elem_energy', g.ntype, g.n, g.conn, p.X, U, g.params, e_c, s)
    torch.cuda.synchronize()
    assert torch.equal(e_op, e_c)


@pytest.mark.parametrize('precond', ['jacobi', 'block'])
@pytest.mark.parametrize('mat', [F64, F32])
def test_pcg_ops_match_ctypes(gpu, precond, mat):
    from fea_amd import solver
    p, U = _plan('C3D8')
    R, K = p.tangent(U, 1.0, F64, val_dtype=mat)
    b = (-R).contiguous()
    pcg = solver.PCG(p.n_free, p.nnz, p.rowptr, p.col, p.diag_slot, mat, p.device, vec_dtype=F64, grp=p.grp,
                     n_groups=p.spmv_groups, chunk=8, precond=precond, spmv_col=p.spmv_col)
    pcg.set_matrix(K)
    x_op = pcg.solve(b, None, tol=1e-18, max_iter=4000).clone()
    it_op = pcg.iterations
    # ctypes path on the same preconditioner and fresh vectors
    s = nat.stream_ptr()
    n = p.n_free
    x, r, z, pp, q = [torch.zeros(n, dtype=torch.float64, device='cuda') for _ in range(5)]
    ws = torch.zeros(int(nat.lib().fea_pcg_ws_size(n)), dtype=torch.float64, device='cuda')
    st = np.zeros(4)
    pc = 1 if precond == 'block' else 0
    call('fea_pcg_start', mat, F64, n, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, K, pc, pcg.dinv, b, x, r, z,
         pp, ws, s)
    done = 0
    while True:
        call('fea_pcg_iterate', mat, F64, n, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, K, pc, pcg.dinv, x, r,
             z, pp, q, ws, 1e-18, 4000, 8, st, s)
        done += 8
        if st[1] != 0.0 or done >= 4000:
            break
    assert int(st[0]) == it_op and 0 < it_op < 4000
    assert torch.equal(x, x_op)
    # the SpMV operator
    y_op = torch.empty_like(x)
    y_c = torch.empty_like(x)
    nat.ops().spmv(mat, F64, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, K, x, y_op)
    call('fea_spmv', mat, F64, n, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, K, x, y_c, s)
    torch.cuda.synchronize()
    assert torch.equal(y_op, y_c)


def test_operators_refuse_bad_tensors(gpu):
    p, U = _plan('C3D4')
    O = nat.ops()
    g = [g for g in p.groups if g.kind == 'elem'][0]
    prm = g.params.tolist()
    ks = p.kstage(F64)
    # wrong dtype, wrong device, non-contiguous, wrong size, wrong parameter count
    with pytest.raises(RuntimeError, match='dtype'):
        O.elem_residual_tangent(g.ntype, F64, g.conn, p.X, U.float(), prm, g.dst_k, ks, g.dst_r, p.rstage)
    with pytest.raises((RuntimeError, NotImplementedError)):
        O.elem_residual_tangent(g.ntype, F64, g.conn, p.X, U.cpu(), prm, g.dst_k, ks, g.dst_r, p.rstage)
    U2 = torch.zeros(2 * U.numel(), dtype=torch.float64, device='cuda')[::2]
    with pytest.raises(RuntimeError, match='contiguous'):
        O.elem_residual_tangent(g.ntype, F64, g.conn, p.X, U2, prm, g.dst_k, ks, g.dst_r, p.rstage)
    with pytest.raises(RuntimeError, match='dst_k'):
        O.elem_residual_tangent(g.ntype, F64, g.conn, p.X, U, prm, g.dst_k[:-1], ks, g.dst_r, p.rstage)
    with pytest.raises(RuntimeError, match='params'):
        O.elem_residual_tangent(g.ntype, F64, g.conn, p.X, U, prm[:2], g.dst_k, ks, g.dst_r, p.rstage)
    with pytest.raises(RuntimeError, match='elem_type'):
        O.elem_residual_tangent(5, F64, g.conn, p.X, U, prm, g.dst_k, ks, g.dst_r, p.rstage)
    x = torch.zeros(p.n_free + 1, dtype=torch.float64, device='cuda')
    val = p.csr_values(F64)
    with pytest.raises(RuntimeError, match='length'):
        O.spmv(F64, F64, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, val, x, x.clone())
    torch.cuda.synchronize()
