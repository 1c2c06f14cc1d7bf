"""fea_pcg_run_noalloc (the allocation-free PCG: workspace in the library's static device
array) against fea_pcg_start + fea_pcg_iterate with a caller workspace, through the C ABI:
the same kernels on the same inputs give bit-identical iterates and counts.  Two calls in a
row, the second on a smaller system, so the second reuses the array the first left behind."""
import numpy as np
import pytest
import torch

from fea_amd import _native as nat
from fea_amd import workloads
from fea_amd._native import F32, F64, call

pytestmark = pytest.mark.gpu

_DT = {F32: torch.float32, F64: torch.float64}


def _system(n_cells, seed, mat):
    fe, _ = workloads.cube_pressure(n_cells, pressure=0.2, jitter=0.05, seed=seed)
    fe.assembly.initialize()
    p = fe.assembly.plan
    U = p.gc_to_u(torch.as_tensor(0.01 * np.random.default_rng(seed).standard_normal(p.n_free), device='cuda'))
    R, K = p.tangent(U, 1.0, F64, val_dtype=mat)
    return p, (-R).contiguous(), K.clone()


def _dinv(p, K, mat, vec, precond):
    s = nat.stream_ptr()
    if precond == 1:
        d = torch.zeros(6 * abs(p.spmv_groups), dtype=torch.float32, device='cuda')
        call('fea_block_jacobi_inv', mat, p.n_free, p.spmv_groups, p.grp, p.rowptr, p.col, p.diag_slot, K, d, s)
    else:
        d = torch.zeros(p.n_free, dtype=_DT[vec], device='cuda')
        call('fea_jacobi_inv', mat, vec, p.n_free, p.diag_slot, K, d, s)
    return d


def _vectors(n, vec):
    return [torch.zeros(n, dtype=_DT[vec], device='cuda') for _ in range(5)]      # x r z p q


def _solve_ws(p, b, K, d, mat, vec, precond, tol, max_iter, chunk):
    x, r, z, pp, q = _vectors(p.n_free, vec)
    ws = torch.zeros(int(nat.lib().fea_pcg_ws_size(p.n_free)), dtype=torch.float64, device='cuda')
    st = np.zeros(4)
    s = nat.stream_ptr()
    call('fea_pcg_start', mat, vec, p.n_free, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, K, precond, d,
         b, x, r, z, pp, ws, s)
    done = 0
    while True:
        call('fea_pcg_iterate', mat, vec, p.n_free, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, K, precond,
             d, x, r, z, pp, q, ws, tol, max_iter, chunk, st, s)
        done += chunk
        if st[1] != 0.0 or done >= max_iter:
            return x, st.copy()


def _solve_noalloc(p, b, K, d, mat, vec, precond, tol, max_iter, chunk):
    x, r, z, pp, q = _vectors(p.n_free, vec)
    st = np.zeros(4)
    s = nat.stream_ptr()
    first, done = True, 0
    while True:
        call('fea_pcg_run_noalloc', mat, vec, p.n_free, p.nnz, p.spmv_groups, p.grp, p.rowptr, p.spmv_col, K,
             precond, d, b if first else None, x, r, z, pp, q, tol, max_iter, chunk, st, s)
        first = False
        done += chunk
        if st[1] != 0.0 or done >= max_iter:
            return x, st.copy()


@pytest.mark.parametrize('precond', [0, 1])
@pytest.mark.parametrize('pair', [(F64, F64), (F32, F64)])
def test_noalloc_pcg_matches_workspace_pcg(gpu, pair, precond):
    mat, vec = pair
    for n_cells, seed in ((4, 1), (2, 2)):                 # second call: a smaller system
        p, b, K = _system(n_cells, seed, mat)
        d = _dinv(p, K, mat, vec, precond)
        x_ws, st_ws = _solve_ws(p, b, K, d, mat, vec, precond, 1e-20, 3000, 16)
        x_na, st_na = _solve_noalloc(p, b, K, d, mat, vec, precond, 1e-20, 3000, 16)
        torch.cuda.synchronize()
        assert st_ws[1] == 1.0 and 0 < st_ws[0] < 3000, st_ws
        assert np.array_equal(st_ws, st_na), (st_ws, st_na)
        assert torch.equal(x_ws, x_na)
        # and the result solves the system
        A = torch.sparse_csr_tensor(p.rowptr.long(), p.col.long(), K.double(), (p.n_free, p.n_free)).to_dense()
        res = torch.linalg.vector_norm(A @ x_na.double() - b.double()) / torch.linalg.vector_norm(b.double())
        assert float(res) < 1e-8


def test_noalloc_pcg_fixed_iteration_counts(gpu):
    """A solve stopped after k iterations (max_iter = k) agrees bit for bit as well."""
    p, b, K = _system(3, 5, F64)
    d = _dinv(p, K, F64, F64, 1)
    for k in (1, 7, 40):
        x_ws, st_ws = _solve_ws(p, b, K, d, F64, F64, 1, 0.0, k, 8)
        x_na, st_na = _solve_noalloc(p, b, K, d, F64, F64, 1, 0.0, k, 8)
        assert st_ws[0] == k and np.array_equal(st_ws, st_na)
        assert torch.equal(x_ws, x_na)


def test_noalloc_pcg_rejects_bad_arguments(gpu):
    with pytest.raises(nat.FeaNativeError):
        call('fea_pcg_run_noalloc', F64, F64, 0, 0, 0, None, None, None, None, 0, None, None, None, None, None,
             None, None, 1e-10, 10, 1, None, nat.stream_ptr())
