"""Linear solvers on the device CSR: Jacobi-preconditioned CG and its FP32/FP64
iterative-refinement variant.

``conjugate_gradient`` keeps the reference signature (_linear_solver.py:4-42).
``PCG`` is the device-resident driver used by the solvers: the HIP kernels keep
every CG scalar on the GPU and the host only checks a ``done`` flag once per
chunk of iterations.
"""
import numpy as np
import torch

from .. import _native as nat
from .._native import F32, F64, call, ptr

_DT = {F32: torch.float32, F64: torch.float64}


class PCG:
    """Jacobi-PCG on a fixed CSR structure (rowptr/col int32 on the device).

    With the Jacobi preconditioner M = |diag K| this is the reference's CG on
    the symmetrically scaled system D^-1/2 K D^-1/2 (static_implicit.py:314-321)
    with identical iterates; its stopping test r.z / b.(M^-1 b) < tol is the
    reference's rsnew / r_r0 < tol on the scaled system (_linear_solver.py:37).

    ``mat_dtype`` is the precision of the matrix values, ``vec_dtype`` that of
    the vectors and recurrences; ``grp`` (optional) the node row groups.
    """

    def __init__(self, n, nnz, rowptr, col, diag_slot, mat_dtype, device, vec_dtype=None, grp=None,
                 n_groups=0, chunk=32):
        self.n, self.nnz, self.chunk = int(n), int(nnz), int(chunk)
        self.mat_dtype = mat_dtype
        self.vec_dtype = mat_dtype if vec_dtype is None else vec_dtype
        self.rowptr, self.col, self.diag_slot = rowptr, col, diag_slot
        self.grp, self.n_groups = grp, int(n_groups)
        T = _DT[self.vec_dtype]
        mk = lambda: torch.zeros(max(1, self.n), dtype=T, device=device)[:self.n]
        self.x, self.r, self.z, self.p, self.q, self.dinv, self.b = mk(), mk(), mk(), mk(), mk(), mk(), mk()
        self.ws = torch.zeros(int(nat.lib().fea_pcg_ws_size(self.n)), dtype=torch.float64, device=device)
        self.status = np.zeros(4, dtype=np.float64)
        self.iterations = 0

    @property
    def dtype(self):
        return self.vec_dtype

    def set_matrix(self, val):
        self.val = val
        call('fea_jacobi_inv', self.mat_dtype, self.vec_dtype, self.n, self.diag_slot, val, self.dinv,
             nat.stream_ptr())

    def solve(self, b, x0=None, tol=1e-5, max_iter=1500):
        """Solve K x = b; returns x (vector precision)."""
        s = nat.stream_ptr()
        if b.dtype != _DT[self.vec_dtype]:
            self.b.copy_(b)
            b = self.b
        if x0 is None:
            self.x.zero_()
        else:
            self.x.copy_(x0)
        call('fea_pcg_start', self.mat_dtype, self.vec_dtype, self.n, self.nnz, self.n_groups, self.grp,
             self.rowptr, self.col, self.val, self.dinv, b, self.x, self.r, self.z, self.p, self.ws, s)
        done_iters = 0
        while True:
            step = min(self.chunk, max(1, max_iter - done_iters))
            call('fea_pcg_iterate', self.mat_dtype, self.vec_dtype, self.n, self.nnz, self.n_groups, self.grp,
                 self.rowptr, self.col, self.val, self.dinv, self.x, self.r, self.z, self.p, self.q, self.ws,
                 float(tol), int(max_iter), int(step), self.status, s)
            done_iters += step
            if self.status[1] != 0.0 or done_iters >= max_iter:
                break
        self.iterations = int(self.status[0])
        return self.x


def conjugate_gradient(A_indices, A_values, b, x0=torch.zeros([0]), tol=1e-3, max_iter=1500):
    """Reference signature (_linear_solver.py:6-11): plain CG on a COO matrix.

    The COO input is coalesced into CSR once (torch sparse conversion is only
    the format change); iterations run in the device PCG kernels with an
    identity preconditioner, as the reference's unpreconditioned loop.
    """
    nat.require_gpu()
    dev = torch.device('cuda')
    n = b.shape[0]
    A = torch.sparse_coo_tensor(A_indices.to(dev), A_values.to(dev), [n, n]).coalesce().to_sparse_csr()
    dtype = F64 if A_values.dtype == torch.float64 else F32
    T = _DT[dtype]
    rowptr = A.crow_indices().to(torch.int32)
    col = A.col_indices().to(torch.int32)
    val = A.values().to(T).contiguous()
    pcg = PCG(n, val.numel(), rowptr, col, torch.full((n,), -1, dtype=torch.int32, device=dev), dtype, dev)
    pcg.val = val
    pcg.dinv.fill_(1.0)
    x0 = None if x0.numel() == 0 else x0.to(dev, T)
    x = pcg.solve(b.to(dev, T).contiguous(), x0, tol=tol, max_iter=max_iter)
    return x.clone()
