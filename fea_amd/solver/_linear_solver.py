"""Linear solvers on the device CSR: Jacobi-preconditioned CG and its FP32/FP64
iterative-refinement variant.

``conjugate_gradient`` keeps the reference signature (_linear_solver.py:4-42).
``PCG`` is the device-resident driver used by the solvers: the HIP kernels keep
every CG scalar on the GPU and the host only checks a ``done`` flag once per
chunk of iterations.
"""
import ctypes

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
    ``precond='block'`` replaces the point Jacobi M = |diag K| by the node-block
    Jacobi M = blockdiag(K_gg) over the row groups (an addition: other iterates,
    the same solution; needs ``grp``, else point Jacobi is used).
    ``spmv_col``: the SpMV's column stream, required for a triple-structured group
    table (``n_groups < 0``: the compact per-triple columns, ``DevicePlan.spmv_col``).
    ``precond='amg'``: the smoothed-aggregation V-cycle of ``amg`` (solver.amg.AMG, built
    for the same pattern; float64 vectors), with the node-block Jacobi of the current
    matrix on its finest level.
    """

    def __init__(self, n, nnz, rowptr, col, diag_slot, mat_dtype, device, vec_dtype=None, grp=None,
                 n_groups=0, chunk=32, precond='jacobi', spmv_col=None, amg=None):
        if precond not in ('jacobi', 'block', 'amg'):
            raise ValueError("precond must be 'jacobi', 'block' or 'amg'")
        self.amg = amg if precond == 'amg' else None
        if precond == 'amg':
            if amg is None or grp is None or int(n_groups) == 0:
                raise ValueError("precond='amg' needs the hierarchy (solver.amg.AMG) and the node row groups")
            if (mat_dtype if vec_dtype is None else vec_dtype) != F64:
                raise ValueError("precond='amg' runs with float64 vectors")
        self.block = precond in ('block', 'amg') and grp is not None and int(n_groups) != 0
        self.n, self.nnz, self.chunk = int(n), int(nnz), int(chunk)
        self.mat_dtype = mat_dtype
        self.vec_dtype = mat_dtype if vec_dtype is None else vec_dtype
        self.rowptr, self.col, self.diag_slot = rowptr, col, diag_slot
        if int(n_groups) < 0 and spmv_col is None:
            raise ValueError('a triple-structured group table needs spmv_col (the compact triple columns)')
        self.spmv_col = col if int(n_groups) >= 0 else spmv_col
        self.grp, self.n_groups = grp, int(n_groups)
        T = _DT[self.vec_dtype]
        mk = lambda: torch.zeros(max(1, self.n), dtype=T, device=device)[:self.n]
        self.x, self.r, self.z, self.p, self.q, self.b = mk(), mk(), mk(), mk(), mk(), mk()
        if self.block:
            self.dinv = torch.zeros(6 * abs(self.n_groups), dtype=torch.float32, device=device)
        else:
            self.dinv = mk()
        self.ws = torch.zeros(int(nat.lib().fea_pcg_ws_size(self.n)), dtype=torch.float64, device=device)
        self.status = np.zeros(4, dtype=np.float64)
        self._status_t = torch.from_numpy(self.status)          # host status, shared with the numpy view
        self.iterations = 0

    @property
    def dtype(self):
        return self.vec_dtype

    def set_matrix(self, val):
        self.val = val
        if self.block:
            nat.ops().block_jacobi_inv(self.mat_dtype, self.n_groups, self.grp, self.rowptr, self.col,
                                       self.diag_slot, val, self.dinv)
        else:
            nat.ops().jacobi_inv(self.mat_dtype, self.vec_dtype, self.diag_slot, val, self.dinv)
        if self.amg is not None:
            # the V-cycle's finest level smooths with this matrix (its binary16 copy):
            # lambda_max of D^-1 A for it, and the copy refreshed
            self.amg.estimate_rho(self.mat_dtype, val, self.dinv)

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
        if self.amg is not None:
            return self._solve_amg(b, tol, max_iter)
        pc = 1 if self.block else 0
        O = nat.ops()
        O.pcg_start(self.mat_dtype, self.vec_dtype, self.n_groups, self.grp, self.rowptr, self.spmv_col, self.val,
                    pc, self.dinv, b, self.x, self.r, self.z, self.p, self.ws)
        done_iters = 0
        while True:
            step = min(self.chunk, max(1, max_iter - done_iters))
            O.pcg_iterate(self.mat_dtype, self.vec_dtype, self.n_groups, self.grp, self.rowptr, self.spmv_col,
                          self.val, pc, self.dinv, self.x, self.r, self.z, self.p, self.q, self.ws, float(tol),
                          int(max_iter), int(step), self._status_t)
            done_iters += step
            if self.status[1] != 0.0 or done_iters >= max_iter:
                break
        self.iterations = int(self.status[0])
        return self.x

    def _solve_amg(self, b, tol, max_iter):
        A = self.amg
        A.pcg_start(self.mat_dtype, self.val, self.dinv, b, self.x, self.r, self.z, self.p, self.ws)
        done_iters = 0
        while True:
            step = min(self.chunk, max(1, max_iter - done_iters))
            A.pcg_iterate(self.mat_dtype, self.val, self.dinv, self.x, self.r, self.z, self.p, self.q, self.ws,
                          float(tol), int(max_iter), int(step), self.status)
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


class PcgSystem(ctypes.Structure):
    """Mirror of ``fea_pcg_system`` (include/fea_amd.h)."""
    _fields_ = [('n', ctypes.c_int32), ('n_groups', ctypes.c_int32), ('spmv_blocks', ctypes.c_int32),
                ('vec_blocks', ctypes.c_int32), ('max_iter', ctypes.c_int32), ('precond', ctypes.c_int32),
                ('nnz', ctypes.c_int64), ('tol', ctypes.c_double),
                ('grp', ctypes.c_void_p), ('rowptr', ctypes.c_void_p), ('col', ctypes.c_void_p),
                ('diag_slot', ctypes.c_void_p), ('val', ctypes.c_void_p), ('dinv', ctypes.c_void_p),
                ('b', ctypes.c_void_p), ('x', ctypes.c_void_p), ('r', ctypes.c_void_p), ('z', ctypes.c_void_p),
                ('p', ctypes.c_void_p), ('q', ctypes.c_void_p), ('ws', ctypes.c_void_p)]


class PcgVectors:
    """Device vectors and workspace of one system of a batched PCG."""

    def __init__(self, n, vec_dtype, device, n_groups=0):
        T = _DT[vec_dtype]
        mk = lambda: torch.zeros(max(1, n), dtype=T, device=device)[:n]
        self.n = int(n)
        self.x, self.r, self.z, self.p, self.q, self.b = mk(), mk(), mk(), mk(), mk(), mk()
        # the preconditioner: n diagonal inverses, or 6 floats per row group (node-block)
        nb = max(self.n * T.itemsize, 24 * abs(int(n_groups)))
        self.dinv = torch.zeros(max(1, (nb + 7) // 8), dtype=torch.float64, device=device)
        self.block_ok = abs(int(n_groups)) > 0
        self.ws = torch.zeros(int(nat.lib().fea_pcg_ws_size(self.n)), dtype=torch.float64, device=device)

    def nbytes(self):
        return sum(t.numel() * t.element_size() for t in (self.x, self.r, self.z, self.p, self.q, self.dinv,
                                                         self.b, self.ws))


class BatchedPCG:
    """Jacobi-PCG over a batch of independent systems, one launch per CG stage.

    ``solve(entries)`` takes a list of ``(plan, val, vectors, tol, max_iter)``:
    ``plan`` a DevicePlan (its CSR, node groups and diagonal slots), ``val``
    its matrix values (``mat_dtype``), ``vectors`` a :class:`PcgVectors` whose
    ``b`` holds the right-hand side and ``x`` the warm start.  Every system is
    iterated until its own test r.z / b.(M^-1 b) < tol (the reference's
    criterion, _linear_solver.py:37) or max_iter; on return each ``x`` holds
    the solution.  The host reads one status array per chunk of ``chunk``
    iterations for the whole batch.  With ``use_graph`` the chunk (3 launches
    per iteration plus the status kernel, fixed grids that loop over the batch's
    virtual workgroups, counts read from a device header) is captured once in a HIP
    graph and replayed for every batch layout.  ``precond='block'``: node-block Jacobi for every
    system whose vectors were made with its row-group count (``PcgVectors(...,
    n_groups=)``), point Jacobi for the others.
    """

    def __init__(self, mat_dtype, vec_dtype, device, chunk=32, use_graph=False, precond='jacobi'):
        if precond not in ('jacobi', 'block'):
            raise ValueError("precond must be 'jacobi' or 'block'")
        self.precond = precond
        self.mat_dtype, self.vec_dtype = mat_dtype, vec_dtype
        self.device = torch.device(device)
        self.chunk = int(chunk)
        self.use_graph = bool(use_graph)
        L = nat.lib()
        if L.fea_pcg_system_size() != ctypes.sizeof(PcgSystem):
            raise nat.FeaNativeError('fea_pcg_system layout mismatch (%d vs %d bytes)'
                                     % (L.fea_pcg_system_size(), ctypes.sizeof(PcgSystem)))
        self.max_sys = 1024
        nb = self.max_sys * ctypes.sizeof(PcgSystem)
        self._sys_dev = torch.zeros(nb, dtype=torch.uint8, device=self.device)
        self._ss_dev = torch.zeros(self.max_sys + 1, dtype=torch.int32, device=self.device)
        self._vs_dev = torch.zeros(self.max_sys + 1, dtype=torch.int32, device=self.device)
        self._status_dev = torch.zeros(2 * self.max_sys, dtype=torch.float64, device=self.device)
        self._hdr_dev = torch.zeros(4, dtype=torch.int32, device=self.device)   # n, spmv / vec totals
        self._table = None             # the bytes of the table on the device
        self._graph = None
        self._graph_key = None
        # the graph path launches fixed grids that loop over a batch's virtual workgroups
        self.spmv_grid, self.vec_grid = 2048, 1024
        self.iterations = []
        self.chunks = 0

    def _entry(self, plan, val, v, tol, max_iter):
        """The system's table entry (bytes) and its SpMV / vector workgroup counts, cached
        on its vectors under every pointer and value the entry holds that can change."""
        key = (id(plan), plan.rowptr.data_ptr(), plan.spmv_col.data_ptr(), plan.grp.data_ptr(),
               plan.diag_slot.data_ptr(), val.data_ptr(), float(tol), int(max_iter), self.precond, self.mat_dtype)
        c = getattr(v, '_pcg_entry', None)
        if c is not None and c[0] == key:
            return c[1]
        if plan.grp is None or plan.n_free != v.n:
            raise ValueError('vectors do not match the plan')
        if val.dtype != _DT[self.mat_dtype] or v.x.dtype != _DT[self.vec_dtype]:
            raise ValueError('dtype mismatch')
        s = PcgSystem()
        s.n, s.n_groups, s.max_iter, s.nnz, s.tol = v.n, int(plan.spmv_groups), int(max_iter), plan.nnz, float(tol)
        s.grp, s.rowptr, s.col, s.diag_slot = (plan.grp.data_ptr(), plan.rowptr.data_ptr(), plan.spmv_col.data_ptr(),
                                               plan.diag_slot.data_ptr())
        s.val, s.dinv, s.b, s.x = val.data_ptr(), v.dinv.data_ptr(), v.b.data_ptr(), v.x.data_ptr()
        # node-block Jacobi needs the row groups and a preconditioner buffer sized for them
        s.precond = 1 if (self.precond == 'block' and getattr(v, 'block_ok', False)) else 0
        s.r, s.z, s.p, s.q, s.ws = v.r.data_ptr(), v.z.data_ptr(), v.p.data_ptr(), v.q.data_ptr(), v.ws.data_ptr()
        ss = np.zeros(2, dtype=np.int32)
        vs = np.zeros(2, dtype=np.int32)
        nat.check(nat.lib().fea_pcg_batch_layout(1, ctypes.byref(s), ss.ctypes.data, vs.ctypes.data),
                  'fea_pcg_batch_layout')
        out = (bytes(s), int(ss[1]), int(vs[1]))
        v._pcg_entry = (key, out)
        return out

    def _layout(self, entries):
        """Device table of the listed systems; uploaded only when it differs from the
        table already on the device (the kernels only read it)."""
        n = len(entries)
        if n > self.max_sys:
            raise ValueError('batch of %d systems exceeds %d' % (n, self.max_sys))
        try:
            parts = [self._entry(*e) for e in entries]
        except ValueError as err:
            raise ValueError('batch entry: %s' % err) from None
        ss = np.zeros(n + 1, dtype=np.int32)
        vs = np.zeros(n + 1, dtype=np.int32)
        ss[1:] = np.cumsum([p[1] for p in parts])
        vs[1:] = np.cumsum([p[2] for p in parts])
        if ss[-1] >= (1 << 30) or vs[-1] >= (1 << 30):
            raise ValueError('batch too large')
        blob = b''.join(p[0] for p in parts)
        if blob != self._table:
            # stream-ordered copies from pinned memory: the host does not wait for the
            # kernels queued before them, which still read the previous table
            raw = np.frombuffer(blob, dtype=np.uint8)
            self._sys_dev[:raw.size].copy_(torch.from_numpy(raw.copy()).pin_memory(), non_blocking=True)
            self._ss_dev[:n + 1].copy_(torch.from_numpy(ss).pin_memory(), non_blocking=True)
            self._vs_dev[:n + 1].copy_(torch.from_numpy(vs).pin_memory(), non_blocking=True)
            hdr = np.array([n, ss[-1], vs[-1], 0], dtype=np.int32)
            self._hdr_dev.copy_(torch.from_numpy(hdr).pin_memory(), non_blocking=True)
            self._table = blob
        return n, int(ss[-1]), int(vs[-1])

    def _iterate(self, n, st, vt, stream):
        call('fea_pcg_batch_iterate', self.mat_dtype, self.vec_dtype, n, self._sys_dev, self._ss_dev, self._vs_dev,
             st, vt, self.chunk, self._status_dev, stream)

    def start(self, entries):
        """Initialise the listed systems (r = b - A x, Jacobi inverse, z, p, counters)."""
        if not entries:
            return
        n, st, vt = self._layout(entries)
        call('fea_pcg_batch_start', self.mat_dtype, self.vec_dtype, n, self._sys_dev, self._ss_dev, self._vs_dev,
             st, vt, nat.stream_ptr())

    def _iterate_dev(self, stream):
        call('fea_pcg_batch_iterate_dev', self.mat_dtype, self.vec_dtype, self._hdr_dev, self._sys_dev, self._ss_dev,
             self._vs_dev, self.spmv_grid, self.vec_grid, self.chunk, self._status_dev, stream)

    def _chunk_graph(self):
        """The chunk captured once (fixed grids, counts read from the device header): one
        graph serves every batch layout; replays read the table uploaded before them."""
        key = ('dev', self.chunk, self.spmv_grid, self.vec_grid)
        if self._graph_key != key:
            self._graph = None
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize()
            with torch.cuda.graph(g):
                self._iterate_dev(nat.stream_ptr())
            torch.cuda.synchronize()
            self._graph, self._graph_key = g, key
        return self._graph

    def iterate(self, entries):
        """Enqueue one chunk of iterations of the listed (started) systems; returns the
        device status tensor [n, 2] = (iterations, done) written after the chunk (the
        next chunk overwrites it; systems already done are skipped by the chunk).  With
        ``use_graph`` the chunk is a replay of the one captured graph."""
        n, st, vt = self._layout(entries)
        if self.use_graph:
            self._chunk_graph().replay()
        else:
            self._iterate(n, st, vt, nat.stream_ptr())
        self.chunks += 1
        return self._status_dev[:2 * n].view(n, 2)

    def status(self, n):
        """The status [n, 2] the last chunk wrote for its n systems (a device view)."""
        return self._status_dev[:2 * n].view(n, 2)

    def solve(self, entries):
        """Start and iterate the listed systems to convergence (one host read per chunk)."""
        if not entries:
            return []
        n, st, vt = self._layout(entries)
        s = nat.stream_ptr()
        call('fea_pcg_batch_start', self.mat_dtype, self.vec_dtype, n, self._sys_dev, self._ss_dev, self._vs_dev,
             st, vt, s)
        max_iter = max(int(e[4]) for e in entries)
        graph = self._chunk_graph() if self.use_graph else None
        done_iters = 0
        status = None
        while done_iters < max_iter:
            if graph is not None:
                graph.replay()
            else:
                self._iterate(n, st, vt, s)
            self.chunks += 1
            done_iters += self.chunk
            status = self._status_dev[:2 * n].view(n, 2).cpu().numpy()
            if np.all(status[:, 1] != 0.0):
                break
        self.iterations = [int(i) for i in status[:, 0]]
        return self.iterations
