"""Smoothed-aggregation multigrid preconditioner for the device PCG.

An addition to the reference, whose CG is scaled by the point-Jacobi diagonal only
(FEA/solver/static_implicit.py:314-321): the same linear systems and the same Newton
solutions, with one to two orders of magnitude fewer CG iterations on the
near-incompressible brick beams.

The hierarchy is built once per model on the host (``fea_amg_build``,
csrc/amg_setup.cpp) from a tangent of the model -- node aggregates, the rigid-body
modes as the near-null space, one damped block-Jacobi smoothing step of the
prolongator, Galerkin coarse operators in 6x6 blocks and a dense inverse on the
coarsest level -- and kept on the device.  Every preconditioner application is a
V-cycle whose finest level uses the *current* matrix and its node-block Jacobi
inverses, so the coarse levels may lag the tangent while each CG solve still has a
fixed symmetric positive definite preconditioner.
"""
import ctypes
import os

import numpy as np
import torch

from .. import _native as nat
from .._native import F32, F64, check

_P = ctypes.c_void_p


class AmgLevel(ctypes.Structure):
    """Mirror of ``fea_amg_level`` (include/fea_amd.h)."""
    _fields_ = [('n', ctypes.c_int32), ('nagg', ctypes.c_int32), ('rho', ctypes.c_double),
                ('rowptr', _P), ('col', _P), ('val', _P), ('dinv', _P), ('prow', _P), ('pagg', _P), ('pval', _P),
                ('pt_ptr', _P), ('pt_row', _P), ('ptval', _P), ('dense_inv', _P),
                ('b', _P), ('x', _P), ('x2', _P), ('d', _P), ('r', _P), ('val16', _P), ('val16_scale', _P)]




def _lib():
    return nat.lib()


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class AMG:
    """The hierarchy of one model's free-DOF CSR, resident on the device.

    ``plan`` is the DevicePlan (its CSR pattern, node row groups and coordinates) and
    ``K`` a float64 tangent in that pattern (values on the device) to build from.
    ``degree``: Chebyshev smoothing degree; ``max_coarse``: largest coarsest level
    (dense inverse).
    """

    def __init__(self, plan, K, degree=2, max_coarse=2000, max_levels=8, threads=None, handle=None,
                 half_level0=False):
        import time
        t0 = time.perf_counter()
        L = _lib()
        if L.fea_amg_level_size() != ctypes.sizeof(AmgLevel):
            raise nat.FeaNativeError('fea_amg_level layout mismatch (%d vs %d bytes)'
                                     % (L.fea_amg_level_size(), ctypes.sizeof(AmgLevel)))
        if plan.grp is None or plan.n_groups == 0:
            raise ValueError('AMG needs the node row groups of the plan')
        self.plan = plan
        self.degree = int(degree)
        self.device = plan.device
        n = plan.n_free
        self.n = n
        if handle is None:
            h = HostBuild(plan, K, max_coarse, max_levels, threads).run()
        else:
            h = handle          # built on the host already (HostBuild.run, e.g. on another thread)
        try:
            self._export(L, h)
        finally:
            L.fea_amg_destroy(h)
        dev = self.device
        # power iteration work: a fixed pseudo-random unit start vector
        x0 = np.random.default_rng(12345).standard_normal(n)
        x0 /= np.linalg.norm(x0)
        self._px = torch.as_tensor(x0.astype(np.float32), device=dev)
        self._pwork = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(2)]
        self._ppart = torch.zeros(2048, dtype=torch.float64, device=dev)
        # half_level0: the level-0 smoothing and residual read a binary16 copy of the current
        # matrix (triple-structured matrices; two thirds of the float values' bytes, a third of
        # their loads; the beam's PCG iteration counts do not change).  Off by default: on the
        # beam those passes get 8-13 % faster but the PCG product, which keeps the float matrix,
        # 26 % slower -- the two copies no longer both fit the 256 MB Infinity Cache -- and the
        # solve time is unchanged
        self.half_level0 = bool(half_level0) and plan.spmv_groups < 0 and plan.nnz % 3 == 0
        if self.half_level0:
            self._val16 = torch.empty(4 * plan.nnz // 3, dtype=torch.int16, device=dev)
            self._val16_scale = torch.zeros(2, dtype=torch.float64, device=dev)
            self._arr[0].val16 = _ptr(self._val16)
            self._arr[0].val16_scale = _ptr(self._val16_scale)
        self.rho0 = None
        self.setup_time = time.perf_counter() - t0

    def _export(self, L, h):
        dev = self.device
        nl = L.fea_amg_num_levels(h)
        self.levels = []
        self._keep = []
        arr = (AmgLevel * nl)()
        sizes = np.zeros(6, dtype=np.int64)
        for l in range(nl):
            check(L.fea_amg_level_sizes(h, l, _np_ptr(sizes)), 'fea_amg_level_sizes')
            n, nagg, nblk, nnz, ngrp, nd = (int(v) for v in sizes)
            last = l == nl - 1
            host = {}
            if not last:
                host['prow'] = np.zeros(n + 1, np.int32)
                host['pagg'] = np.zeros(nblk, np.int32)
                host['pval'] = np.zeros(6 * nblk, np.float32)
                host['pt_ptr'] = np.zeros(nagg + 1, np.int32)
                host['pt_blk'] = np.zeros(nblk, np.int32)
                host['pt_row'] = np.zeros(nblk, np.int32)
            if l > 0 and not last:
                host['rowptr'] = np.zeros(n + 1, np.int32)
                host['col'] = np.zeros(nnz, np.int32)
                host['val'] = np.zeros(nnz, np.float64)
                host['dinv'] = np.zeros(36 * ngrp, np.float32)
            if last:
                host['dense'] = np.zeros(nd * nd, np.float32)
            rho = ctypes.c_double(0.0)
            g = host.get
            check(L.fea_amg_level_export(h, l, _np_ptr(g('prow')), _np_ptr(g('pagg')), _np_ptr(g('pval')),
                                         _np_ptr(g('pt_ptr')), _np_ptr(g('pt_blk')), _np_ptr(g('pt_row')),
                                         _np_ptr(g('rowptr')), _np_ptr(g('col')), _np_ptr(g('val')),
                                         _np_ptr(g('dinv')), ctypes.byref(rho), _np_ptr(g('dense'))),
                  'fea_amg_level_export')
            if not last:
                # prolongator values in binary16, and a copy in transposed-entry order for restriction
                p6 = host.pop('pval').reshape(-1, 6)
                host['pval16'] = p6.astype(np.float16).view(np.uint16).reshape(-1)
                host['ptval16'] = p6[host.pop('pt_blk')].astype(np.float16).view(np.uint16).reshape(-1)
            d = {k: torch.as_tensor(v.astype(np.float32) if k == 'val' else v.view(np.int16)
                                    if v.dtype == np.uint16 else v, device=dev) for k, v in host.items()}
            for k in ('b', 'x', 'x2', 'd', 'r'):
                d[k] = torch.zeros(max(1, n), dtype=torch.float64, device=dev)
            lv = arr[l]
            lv.n, lv.nagg, lv.rho = n, nagg if not last else 0, float(rho.value)
            lv.rowptr, lv.col, lv.val, lv.dinv = (_ptr(d.get(k)) for k in ('rowptr', 'col', 'val', 'dinv'))
            lv.prow, lv.pagg, lv.pval = _ptr(d.get('prow')), _ptr(d.get('pagg')), _ptr(d.get('pval16'))
            lv.pt_ptr, lv.pt_row, lv.ptval = _ptr(d.get('pt_ptr')), _ptr(d.get('pt_row')), _ptr(d.get('ptval16'))
            lv.dense_inv = _ptr(d.get('dense'))
            lv.b, lv.x, lv.x2, lv.d, lv.r = (_ptr(d[k]) for k in ('b', 'x', 'x2', 'd', 'r'))
            self.levels.append({'n': n, 'nagg': nagg, 'nnz': nnz, 'rho': float(rho.value), 'blocks': nblk})
            self._keep.append(d)
        self._arr = arr
        self.n_levels = nl
        self.operator_complexity = (self.plan.nnz + sum(v['nnz'] for v in self.levels[1:])) / self.plan.nnz

    def nbytes(self):
        return sum(t.numel() * t.element_size() for d in self._keep for t in d.values())

    def set_level0(self, mat_dtype, val):
        """Refresh the binary16 copy of the level-0 matrix (fea_amg_half_values) after the
        matrix values changed; without half_level0 nothing to do."""
        if self.half_level0:
            check(_lib().fea_amg_half_values(mat_dtype, self.plan.nnz, _ptr(val), _ptr(self._val16),
                                             _ptr(self._val16_scale), nat.stream_ptr()), 'fea_amg_half_values')

    def nbytes_level0(self):
        return self._val16.numel() * 2 if self.half_level0 else 0

    def estimate_rho(self, mat_dtype, val, binv, iters=15):
        """lambda_max of D^-1 A for the current level-0 matrix (power iteration on the device),
        and the level-0 binary16 copy refreshed (``set_level0``): called whenever the matrix
        values change, before the V-cycle uses them."""
        self.set_level0(mat_dtype, val)
        p = self.plan
        rho = ctypes.c_double(0.0)
        x, y = self._pwork
        x.copy_(self._px)
        check(_lib().fea_amg_power(mat_dtype, self.n, p.nnz, p.spmv_groups, _ptr(p.grp), _ptr(p.rowptr),
                                   _ptr(p.spmv_col), _ptr(val), _ptr(binv), _ptr(x), _ptr(y), _ptr(self._ppart),
                                   int(iters), ctypes.byref(rho), nat.stream_ptr()), 'fea_amg_power')
        self.rho0 = float(rho.value)
        return self.rho0

    def vcycle(self, mat_dtype, val, binv, b, z):
        """z = V(b) (float64 vectors, z not aliasing b) with the current level-0 matrix ``val``."""
        p = self.plan
        check(_lib().fea_amg_vcycle(mat_dtype, self.n, p.nnz, p.spmv_groups, _ptr(p.grp), _ptr(p.rowptr),
                                    _ptr(p.spmv_col), _ptr(val), _ptr(binv), float(self.rho0), self._arr,
                                    self.n_levels, self.degree, _ptr(b), _ptr(z), nat.stream_ptr()),
              'fea_amg_vcycle')

    def pcg_start(self, mat_dtype, val, binv, b, x, r, z, p_, ws):
        p = self.plan
        check(_lib().fea_pcg_start_amg(mat_dtype, self.n, p.nnz, p.spmv_groups, _ptr(p.grp), _ptr(p.rowptr),
                                       _ptr(p.spmv_col), _ptr(val), _ptr(binv), float(self.rho0), self._arr,
                                       self.n_levels, self.degree, _ptr(b), _ptr(x), _ptr(r), _ptr(z), _ptr(p_),
                                       _ptr(ws), nat.stream_ptr()),
              'fea_pcg_start_amg')

    def pcg_iterate(self, mat_dtype, val, binv, x, r, z, p_, q, ws, tol, max_iter, iters, status):
        p = self.plan
        check(_lib().fea_pcg_iterate_amg(mat_dtype, self.n, p.nnz, p.spmv_groups, _ptr(p.grp), _ptr(p.rowptr),
                                         _ptr(p.spmv_col), _ptr(val), _ptr(binv), float(self.rho0), self._arr,
                                         self.n_levels, self.degree, _ptr(x), _ptr(r), _ptr(z), _ptr(p_), _ptr(q),
                                         _ptr(ws), float(tol), int(max_iter),
                                         int(iters), _np_ptr(status), nat.stream_ptr()),
              'fea_pcg_iterate_amg')


class HostBuild:
    """The host part of the set-up (fea_amg_build) separated from the device export: the
    inputs are gathered on the calling thread (the tangent's values copied to the host),
    and ``run()`` -- one native call, which releases the interpreter lock -- may then run
    on another thread; its handle goes to ``AMG(plan, None, handle=...)`` on the thread
    that owns the device work."""

    def __init__(self, plan, K, max_coarse=2000, max_levels=8, threads=None):
        n = plan.n_free
        # at least two levels (a coarsest level of at most half the rows); below max_coarse the
        # first coarse level is the coarsest, solved by its dense inverse: on a 3 630-DOF cube two
        # levels (3 630 -> 384) take 0.045 s per solve where three (384 -> 6) took 0.080 s, the
        # V-cycle's launches setting its cost at that size
        self.max_coarse = int(min(max_coarse, max(60, n // 2)))
        self.max_levels = int(max_levels)
        self.threads = threads
        self.args = (plan.rowptr_np, plan.col_np, K.detach().to(torch.float64).cpu().numpy(), plan.grp.cpu().numpy(),
                     *dof_geometry(plan.dof_map_np, plan.X.cpu().numpy()))

    def run(self):
        return _build(*self.args, self.max_coarse, self.max_levels, self.threads)


def destroy_handle(h):
    """Free a host hierarchy that will not be exported (HostBuild.run's handle)."""
    _lib().fea_amg_destroy(h)


def dof_geometry(dof_map, X):
    """(coordinates of each free DOF's node [n, 3], its component 0..2 [n]) from the
    free-DOF map (3 N entries, -1 where fixed) and the node coordinates."""
    dof_map = np.asarray(dof_map).reshape(-1)
    X = np.asarray(X, dtype=np.float64).reshape(-1, 3)
    gdof = np.nonzero(dof_map >= 0)[0]
    order = dof_map[gdof]
    n = gdof.size
    dof_xyz = np.empty((n, 3))
    dof_comp = np.empty(n, dtype=np.int32)
    dof_xyz[order] = X[gdof // 3]
    dof_comp[order] = gdof % 3
    return dof_xyz, dof_comp


def _build(rowptr, col, val, grp, dof_xyz, dof_comp, max_coarse, max_levels, threads):
    L = _lib()
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    grp = np.ascontiguousarray(grp, dtype=np.int32)
    dof_xyz = np.ascontiguousarray(dof_xyz, dtype=np.float64)
    dof_comp = np.ascontiguousarray(dof_comp, dtype=np.int32)
    n = rowptr.size - 1
    threads = threads or max(1, min(16, len(os.sched_getaffinity(0))))
    h = ctypes.c_void_p()
    check(L.fea_amg_build(n, _np_ptr(rowptr), _np_ptr(col), _np_ptr(val), grp.size - 1, _np_ptr(grp),
                          _np_ptr(dof_xyz), _np_ptr(dof_comp), int(max_coarse), int(max_levels), int(threads),
                          ctypes.byref(h)), 'fea_amg_build')
    return h


def host_hierarchy(rowptr, col, val, grp, dof_xyz, dof_comp, max_coarse=2000, max_levels=8, threads=None):
    """The host-built hierarchy as numpy / scipy arrays (to check the set-up without a
    GPU): per level a dict with the operator A (scipy CSR; level 0 is the input), the
    prolongator P (rows x 6 nagg, all but the coarsest), the smoother's lambda_max
    estimate rho, the smoother's 6x6 group inverses (levels >= 1) and, on the coarsest
    level, the dense inverse."""
    import scipy.sparse as sp
    L = _lib()
    h = _build(rowptr, col, val, grp, dof_xyz, dof_comp, max_coarse, max_levels, threads)
    out = []
    try:
        nl = L.fea_amg_num_levels(h)
        sizes = np.zeros(6, dtype=np.int64)
        for l in range(nl):
            check(L.fea_amg_level_sizes(h, l, _np_ptr(sizes)), 'fea_amg_level_sizes')
            n_l, nagg, nblk, nnz, ngrp, nd = (int(v) for v in sizes)
            a = {'prow': np.zeros(n_l + 1, np.int32), 'pagg': np.zeros(max(1, nblk), np.int32),
                 'pval': np.zeros(max(1, 6 * nblk), np.float32), 'pt_ptr': np.zeros(max(1, nagg + 1), np.int32),
                 'pt_blk': np.zeros(max(1, nblk), np.int32), 'pt_row': np.zeros(max(1, nblk), np.int32),
                 'rowptr': np.zeros(n_l + 1, np.int32), 'col': np.zeros(max(1, nnz), np.int32),
                 'val': np.zeros(max(1, nnz), np.float64), 'dinv': np.zeros(max(1, 36 * ngrp), np.float32),
                 'dense': np.zeros(max(1, nd * nd), np.float32)}
            rho = ctypes.c_double(0.0)
            check(L.fea_amg_level_export(h, l, _np_ptr(a['prow']), _np_ptr(a['pagg']), _np_ptr(a['pval']),
                                         _np_ptr(a['pt_ptr']), _np_ptr(a['pt_blk']), _np_ptr(a['pt_row']),
                                         _np_ptr(a['rowptr']), _np_ptr(a['col']), _np_ptr(a['val']),
                                         _np_ptr(a['dinv']), ctypes.byref(rho), _np_ptr(a['dense'])),
                  'fea_amg_level_export')
            A = sp.csr_matrix((a['val'][:nnz], a['col'][:nnz], a['rowptr']), shape=(n_l, n_l))
            lev = {'n': n_l, 'A': A, 'rho': rho.value, 'nagg': nagg, 'raw': a}
            if l > 0:
                lev['dinv'] = a['dinv'][:36 * ngrp].reshape(ngrp, 6, 6).astype(np.float64)
            if nagg:
                rows = np.repeat(np.arange(n_l), np.diff(a['prow']))
                r6 = np.repeat(rows, 6)
                c6 = (6 * a['pagg'][:nblk][:, None] + np.arange(6)[None, :]).reshape(-1)
                lev['P'] = sp.csr_matrix((a['pval'][:6 * nblk].astype(np.float64), (r6, c6)), shape=(n_l, 6 * nagg))
            if nd:
                lev['dense_inv'] = a['dense'][:nd * nd].reshape(nd, nd).astype(np.float64)
            out.append(lev)
    finally:
        L.fea_amg_destroy(h)
    return out
