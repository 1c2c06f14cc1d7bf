"""Device-side assembly plan: HBM layout + orchestration of the HIP kernels.

A ``DevicePlan`` owns, on one GPU:

* ``X``        float64 [N,3]  reference coordinates of every node of every
                              instance (instances concatenated in the RGC block
                              order of assembly.py:104-128);
* ``dof_map``  int32 [3N]     free-DOF index (GC numbering) or -1;
* the free-DOF CSR (``rowptr``, ``col``, ``diag_slot``) and the node-pair block
  tables + staging maps produced by ``fea_pattern_create``;
* one record per *group*: an element group (C3D4/C3D8/C3D10 + material), a
  pressure facet group or a traction facet group, with its int32
  connectivity and its ``dst_k`` / ``dst_r`` staging maps;
* staging buffers ``kstage`` (per precision) and ``rstage`` float64.

Every array is allocated once when the plan is built; assembly calls only
launch kernels into these buffers.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat
from ._native import F32, F64, call, ptr

_TORCH_DT = {F32: torch.float32, F64: torch.float64}

# Direct element assembly (fea_elem_tangent_direct: element colours adding their blocks
# into the CSR, no element staging) for plans built while this is True.  Off by default:
# on the 131k-brick beam the eight colour launches (0.83 ms per tangent) are slower than
# staging and the gather (0.62 ms), although they move a fraction of the bytes.
DIRECT_ASSEMBLY = False


def _i32(a, device):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)


def row_group_tables(rowptr, col, sizes):
    """Group table grp (group g = rows grp[g] .. grp[g+1]-1, ``sizes`` rows each, at most 3)
    of a CSR whose rows in a group share one column list, and whether the pattern is
    triple-structured (every group 3 rows, its columns whole aligned triples: the
    triple-blocked SpMV).  Returns (None, False) when the rows of a group differ in length."""
    sizes = np.asarray(sizes, dtype=np.int64)
    n = rowptr.shape[0] - 1
    if sizes.sum() != n or np.any(sizes < 1) or np.any(sizes > 3):
        return None, False
    grp = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    rl = np.diff(rowptr)
    for j in (1, 2):
        sel = sizes > j
        if not np.array_equal(rl[grp[:-1][sel] + j], rl[grp[:-1][sel]]):
            return None, False
    nnz = int(rowptr[-1])
    triple = bool(np.all(sizes == 3)) and nnz > 0
    if triple:
        triple = bool(np.all(rl % 3 == 0))
    if triple:
        # every row a whole number of triples, so the triples of all rows are the rows of
        # col reshaped [nnz/3, 3]: each must read c, c+1, c+2
        C = np.asarray(col[:nnz]).reshape(-1, 3)
        triple = bool(np.all(C[:, 1] == C[:, 0] + 1) and np.all(C[:, 2] == C[:, 0] + 2))
    return grp, triple


def triple_columns(rowptr, col, grp):
    """The compact column array of a triple-structured CSR that the triple-blocked SpMV reads:
    the first column of each neighbour triple of each group's head row, group g's triples at
    rowptr[grp[g]] / 9 + j (every group holds 3 rows whose lengths are multiples of 3)."""
    heads = np.asarray(grp[:-1], dtype=np.int64)
    s0 = np.asarray(rowptr, dtype=np.int64)[heads]
    cnt = (np.asarray(rowptr, dtype=np.int64)[heads + 1] - s0) // 3
    start = np.cumsum(cnt) - cnt
    if np.any(s0 % 9) or not np.array_equal(start, s0 // 9):
        raise ValueError('not a triple-structured CSR')
    idx = np.repeat(s0, cnt) + 3 * (np.arange(int(cnt.sum())) - np.repeat(start, cnt))
    return np.ascontiguousarray(np.asarray(col)[idx], dtype=np.int32)


class Group:
    """One contiguous batch of elements or facets sharing a kernel configuration."""

    def __init__(self, kind, ntype, conn, params=None, load=None, in_k=True, rho=1.0, name=''):
        self.kind = kind            # 'elem' | 'pressure' | 'traction' | 'barrier' (contact pairs)
        self.ntype = int(ntype)     # nodes per element / facet (= C-ABI type id)
        self.conn_np = np.ascontiguousarray(conn, dtype=np.int32)
        self.n = int(self.conn_np.shape[0])
        self.params = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        self.load = load            # the load object (pressure magnitude / traction vector)
        self.in_k = bool(in_k)
        self.rho = float(rho)
        self.name = name
        self.conn = None
        self.dst_k = None
        self.dst_r = None


class DevicePlan:
    def __init__(self, X, free_mask, groups, device=None):
        nat.require_gpu()
        self.device = torch.device('cuda') if device is None else torch.device(device)
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(-1, 3))
        self.n_nodes = X.shape[0]
        free = np.asarray(free_mask, dtype=bool).reshape(-1)
        if free.shape[0] != 3 * self.n_nodes:
            raise ValueError('free mask has %d entries, expected %d' % (free.shape[0], 3 * self.n_nodes))
        dof = np.full(3 * self.n_nodes, -1, dtype=np.int32)
        dof[free] = np.arange(int(free.sum()), dtype=np.int32)
        self.free_mask = free
        self.dof_map_np = dof
        self.groups = list(groups)
        self.X = torch.as_tensor(X, device=self.device)
        self.dof_map = _i32(dof, self.device)
        self._build_pattern()
        self._kstage = {}
        self._csr_val = {}
        self.rstage = torch.zeros(max(1, 3 * self.n_rstage), dtype=torch.float64, device=self.device)
        self.R = torch.zeros(max(1, self.n_free), dtype=torch.float64, device=self.device)[:self.n_free]
        # energy partial segments: one per group
        self._epart_off = []
        off = 0
        for g in self.groups:
            if g.kind == 'elem':
                nb = nat.lib().fea_elem_energy_blocks(g.n)
            else:
                nb = nat.lib().fea_surface_blocks(g.n)
            self._epart_off.append((off, nb))
            off += nb
        self._epart = torch.zeros(max(1, off), dtype=torch.float64, device=self.device)
        self._eout = torch.zeros(max(1, len(self.groups)), dtype=torch.float64, device=self.device)
        self._nodal = None

    # ------------------------------------------------------------------ pattern
    def _build_pattern(self):
        L = nat.lib()
        ng = len(self.groups)
        conns = [g.conn_np for g in self.groups]
        conn_ptrs = (ctypes.c_void_p * max(1, ng))(*[c.ctypes.data for c in conns])
        n_elem = np.array([g.n for g in self.groups], dtype=np.int32)
        nen = np.array([g.ntype for g in self.groups], dtype=np.int32)
        in_k = np.array([1 if g.in_k else 0 for g in self.groups], dtype=np.int32)
        handle = ctypes.c_void_p()
        rc = L.fea_pattern_create(self.n_nodes, self.dof_map_np.ctypes.data, ng, conn_ptrs,
                                  n_elem.ctypes.data, nen.ctypes.data, in_k.ctypes.data,
                                  ctypes.byref(handle))
        nat.check(rc, 'fea_pattern_create')
        try:
            sizes = np.zeros(5, dtype=np.int64)
            nat.check(L.fea_pattern_sizes(handle, sizes.ctypes.data), 'fea_pattern_sizes')
            self.n_free, self.nnz, self.n_blocks, self.n_kstage, self.n_rstage = [int(s) for s in sizes]
            rowptr = np.zeros(self.n_free + 1, dtype=np.int32)
            col = np.zeros(max(1, self.nnz), dtype=np.int32)
            diag = np.zeros(max(1, self.n_free), dtype=np.int32)
            nat.check(L.fea_pattern_export_csr(handle, rowptr.ctypes.data, col.ctypes.data, diag.ctypes.data),
                      'fea_pattern_export_csr')
            blk_a = np.zeros(max(1, self.n_blocks), dtype=np.int32)
            blk_b = np.zeros_like(blk_a)
            blk_cp = np.zeros_like(blk_a)
            kseg = np.zeros(self.n_blocks + 1, dtype=np.int32)
            rseg = np.zeros(self.n_nodes + 1, dtype=np.int32)
            nat.check(L.fea_pattern_export_blocks(handle, blk_a.ctypes.data, blk_b.ctypes.data,
                                                  blk_cp.ctypes.data, kseg.ctypes.data, rseg.ctypes.data),
                      'fea_pattern_export_blocks')
            dev = self.device
            self.rowptr_np, self.col_np = rowptr, col[:self.nnz]
            # node row groups: the free rows of one node are consecutive and share columns
            nf = self.free_mask.reshape(-1, 3).sum(1)
            grp, triple = row_group_tables(rowptr, col[:self.nnz], nf[nf > 0])
            if grp is None:
                raise RuntimeError('CSR rows of a node differ in length')
            self.n_groups = int(grp.shape[0] - 1)
            self.grp = _i32(grp, dev)
            self.triple = triple
            self.spmv_groups = -self.n_groups if triple else self.n_groups
            self.rowptr, self.col = _i32(rowptr, dev), _i32(col, dev)
            # the column stream of the SpMV: compact per-triple columns on a triple-structured CSR
            self.spmv_col = _i32(triple_columns(rowptr, col, grp), dev) if triple else self.col
            self.diag_slot = _i32(diag, dev)
            self.blk_a, self.blk_b, self.blk_colpos = _i32(blk_a, dev), _i32(blk_b, dev), _i32(blk_cp, dev)
            self.kseg, self.rseg = _i32(kseg, dev), _i32(rseg, dev)
            rest_dk = {}
            for gi, g in enumerate(self.groups):
                dk = np.zeros(max(1, g.n * g.ntype * g.ntype), dtype=np.int32) if g.in_k else None
                dr = np.zeros(max(1, g.n * g.ntype), dtype=np.int32)
                nat.check(L.fea_pattern_export_group(handle, gi, None if dk is None else dk.ctypes.data,
                                                     dr.ctypes.data), 'fea_pattern_export_group')
                g.conn = _i32(g.conn_np if g.n else np.zeros((1, g.ntype)), dev)
                g.dst_k = None if dk is None else _i32(dk, dev)
                g.dst_r = _i32(dr, dev)
                g.colpos = g.order = None
                if DIRECT_ASSEMBLY and g.kind == 'elem' and g.in_k and g.n > 0:
                    # direct assembly: the CSR column of each (element, a, b) block and a
                    # colouring of the elements (no two of a colour share a node)
                    cp = np.zeros(g.n * g.ntype * g.ntype, dtype=np.int32)
                    nat.check(L.fea_pattern_export_group_colpos(handle, gi, cp.ctypes.data),
                              'fea_pattern_export_group_colpos')
                    color = np.zeros(g.n, dtype=np.int32)
                    nc = ctypes.c_int32(0)
                    nat.check(L.fea_color_elements(self.n_nodes, g.n, g.ntype, g.conn_np.ctypes.data,
                                                   color.ctypes.data, ctypes.byref(nc)), 'fea_color_elements')
                    g.colpos = _i32(cp, dev)
                    g.order = _i32(np.argsort(color, kind='stable'), dev)
                    g.color_ptr = np.concatenate([[0], np.cumsum(np.bincount(color, minlength=nc.value))]).tolist()
                elif g.in_k and dk is not None:
                    rest_dk[gi] = dk
            # direct assembly of the element stiffness when every element group has its
            # colouring (set False to use the staged path on such a plan)
            elems = [g for g in self.groups if g.kind == 'elem' and g.n > 0]
            self.direct_assembly = bool(elems) and all(g.colpos is not None for g in elems)
            # the other stiffness groups (pressure / traction facets, contact pairs) then keep
            # a staging of their own: their slots ranked in order, which keeps each block's
            # (group, element, a, b) order, with kseg_rest the rank of each block's start
            self.n_kstage_rest, self.kseg_rest = 0, None
            for g in self.groups:
                g.dst_k_rest = None
            if self.direct_assembly:
                slots = np.sort(np.concatenate([d[d >= 0] for d in rest_dk.values()]) if rest_dk
                                else np.zeros(0, dtype=np.int32))
                self.n_kstage_rest = int(slots.size)
                self.kseg_rest = _i32(np.searchsorted(slots, kseg), dev)
                for gi, d in rest_dk.items():
                    self.groups[gi].dst_k_rest = _i32(np.where(d >= 0, np.searchsorted(slots, d), -1), dev)
        finally:
            L.fea_pattern_destroy(handle)

    def nodal_staging(self):
        """All-node staging for sensitivities: (seg [N+1] int32, dst per group,
        two stage buffers).  Rows are ordered (group, element, node) inside each
        node, like the residual staging, but constrained nodes keep their rows."""
        if self._nodal is None:
            flat = [g.conn_np.reshape(-1).astype(np.int64) for g in self.groups]
            allc = np.concatenate(flat) if flat else np.zeros(0, dtype=np.int64)
            order = np.argsort(allc, kind='stable')
            slot = np.empty(allc.shape[0], dtype=np.int64)
            slot[order] = np.arange(allc.shape[0])
            seg = np.concatenate([[0], np.cumsum(np.bincount(allc, minlength=self.n_nodes))])
            dsts, off = [], 0
            for f in flat:
                dsts.append(_i32(slot[off:off + f.shape[0]] if f.shape[0] else np.zeros(1), self.device))
                off += f.shape[0]
            stage = [torch.zeros(max(1, 3 * allc.shape[0]), dtype=torch.float64, device=self.device)
                     for _ in range(2)]
            self._nodal = (_i32(seg, self.device), dsts, stage)
        return self._nodal

    def set_coordinates(self, X):
        """Move the nodes (shape changes; the sparsity pattern is unchanged)."""
        X = torch.as_tensor(X, dtype=torch.float64, device=self.device).reshape(-1)
        if X.numel() != 3 * self.n_nodes:
            raise ValueError('X must have %d entries' % (3 * self.n_nodes))
        self.X.view(-1).copy_(X)
        self._mass = None

    def kstage_rest(self, dtype):
        """Staging of the non-element stiffness groups under direct assembly."""
        key = ('rest', dtype)
        if key not in self._kstage:
            self._kstage[key] = torch.zeros(max(1, 9 * self.n_kstage_rest), dtype=_TORCH_DT[dtype],
                                            device=self.device)
        return self._kstage[key]

    def kstage(self, dtype):
        if dtype not in self._kstage:
            self._kstage[dtype] = torch.zeros(max(1, 9 * self.n_kstage), dtype=_TORCH_DT[dtype],
                                              device=self.device)
        return self._kstage[dtype]

    def csr_values(self, dtype):
        if dtype not in self._csr_val:
            self._csr_val[dtype] = torch.zeros(max(1, self.nnz), dtype=_TORCH_DT[dtype],
                                               device=self.device)
        return self._csr_val[dtype]

    # ------------------------------------------------------------------ kernels
    @staticmethod
    def _s():
        return nat.stream_ptr()

    def _check_u(self, U):
        if U.dtype != torch.float64 or not U.is_cuda or U.numel() != 3 * self.n_nodes or not U.is_contiguous():
            raise ValueError('U must be a contiguous float64 CUDA tensor with %d entries' % (3 * self.n_nodes))

    def residual(self, U, load_scale=1.0, dtype=F64, out=None):
        """R = f_int(U) - f_ext on free DOFs (assemble_force, assembly.py:230-281)."""
        self.residual_stage(U, load_scale, dtype)
        R = self.R if out is None else out
        nat.ops().gather_residual(self.rseg, self.dof_map, self.rstage, R)
        return R

    def residual_stage(self, U, load_scale=1.0, dtype=F64):
        """Launch every group's residual kernel into the staging array (no gather)."""
        self._check_u(U)
        s = self._s()
        for g in self.groups:
            if g.n == 0:
                continue
            if g.kind == 'elem':
                nat.ops().elem_residual(g.ntype, dtype, g.conn, self.X, U, g.params.tolist(), g.dst_r,
                                        self.rstage)
            elif g.kind == 'pressure':
                call('fea_surface_pressure', g.ntype, dtype, g.n, g.conn, self.X, U,
                     float(g.load.pressure) * load_scale, None, None, g.dst_r, self.rstage, s)
            elif g.kind == 'traction':
                t = np.asarray(g.load.traction, dtype=np.float64)
                call('fea_surface_traction', g.ntype, g.n, g.conn, self.X, U,
                     t, float(load_scale), g.dst_r, self.rstage, None, s)
            elif g.kind == 'barrier':
                g.load.assemble(U, None, None, F64, g.dst_r, self.rstage)

    def tangent(self, U, load_scale=1.0, dtype=F64, with_residual=True, val_dtype=None, stage_dtype=None):
        """Tangent CSR values (``val_dtype``, default ``dtype``) and optionally the float64 residual.

        Replaces _assemble_generalized_Matrix + _assemble_reduced_Matrix
        (assembly.py:311-387).  ``dtype`` is the precision of the element
        contraction, ``stage_dtype`` (default ``dtype``; F32 under F64 rounds each
        element block once as it is staged) that of the staged blocks, which are
        summed in it and rounded once to ``val_dtype``.  With ``with_residual`` the
        residual comes from the same element pass, accumulated in float64 whatever
        ``dtype``; a pressure facet computes in the staging precision, so with F32
        staging its residual comes from a second, float64 residual-only launch.
        """
        self._check_u(U)
        s = self._s()
        direct = self.direct_assembly
        val_dtype = dtype if val_dtype is None else val_dtype
        stage = dtype if stage_dtype is None else stage_dtype
        if stage != dtype and not (dtype == F64 and stage == F32):
            raise ValueError('staging precision F%d under element arithmetic F%d' % (stage, dtype))
        elem_code = nat.F64_STAGE_F32 if stage != dtype else dtype
        ks = self.kstage_rest(stage) if direct else self.kstage(stage)
        rs = self.rstage if with_residual else None
        for g in self.groups:
            if g.n == 0:
                continue
            dk = g.dst_k_rest if direct else g.dst_k
            if g.kind == 'elem':
                if not direct:
                    nat.ops().elem_residual_tangent(g.ntype, elem_code, g.conn, self.X, U, g.params.tolist(),
                                                    g.dst_k, ks, g.dst_r if with_residual else None, rs)
            elif g.kind == 'pressure':
                pr = float(g.load.pressure) * load_scale
                if stage == F64 or not with_residual:
                    call('fea_surface_pressure', g.ntype, stage, g.n, g.conn, self.X, U, pr, dk, ks,
                         g.dst_r if with_residual else None, rs, s)
                else:
                    call('fea_surface_pressure', g.ntype, stage, g.n, g.conn, self.X, U, pr, dk, ks, None,
                         None, s)
                    call('fea_surface_pressure', g.ntype, F64, g.n, g.conn, self.X, U, pr, None, None, g.dst_r,
                         rs, s)
            elif g.kind == 'traction' and with_residual:
                t = np.asarray(g.load.traction, dtype=np.float64)
                call('fea_surface_traction', g.ntype, g.n, g.conn, self.X, U,
                     t, float(load_scale), g.dst_r, self.rstage, None, s)
            elif g.kind == 'barrier':
                g.load.assemble(U, dk, ks, stage, g.dst_r if with_residual else None, rs)
        val = self.csr_values(val_dtype)
        if direct and self.n_kstage_rest == 0:
            val.zero_()                     # the element colours add into it
        else:
            # staged sums (under direct assembly only the non-element groups', zero elsewhere:
            # this also clears the values for the element colours, which add to them)
            nat.ops().gather_csr(stage, val_dtype, self.blk_a, self.blk_b, self.blk_colpos,
                                 self.kseg_rest if direct else self.kseg, self.dof_map, self.rowptr, ks, val)
        if direct:
            for g in self.groups:
                if g.kind != 'elem' or g.n == 0:
                    continue
                params = g.params.tolist()
                for c in range(len(g.color_ptr) - 1):
                    nat.ops().elem_tangent_direct(g.ntype, dtype, val_dtype, g.order[g.color_ptr[c]:g.color_ptr[c + 1]],
                                                  g.conn, self.X, U, params, self.dof_map, self.rowptr, g.colpos, val,
                                                  self.triple, g.dst_r if with_residual else None, rs)
        R = None
        if with_residual:
            R = self.R
            nat.ops().gather_residual(self.rseg, self.dof_map, self.rstage, R)
        return R, val

    def energy_terms(self, U, load_scale=1.0):
        """Per-group energy terms device: elem -> strain energy, pressure -> V, traction -> f.u."""
        self._check_u(U)
        s = self._s()
        for gi, g in enumerate(self.groups):
            off, nb = self._epart_off[gi]
            if g.n == 0:
                continue
            seg = self._epart[off:off + nb]
            if g.kind == 'elem':
                nat.ops().elem_energy(g.ntype, g.conn, self.X, U, g.params.tolist(), seg)
            elif g.kind == 'pressure':
                call('fea_surface_volume', g.ntype, g.n, g.conn, self.X, U, seg, s)
            elif g.kind == 'traction':
                t = np.asarray(g.load.traction, dtype=np.float64)
                call('fea_surface_traction', g.ntype, g.n, g.conn, self.X, U,
                     t, 1.0, None, None, seg, s)
            else:
                continue                    # barrier energy: evaluated over the exact pairs by LivePlan
            call('fea_reduce_sum', nb, seg, self._eout[gi:gi + 1], s)
        return self._eout

    def energy(self, U, load_scale=1.0):
        """Total potential energy (_total_Potential_Energy, assembly.py:389-415), host float."""
        terms = self.energy_terms(U, load_scale).cpu().numpy()
        E = 0.0
        for gi, g in enumerate(self.groups):
            if g.n == 0 or g.kind == 'barrier':
                continue
            if g.kind == 'elem':
                E += terms[gi]
            elif g.kind == 'pressure':
                E += float(g.load.pressure) * load_scale * terms[gi]     # - (-p V)
            elif g.kind == 'traction':
                E -= load_scale * terms[gi]                              # - f.u
        return float(E)

    def mass_matrix(self):
        """Consistent mass in the tangent's free-DOF CSR pattern, float64 values
        (Assembly.assemble_mass_matrix, assembly.py:421-443); computed once."""
        if getattr(self, '_mass', None) is None:
            s = self._s()
            ks = self.kstage(F64)
            ks.zero_()
            for g in self.groups:
                if g.kind == 'elem' and g.n:
                    call('fea_elem_mass', g.ntype, g.n, g.conn, self.X, g.rho, g.dst_k, ks, s)
            val = torch.zeros(max(1, self.nnz), dtype=torch.float64, device=self.device)
            call('fea_gather_csr', F64, F64, self.n_blocks, self.blk_a, self.blk_b, self.blk_colpos,
                 self.kseg, self.dof_map, self.rowptr, ks, val, s)
            self._mass = val
        return self._mass

    def lumped_mass(self):
        """Row-sum lumped mass on free DOFs (dynamic_explicit.py:147-166)."""
        s = self._s()
        self.rstage.zero_()
        for g in self.groups:
            if g.kind == 'elem' and g.n:
                call('fea_elem_lumped_mass', g.ntype, g.n, g.conn, self.X, self.dof_map,
                     g.rho, g.dst_r, self.rstage, s)
        m = torch.zeros(max(1, self.n_free), dtype=torch.float64, device=self.device)[:self.n_free]
        call('fea_gather_residual', self.n_nodes, self.rseg, self.dof_map, self.rstage, m, s)
        return m

    # ------------------------------------------------------------------ vectors
    def dot(self, a, b):
        """a.b of float64 vectors via the deterministic partial + final reduction (device scalar)."""
        n = a.numel()
        nb = nat.lib().fea_vec_blocks(n)
        part = self._scratch('dot', nb)
        out = self._scratch('dot_out', 1)
        s = self._s()
        call('fea_dot_partial', n, a, b, part, s)
        call('fea_reduce_sum', nb, part, out, s)
        return out

    def _scratch(self, key, n):
        if not hasattr(self, '_scr'):
            self._scr = {}
        t = self._scr.get(key)
        if t is None or t.numel() < n:
            t = torch.zeros(max(1, n), dtype=torch.float64, device=self.device)
            self._scr[key] = t
        return t[:n]

    def dof_axpy(self, U0, alpha, d, out):
        """out = U0 + alpha * d scattered through dof_map (GC0 + alpha dGC as nodal field)."""
        call('fea_dof_axpy', self.n_nodes, self.dof_map, U0, float(alpha), d, out, self._s())
        return out

    def gc_to_u(self, GC, out=None):
        U = torch.zeros(self.n_nodes * 3, dtype=torch.float64, device=self.device) if out is None else out
        zero = torch.zeros_like(U)
        return self.dof_axpy(zero, 1.0, GC, U)

    def u_to_gc(self, U):
        if getattr(self, '_free_idx', None) is None:
            self._free_idx = torch.as_tensor(np.nonzero(self.free_mask)[0], device=self.device)
        return U.reshape(-1)[self._free_idx].clone()

    # ------------------------------------------------------------------ linear
    def spmv(self, val, x, y, mat_dtype, vec_dtype=None):
        vec_dtype = mat_dtype if vec_dtype is None else vec_dtype
        call('fea_spmv', mat_dtype, vec_dtype, self.n_free, self.nnz, self.spmv_groups, self.grp, self.rowptr,
             self.spmv_col, val, x, y, self._s())
        return y

    def csr_indices_coo(self):
        """(row, col) int64 [2, nnz] of the CSR (COO view in CSR order)."""
        rows = np.repeat(np.arange(self.n_free, dtype=np.int64), np.diff(self.rowptr_np))
        return torch.as_tensor(np.stack([rows, self.col_np.astype(np.int64)]), device=self.device)
