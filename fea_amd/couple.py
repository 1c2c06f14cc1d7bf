"""Device plan of a model with Couple constraints (reference: constraints/couple.py:7-292,
reference_points.py:5-24, assembly.py:76-499).

A ``CoupledPlan`` presents the interface of ``DevicePlan`` to the solvers, over

* the *state* vector ``U`` = [3N nodal displacements | 6 per reference point (v, z)]:
  coupled node displacements are always the Rodrigues image of their point's (v, z)
  (``fea_couple_kinematics``);
* the reduced system of the generalized coordinates GC = (free node DOFs in node order,
  then free reference-point DOFs in point order), the reference's GC order.

The element, pressure and traction kernels run unchanged on an inner ``DevicePlan``
whose free set also holds the coupled node DOFs (the "extended" system, assembled in
FP64); ``fea_couple_condense`` forms the reduced residual and tangent from it by the
chain rule.  The reduced CSR pattern and the index maps are built here once, on the host.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat
from ._native import F32, F64, call
from .device import DevicePlan, _i32, row_group_tables, triple_columns

_TORCH_DT = {F32: torch.float32, F64: torch.float64}


class CoupleMap(ctypes.Structure):
    """Mirror of ``fea_couple_map`` (include/fea_amd.h)."""
    _fields_ = [('n_couples', ctypes.c_int32), ('n_cnodes', ctypes.c_int32), ('n_pairs', ctypes.c_int32),
                ('n_red', ctypes.c_int32), ('n_red_nodes', ctypes.c_int32), ('n_state_nodes', ctypes.c_int32),
                ('nnz_red', ctypes.c_int64)] + [(f, ctypes.c_void_p) for f in (
                    'cnode_ptr', 'rp_state', 'cnode', 'y', 'crow_ext', 'cr_ptr', 'ce_slot', 'ce_dof',
                    'pair_ptr', 'pe_slot', 'pe_dof', 'pair_row_slot', 'pair_col_slot', 'rp_slot', 'rp_row',
                    'rp_load', 'ff_src', 'rf_src', 'ydot', 'ydot2', 'partial')]


def couple_maps(X, node_free, dof_ext, rowptr_e, col_e, couples, rps):
    """Reduced CSR pattern and the index maps of ``fea_couple_map`` (numpy), from the
    extended CSR (rowptr_e, col_e; ext numbering ``dof_ext`` [3N]) in which the coupled
    node DOFs are rows.  ``node_free`` excludes the coupled nodes; ``couples`` are sorted by
    point and cover every point once.  Every (reduced row, column) has one source:
    node-node entries copy an ext entry, (row, point) entries and their transposes come
    from one (row, couple) pair, and the point-point block from one couple."""
    N = X.shape[0]
    n_rp = len(rps)
    # ------------------------------------------------ numbering
    n_ext = int(rowptr_e.shape[0] - 1)
    n_rn = int(node_free.sum())
    red_node = np.full(3 * N, -1, dtype=np.int64)
    red_node[node_free] = np.arange(n_rn)
    rp_free = np.array([np.asarray(r['free'], dtype=bool) for r in rps]).reshape(-1, 6) \
        if rps else np.zeros((0, 6), dtype=bool)
    red_rp = np.full(6 * n_rp, -1, dtype=np.int64)
    red_rp[rp_free.reshape(-1)] = n_rn + np.arange(int(rp_free.sum()))
    n_red = n_rn + int(rp_free.sum())
    couples = sorted(couples, key=lambda t: t[1])     # couple index = point index
    cnodes = [np.asarray(n, dtype=np.int64).reshape(-1) for n, _ in couples]
    cnode = np.concatenate(cnodes) if cnodes else np.zeros(0, dtype=np.int64)
    cnode_ptr = np.concatenate([[0], np.cumsum([c.shape[0] for c in cnodes])]).astype(np.int64)
    ncn = cnode.shape[0]
    q_couple = np.repeat(np.arange(len(couples)), 3 * np.diff(cnode_ptr))
    ext_to_red = np.full(n_ext, -1, dtype=np.int64)
    ext_to_red[dof_ext[node_free]] = red_node[node_free]
    crow_ext = dof_ext[(3 * cnode[:, None] + np.arange(3)[None, :]).reshape(-1)].astype(np.int64)
    ext_to_q = np.full(n_ext, -1, dtype=np.int64)
    ext_to_q[crow_ext] = np.arange(3 * ncn)

    # ------------------------------------------------ entries of the extended CSR
    rp_e = np.asarray(rowptr_e, dtype=np.int64)
    col_e = np.asarray(col_e, dtype=np.int64)
    rows_e = np.repeat(np.arange(n_ext), np.diff(rp_e))
    slot_e = np.arange(col_e.shape[0], dtype=np.int64)
    rf = ext_to_red[rows_e] >= 0
    cf = ext_to_red[col_e] >= 0
    ff = rf & cf
    ff_r, ff_c, ff_s = ext_to_red[rows_e[ff]], ext_to_red[col_e[ff]], slot_e[ff]
    fc = rf & ~cf
    fc_r, fc_q, fc_s = ext_to_red[rows_e[fc]], ext_to_q[col_e[fc]], slot_e[fc]
    nk = max(1, len(couples))
    key = fc_r * nk + q_couple[fc_q]
    o = np.argsort(key, kind='stable')
    key, fc_q, fc_s = key[o], fc_q[o], fc_s[o]
    pkeys, pstart = np.unique(key, return_index=True)
    pair_ptr = np.concatenate([pstart, [key.shape[0]]]).astype(np.int64)
    pair_row, pair_k = pkeys // nk, pkeys % nk
    cc = ~rf & ~cf
    cc_q, cc_q2, cc_s = ext_to_q[rows_e[cc]], ext_to_q[col_e[cc]], slot_e[cc]
    if np.any(q_couple[cc_q] != q_couple[cc_q2]):
        raise NotImplementedError('two Couple node sets share an element')
    o = np.argsort(cc_q, kind='stable')
    cc_q, cc_q2, cc_s = cc_q[o], cc_q2[o], cc_s[o]
    cr_ptr = np.concatenate([[0], np.cumsum(np.bincount(cc_q, minlength=3 * ncn))]).astype(np.int64)

    # ------------------------------------------------ reduced pattern
    R_rows, R_cols = [ff_r], [ff_c]
    pr_red = red_rp.reshape(-1, 6)[pair_k] if pair_k.shape[0] else np.zeros((0, 6), dtype=np.int64)
    pm = pr_red >= 0
    pr_row = np.repeat(pair_row[:, None], 6, axis=1)
    R_rows += [pr_row[pm], pr_red[pm]]
    R_cols += [pr_red[pm], pr_row[pm]]
    blk = []
    for k in range(len(couples)):
        r = red_rp[6 * k:6 * k + 6]
        a, b = np.meshgrid(r, r, indexing='ij')
        blk.append((a.reshape(-1), b.reshape(-1)))
        m = (a >= 0) & (b >= 0)
        R_rows.append(a[m])
        R_cols.append(b[m])
    rr = np.concatenate(R_rows)
    rc = np.concatenate(R_cols)
    keys = rr * n_red + rc
    ks = np.sort(keys)
    if np.any(ks[1:] == ks[:-1]):
        raise RuntimeError('duplicate entries in the reduced pattern')
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(ks // n_red, minlength=n_red))])
    col = (ks % n_red).astype(np.int32)

    def slot_of(r, c):
        r = np.asarray(r, dtype=np.int64)
        c = np.asarray(c, dtype=np.int64)
        out = np.full(r.shape, -1, dtype=np.int64)
        m = (r >= 0) & (c >= 0)
        out[m] = np.searchsorted(ks, r[m] * n_red + c[m])
        return out

    nnz = ks.shape[0]
    ff_src = np.full(nnz, -1, dtype=np.int64)
    ff_src[slot_of(ff_r, ff_c)] = ff_s
    pair_row_slot = slot_of(pr_row, pr_red)
    pair_col_slot = slot_of(pr_red, pr_row)
    rp_slot = np.concatenate([slot_of(a, b) for a, b in blk]) if blk else np.zeros(0, dtype=np.int64)
    diag = slot_of(np.arange(n_red), np.arange(n_red))
    if np.any(diag < 0):
        raise RuntimeError('reduced pattern misses a diagonal entry')
    rf_src = dof_ext[node_free].astype(np.int64)      # ext row of each reduced node row (node order)

    # row groups of the reduced CSR: the free DOFs of a node, then the (v) and (z)
    # halves of each point
    nf = node_free.reshape(-1, 3).sum(1)
    sizes = [nf[nf > 0]]
    for k in range(n_rp):
        for h in range(2):
            s = int(rp_free[k, 3 * h:3 * h + 3].sum())
            if s:
                sizes.append(np.array([s]))
    sizes = np.concatenate(sizes) if sizes else np.zeros(0, dtype=np.int64)
    grp, triple = row_group_tables(rowptr, col, sizes)

    Xr = np.array([np.asarray(r['node'], dtype=np.float64) for r in rps]).reshape(-1, 3)
    y = X[cnode] - np.repeat(Xr[[k for _, k in couples]], np.diff(cnode_ptr), axis=0) if ncn else \
        np.zeros((0, 3))
    load = np.array([np.asarray(r.get('load', np.zeros(6)), dtype=np.float64) for r in rps]).reshape(-1)
    return dict(n_red=n_red, n_red_nodes=n_rn, nnz=int(nnz), rowptr=rowptr, col=col, diag=diag, grp=grp,
                triple=triple, red_node=red_node, red_rp=red_rp, rp_free=rp_free, cnode=cnode,
                cnode_ptr=cnode_ptr, y=y, load=load, crow_ext=crow_ext, cr_ptr=cr_ptr, ce_slot=cc_s,
                ce_dof=cc_q2, pair_ptr=pair_ptr, pe_slot=fc_s, pe_dof=fc_q, pair_row=pair_row, pair_k=pair_k,
                pair_row_slot=pair_row_slot, pair_col_slot=pair_col_slot, rp_slot=rp_slot, ff_src=ff_src,
                rf_src=rf_src)


class CoupledPlan:
    """``couples``: [(coupled node ids, reference point index)], one couple per point;
    ``rps``: [{'node': (3,), 'free': 6 bools, 'load': 6 floats (force, moment)}]."""

    def __init__(self, X, node_free, groups, couples, rps, device=None):
        nat.require_gpu()
        if nat.lib().fea_couple_map_size() != ctypes.sizeof(CoupleMap):
            raise nat.FeaNativeError('fea_couple_map layout differs from the library')
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float64).reshape(-1, 3))
        N = X.shape[0]
        self.n_nodes = N
        self.n_rp = len(rps)
        node_free = np.asarray(node_free, dtype=bool).reshape(-1).copy()
        used = sorted(k for _, k in couples)
        if len(set(used)) != len(used):
            raise NotImplementedError('a reference point coupled to more than one node set')
        if set(range(self.n_rp)) - set(used):
            raise NotImplementedError('a reference point without a Couple has no stiffness')
        cnodes = [np.asarray(n, dtype=np.int64).reshape(-1) for n, _ in couples]
        allc = np.concatenate(cnodes) if cnodes else np.zeros(0, dtype=np.int64)
        if np.unique(allc).shape[0] != allc.shape[0]:
            raise NotImplementedError('a node in two Couple sets (or twice in one)')
        in_elem = np.zeros(N, dtype=bool)
        for g in groups:
            if g.kind == 'elem':
                in_elem[g.conn_np.reshape(-1)] = True
        if not np.all(in_elem[allc]):
            raise ValueError('coupled nodes must belong to an element')
        node_free.reshape(-1, 3)[allc] = False
        ext = node_free.copy()
        ext.reshape(-1, 3)[allc] = True
        self.inner = DevicePlan(X, ext, groups, device=device)
        self.device = self.inner.device
        dev = self.device
        self.groups = self.inner.groups

        c = couple_maps(X, node_free, self.inner.dof_map_np, self.inner.rowptr_np, self.inner.col_np,
                        sorted(couples, key=lambda t: t[1]), rps)
        self.maps = c
        n_red, n_rn, nnz, ncn = c['n_red'], c['n_red_nodes'], c['nnz'], c['cnode'].shape[0]
        rowptr, col, diag, grp, triple = c['rowptr'], c['col'], c['diag'], c['grp'], c['triple']
        red_node, red_rp, rp_free, cnode = c['red_node'], c['red_rp'], c['rp_free'], c['cnode']
        couples = sorted(couples, key=lambda t: t[1])

        # ------------------------------------------------ device arrays
        self.n_free, self.nnz, self.n_red_nodes = n_red, int(nnz), n_rn
        self.rowptr_np, self.col_np = rowptr.astype(np.int32), col
        self.rowptr, self.col = _i32(rowptr, dev), _i32(col, dev)
        self.diag_slot = _i32(diag, dev)
        if grp is None:
            self.grp, self.n_groups, self.spmv_groups, self.triple = None, 0, 0, False
            self.spmv_col = self.col
        else:
            self.grp = _i32(grp, dev)
            self.n_groups = int(grp.shape[0] - 1)
            self.triple = triple
            self.spmv_groups = -self.n_groups if triple else self.n_groups
            # the column stream of the SpMV: compact per-triple columns on a triple-structured CSR
            self.spmv_col = _i32(triple_columns(rowptr, col, grp), dev) if triple else self.col
        self._load = c['load']
        arr = dict(
            cnode_ptr=_i32(c['cnode_ptr'], dev), rp_state=_i32(3 * N + 6 * np.arange(len(couples)), dev),
            cnode=_i32(cnode, dev), y=torch.as_tensor(np.ascontiguousarray(c['y']).reshape(-1), device=dev),
            crow_ext=_i32(c['crow_ext'], dev), cr_ptr=_i32(c['cr_ptr'], dev), ce_slot=_i32(c['ce_slot'], dev),
            ce_dof=_i32(c['ce_dof'], dev), pair_ptr=_i32(c['pair_ptr'], dev), pe_slot=_i32(c['pe_slot'], dev),
            pe_dof=_i32(c['pe_dof'], dev), pair_row_slot=_i32(c['pair_row_slot'].reshape(-1), dev),
            pair_col_slot=_i32(c['pair_col_slot'].reshape(-1), dev), rp_slot=_i32(c['rp_slot'], dev),
            rp_row=_i32(red_rp, dev), rp_load=torch.as_tensor(self._load, device=dev),
            ff_src=_i32(c['ff_src'], dev), rf_src=_i32(c['rf_src'], dev),
            ydot=torch.zeros(max(1, 9 * ncn), dtype=torch.float64, device=dev),
            ydot2=torch.zeros(max(1, 27 * ncn), dtype=torch.float64, device=dev),
            partial=torch.zeros(max(1, 126 * ncn), dtype=torch.float64, device=dev))
        self._arr = arr                                    # keeps every device array alive
        m = CoupleMap()
        m.n_couples, m.n_cnodes, m.n_pairs = len(couples), ncn, int(c['pair_row'].shape[0])
        m.n_red, m.n_red_nodes, m.n_state_nodes, m.nnz_red = n_red, n_rn, N, int(nnz)
        for name, t in arr.items():
            setattr(m, name, t.data_ptr())
        self._map = m

        # state <-> GC maps: the augmented state has N + 2 n_rp "nodes" of 3 entries
        aug = np.concatenate([red_node, red_rp]).astype(np.int32)
        self.n_state = 3 * N + 6 * self.n_rp
        self.dof_map_aug = _i32(aug, dev)
        self._gc_pos = torch.as_tensor(np.nonzero(aug >= 0)[0][np.argsort(aug[aug >= 0])], device=dev)
        self.node_free = node_free
        self.rp_free = rp_free.reshape(-1)
        self.cnode_np = cnode
        self.R = torch.zeros(max(1, n_red), dtype=torch.float64, device=dev)[:n_red]
        self._val = {}

    # ------------------------------------------------------------------ helpers
    @staticmethod
    def _s():
        return nat.stream_ptr()

    def _mp(self):
        return ctypes.addressof(self._map)

    def _check_state(self, U):
        if U.dtype != torch.float64 or not U.is_cuda or U.numel() != self.n_state or not U.is_contiguous():
            raise ValueError('the state must be a contiguous float64 CUDA tensor with %d entries '
                             '(3 per node, then 6 per reference point)' % self.n_state)

    def kinematics(self, U, with_derivs=False):
        """Coupled node displacements from the points' (v, z) (and their z-derivatives)."""
        self._check_state(U)
        call('fea_couple_kinematics', self._mp(), U, 1 if with_derivs else 0, self._s())

    def sync(self, U):
        """Restore the coupled node displacements of a state built by linear combination."""
        self.kinematics(U, False)

    def nodal(self, U):
        return U[:3 * self.n_nodes]

    def csr_values(self, dtype):
        if dtype not in self._val:
            self._val[dtype] = torch.zeros(max(1, self.nnz), dtype=_TORCH_DT[dtype], device=self.device)
        return self._val[dtype]

    # ------------------------------------------------------------------ assembly
    def residual(self, U, load_scale=1.0, dtype=F64, out=None):
        """Reduced residual (assemble_force with Couple.modify_R_K, assembly.py:230-281)."""
        self.kinematics(U, True)
        R_ext = self.inner.residual(self.nodal(U), load_scale, dtype)
        R = self.R if out is None else out
        call('fea_couple_condense', self._mp(), F64, None, R_ext, float(load_scale), None, R, self._s())
        return R

    def tangent(self, U, load_scale=1.0, dtype=F64, with_residual=True, val_dtype=None):
        """Reduced (R, K) (assemble_Stiffness_Matrix with Couple.modify_R_K, assembly.py:283-387).
        The extended system is assembled with FP64 values; ``val_dtype`` is the reduced
        matrix's storage."""
        if dtype != F64:
            raise NotImplementedError('Couple models assemble the extended system in FP64 arithmetic')
        self.kinematics(U, True)
        R_ext, K_ext = self.inner.tangent(self.nodal(U), load_scale, F64, with_residual=True, val_dtype=F64)
        val_dtype = dtype if val_dtype is None else val_dtype
        val = self.csr_values(val_dtype)
        call('fea_couple_condense', self._mp(), val_dtype, K_ext, R_ext, float(load_scale), val, self.R,
             self._s())
        return (self.R if with_residual else None), val

    def energy(self, U, load_scale=1.0):
        """Total potential energy with the points' load potentials F.v + M.z
        (concentrate_force.py:38-43, moment.py:38-43)."""
        self._check_state(U)
        E = self.inner.energy(self.nodal(U), load_scale)
        if self.n_rp and np.any(self._load != 0.0):
            rp = U[3 * self.n_nodes:].cpu().numpy()
            E -= load_scale * float(self._load @ rp)
        return E

    def mass_matrix(self):
        raise NotImplementedError('the reference has no mass for Couple constraints '
                                  '(Couple.modify_mass_matrix returns nothing, couple.py:45-55)')

    def lumped_mass(self):
        self.mass_matrix()

    def set_coordinates(self, X):
        raise NotImplementedError('moving the nodes of a model with Couple constraints')

    # ------------------------------------------------------------------ vectors
    def dot(self, a, b):
        return self.inner.dot(a, b)

    def _scratch(self, key, n):
        return self.inner._scratch('c_' + key, n)

    def dof_axpy(self, U0, alpha, d, out):
        """out = U0 + alpha * d on the free generalized coordinates, then the coupled
        nodes from the points (_GC2RGC with Couple.modify_RGC)."""
        call('fea_dof_axpy', self.n_state // 3, self.dof_map_aug, U0, float(alpha), d, out, self._s())
        self.kinematics(out, False)
        return out

    def gc_to_u(self, GC, out=None):
        U = torch.zeros(self.n_state, dtype=torch.float64, device=self.device) if out is None else out
        zero = torch.zeros_like(U)
        return self.dof_axpy(zero, 1.0, GC, U)

    def u_to_gc(self, U):
        return U.reshape(-1)[self._gc_pos].clone()

    # ------------------------------------------------------------------ linear
    def spmv(self, val, x, y, mat_dtype, vec_dtype=None):
        vec_dtype = mat_dtype if vec_dtype is None else vec_dtype
        call('fea_spmv', mat_dtype, vec_dtype, self.n_free, self.nnz, self.spmv_groups, self.grp, self.rowptr,
             self.spmv_col, val, x, y, self._s())
        return y

    def csr_indices_coo(self):
        rows = np.repeat(np.arange(self.n_free, dtype=np.int64), np.diff(self.rowptr_np))
        return torch.as_tensor(np.stack([rows, self.col_np.astype(np.int64)]), device=self.device)
