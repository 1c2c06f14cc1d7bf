"""Global residual / tangent assembly and reduction to free DOFs.

Test infrastructure only (see oracle/__init__.py).

A model is a plain dict (instances already concatenated in name-sorted order,
so global DOF = 3*node + component as in the flattened RGC of assembly.py:97-128):

    {'nodes': [N,3] f64,
     'groups': [{'etype': 'C3D4'|'C3D8'|'C3D10', 'conn': [E,nen] int, 'material': obj,
                 'rho': float (optional)}],
     'pressures': [{'stype': 'T3'|'T6'|'Q4', 'conn': [Es,ns] int, 'p': float}],
     'tractions': [{'stype': ..., 'conn': ..., 't': (tx,ty,tz)}],
     'barriers': [{'conn1': [F1,3] int, 'conn2': [F2,3] int or None (self contact),
                   'params': dict (optional, oracle.barrier.DEFAULTS)}],
     'fixed': [N,3] bool}
"""
import numpy as np
import scipy.sparse as sp

from . import elements as el
from . import surfaces as sf


def prepare(model):
    """Per-group Gauss precompute (Element_3D.initialize / _pre_load_gaussian)."""
    for g in model['groups']:
        if 'dNdX' not in g:
            g['dNdX'], g['wdet'], g['N'] = el.pre_load_gaussian(g['etype'], model['nodes'], g['conn'])
    for t in model.get('tractions', []):
        if 'f' not in t:
            t['f'] = sf.traction_force(t['stype'], model['nodes'], t['conn'], t['t'])
    return model


def free_mask(model):
    """RGC_remain_index after elements -> loads -> boundarys (assembly.py:157-168)."""
    N = model['nodes'].shape[0]
    m = np.zeros((N, 3), dtype=bool)
    for g in model['groups']:
        m |= el.required_dofs(g['conn'], N)
    for p in model.get('pressures', []):                  # pressure.py:161-169
        m[np.unique(p['conn'])] = True
    m &= ~model['fixed']                                 # boundary_condition.py:35-41
    return m.reshape(-1)


def assemble_full(model, U, onlyforce=False):
    """_assemble_generalized_Matrix (assembly.py:311-355): R0 [3N] and COO K (with duplicates)."""
    prepare(model)
    N = model['nodes'].shape[0]
    R = np.zeros(3 * N)
    rows, cols, vals = [], [], []
    for g in model['groups']:
        Re, Ke = el.structural_force(g['dNdX'], g['wdet'], g['conn'], U, g['material'], onlyforce)
        np.add.at(R, el.force_indices(g['conn']), Re)
        if not onlyforce:
            r, c = el.coo_indices(g['conn'])
            rows.append(r), cols.append(c), vals.append(Ke)
    pos = model['nodes'] + U
    for p in model.get('pressures', []):
        Fv, Kv = sf.pressure_force_stiffness(p['stype'], pos, p['conn'], p['p'], onlyforce)
        np.add.at(R, sf.pressure_force_idx(p['conn']), -Fv)             # assembly.py:337-341
        if not onlyforce:
            idx = sf.pressure_stiff_idx(p['conn'])
            rows.append(idx[0]), cols.append(idx[1]), vals.append(-Kv)
    for t in model.get('tractions', []):
        R -= t['f'].reshape(-1)
    for b in model.get('barriers', []):
        Eb, gb, Hb, nodes = barrier_terms(model, b, U, 1 if onlyforce else 2)
        if nodes is None:
            continue
        dof = (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(nodes.shape[0], 18)
        np.add.at(R, dof.reshape(-1), gb.reshape(-1))
        if not onlyforce:
            rows.append(np.repeat(dof, 18, axis=1).reshape(-1))
            cols.append(np.tile(dof, (1, 18)).reshape(-1))
            vals.append(Hb.reshape(-1))
    if onlyforce:
        return R, None
    return R, (np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def reduce_to_free(model, R, coo):
    """_assemble_reduced_Matrix (assembly.py:357-387) + COO->CSR coalesce (_linear_solver.py:18-19)."""
    free = free_mask(model)
    rows, cols, vals = coo
    keep = free[rows] & free[cols]
    r, c, v = rows[keep], cols[keep], vals[keep]
    r = np.unique(r, return_inverse=True)[1]
    c = np.unique(c, return_inverse=True)[1]
    n = int(free.sum())
    K = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    K.sum_duplicates()
    K.sort_indices()
    return R[free], K


def assemble_reduced(model, U):
    """assemble_Stiffness_Matrix (assembly.py:283-309): R [n_free], K csr."""
    R, coo = assemble_full(model, U)
    return reduce_to_free(model, R, coo)


def assemble_force(model, U):
    """assemble_force (assembly.py:230-281)."""
    R, _ = assemble_full(model, U, onlyforce=True)
    return R[free_mask(model)]


def total_energy(model, U):
    """_total_Potential_Energy (assembly.py:389-415): internal - sum(load potentials)."""
    prepare(model)
    E = 0.0
    for g in model['groups']:
        E += el.potential_energy(g['dNdX'], g['wdet'], g['conn'], U, g['material'])
    pos = model['nodes'] + U
    for p in model.get('pressures', []):
        E -= -p['p'] * sf.pressure_volume(p['stype'], pos, p['conn'])     # pressure.py:159
    for t in model.get('tractions', []):
        E -= float((t['f'] * U).sum())
    for b in model.get('barriers', []):
        E += float(barrier_terms(model, b, U, 0)[0].sum())          # -(-sum phi), contact.py:386-387
    return E


def barrier_terms(model, b, U, order):
    """ContactSelf / Contact (contact.py:300-711, :771-1207) at the current configuration:
    candidate pairs from the current centroids, the self-contact ratio from the initial one;
    returns (energy per pair, gradient, Hessian, pair nodes [P, 6])."""
    from . import barrier as ob
    prm = dict(ob.DEFAULTS, **b.get('params', {}))
    X0 = model['nodes']
    x = X0 + U.reshape(-1, 3)
    c1 = b['conn1']
    c2 = c1 if b.get('conn2') is None else b['conn2']
    r = prm['search_ratio'] * prm['h_t']
    if b.get('conn2') is None:
        pairs = ob.candidate_pairs(ob.centroids(x, c1), r)
        ratio = ob.self_ratio(X0, c1, pairs, prm)
        keep = ratio.sum((1, 2)) > 0
        pairs, ratio = pairs[keep], ratio[keep]
    else:
        pairs = ob.candidate_pairs(ob.centroids(x, c1), r, ob.centroids(x, c2))
        ratio = np.ones((pairs.shape[0], 3, 3))
    if pairs.shape[0] == 0:
        return np.zeros(0), None, None, None
    out = ob.barrier((X0, X0), x, c1, x, c2, pairs, ratio, prm, order=order)
    nodes = np.concatenate([c1[pairs[:, 0]], c2[pairs[:, 1]]], axis=1)
    if order == 0:
        return out, None, None, nodes
    if order == 1:
        return out[0], out[1], None, nodes
    return out[0], out[1], out[2], nodes


def gc_to_u(model, GC):
    """_GC2RGC (assembly.py:470-492): scatter free values, zero the rest."""
    free = free_mask(model)
    u = np.zeros(free.shape[0])
    u[free] = GC
    return u.reshape(-1, 3)


def csr_pattern_bruteforce(model):
    """CSR (indptr, indices) of the reduced matrix built from a dense boolean matrix.

    Independent construction used to pin the pattern: K_ij is structurally
    nonzero iff DOFs i and j share an element or a pressure facet, both free.
    """
    N = model['nodes'].shape[0]
    free = free_mask(model)
    A = np.zeros((N, N), dtype=bool)
    conns = [g['conn'] for g in model['groups']] + [p['conn'] for p in model.get('pressures', [])]
    for conn in conns:
        for e in range(conn.shape[0]):
            nodes = conn[e]
            A[np.ix_(nodes, nodes)] = True
    D = np.kron(A, np.ones((3, 3), dtype=bool))
    D = D[np.ix_(free, free)]
    indptr = np.concatenate([[0], np.cumsum(D.sum(1))]).astype(np.int64)
    indices = np.nonzero(D)[1].astype(np.int64)
    return indptr, indices
