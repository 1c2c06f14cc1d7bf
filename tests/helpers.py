"""Test helpers: oracle model dicts built from the same meshes the product uses."""
import numpy as np

from oracle import materials as omat


def oracle_model(mesh, material, fixed_set, pressure=None, traction=None, load_set=None, rho=1.0):
    N = mesh['nodes'].shape[0]
    fixed = np.zeros((N, 3), dtype=bool)
    fixed[mesh['node_sets'][fixed_set]] = True
    m = {'nodes': mesh['nodes'].astype(np.float64), 'fixed': fixed,
         'groups': [{'etype': mesh['etype'], 'conn': mesh['conn'], 'material': material, 'rho': rho}],
         'pressures': [], 'tractions': []}
    if pressure is not None or traction is not None:
        conns, stype = facets(mesh, load_set)
        if pressure is not None:
            m['pressures'].append({'stype': stype, 'conn': conns, 'p': pressure})
        if traction is not None:
            m['tractions'].append({'stype': stype, 'conn': conns, 't': tuple(traction)})
    return m


FACE_NODES = {
    'C3D4': {0: [0, 2, 1], 1: [0, 1, 3], 2: [1, 2, 3], 3: [0, 3, 2]},
    'C3D10': {0: [0, 2, 1, 6, 5, 4], 1: [0, 1, 3, 4, 8, 7], 2: [1, 2, 3, 5, 9, 8], 3: [0, 3, 2, 7, 9, 6]},
    'C3D8': {0: [0, 3, 2, 1], 1: [4, 5, 6, 7], 2: [0, 1, 5, 4], 3: [1, 2, 6, 5], 4: [2, 3, 7, 6],
             5: [0, 4, 7, 3]},
}


def facets(mesh, name):
    """Facet connectivity of a surface set, in the reference's face node order."""
    out = []
    for idx, fid in mesh['face_sets'][name]:
        out.append(mesh['conn'][idx][:, FACE_NODES[mesh['etype']][fid]])
    conn = np.concatenate(out, axis=0)
    stype = {3: 'T3', 6: 'T6', 4: 'Q4'}[conn.shape[1]]
    return conn, stype


def nh(mu, kappa):
    return omat.NeoHookean(mu, kappa)


def mr(c10, c01, d1):
    return omat.MooneyRivlin(c10, c01, d1)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
