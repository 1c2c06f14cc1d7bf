"""Linear / Newton / explicit solvers restated from the reference (FP64, CPU).

Test infrastructure only (see oracle/__init__.py).
"""
from math import comb

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import assembly as asm


def conjugate_gradient(A, b, x0=None, tol=1e-3, max_iter=1500):
    """conjugate_gradient (_linear_solver.py:4-42); tol is on r.r / b.b (:37)."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r_r0 = b @ b
    r = b - A @ x
    p = r.copy()
    rsold = r @ r
    it = 0
    for it in range(max_iter):
        Ap = A @ p
        alpha = rsold / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        rsnew = r @ r
        if rsnew / r_r0 < tol:
            break
        p = r + rsnew / rsold * p
        rsold = rsnew
    return x, it + 1


def block_jacobi_inverse(A, grp):
    """Node-block Jacobi blocks (an addition to the reference's point Jacobi): for each row
    group g (rows grp[g] .. grp[g+1]-1) the inverse of the symmetric part of its diagonal
    block by cofactors, rounded to float32 as stored; a block that is not positive definite
    (leading minors) falls back to 1 / |A_ii|.  Returns [G, 3, 3] (zero outside nr x nr)."""
    A = sp.csr_matrix(A)
    out = np.zeros((len(grp) - 1, 3, 3))
    for g in range(len(grp) - 1):
        r0, nr = grp[g], grp[g + 1] - grp[g]
        B = np.eye(3)
        B[:nr, :nr] = A[r0:r0 + nr, r0:r0 + nr].toarray()
        B = 0.5 * (B + B.T)
        c = np.array([[B[1, 1] * B[2, 2] - B[1, 2] ** 2, B[0, 2] * B[1, 2] - B[0, 1] * B[2, 2],
                       B[0, 1] * B[1, 2] - B[0, 2] * B[1, 1]],
                      [0.0, B[0, 0] * B[2, 2] - B[0, 2] ** 2, B[0, 1] * B[0, 2] - B[0, 0] * B[1, 2]],
                      [0.0, 0.0, B[0, 0] * B[1, 1] - B[0, 1] ** 2]])
        c[1, 0], c[2, 0], c[2, 1] = c[0, 1], c[0, 2], c[1, 2]
        det = B[0, 0] * c[0, 0] + B[0, 1] * c[0, 1] + B[0, 2] * c[0, 2]
        if B[0, 0] > 0 and c[2, 2] > 0 and det > 0:
            inv = c / det
        else:
            d = np.abs(np.diag(B))
            inv = np.diag(np.where(d == 0, 1.0, 1.0 / np.where(d == 0, 1.0, d)))
        inv[nr:, :] = 0.0
        inv[:, nr:] = 0.0
        out[g] = inv.astype(np.float32).astype(np.float64)
    return out


def pcg_block_jacobi(A, b, grp, x0=None, tol=1e-5, max_iter=1500):
    """Preconditioned CG with M^-1 = blockdiag(block_jacobi_inverse): tol on r.z / b.(M^-1 b)."""
    Minv = block_jacobi_inverse(A, grp)
    rows = np.concatenate([np.arange(grp[g], grp[g] + 3) for g in range(len(grp) - 1)])
    n = b.size

    def apply(v):
        vv = np.zeros(rows.size)
        ok = rows < n
        vv[ok] = v[rows[ok]]
        zz = np.einsum('gab,gb->ga', Minv, vv.reshape(-1, 3)).reshape(-1)
        z = np.zeros(n)
        for g in range(len(grp) - 1):
            nr = grp[g + 1] - grp[g]
            z[grp[g]:grp[g] + nr] = zz[3 * g:3 * g + nr]
        return z
    x = np.zeros_like(b) if x0 is None else x0.copy()
    bz = b @ apply(b)
    r = b - A @ x
    z = apply(r)
    p = z.copy()
    rz = r @ z
    it = 0
    for it in range(max_iter):
        q = A @ p
        alpha = rz / (p @ q)
        x = x + alpha * p
        r = r - alpha * q
        z = apply(r)
        rz_new = r @ z
        if rz_new / bz < tol:
            break
        p = z + rz_new / rz * p
        rz = rz_new
    return x, it + 1


def jacobi_scaled_solve(K, rhs, x0, linear='cg', tol=1e-5, max_iter=1500):
    """_solve_linear_equation (static_implicit.py:297-350).

    diag = sqrt|K_ii| (0 -> 1); solve D^-1/2 K D^-1/2 y = D^-1/2 rhs; x = y / diag.
    ``linear='direct'`` stands in for the reference's Pardiso branch
    (scipy spsolve; pypardiso is not installed).
    """
    diag = np.sqrt(np.abs(K.diagonal()))
    diag[diag == 0] = 1.0
    Dinv = sp.diags(1.0 / diag)
    Ks = (Dinv @ K @ Dinv).tocsr()
    bs = rhs / diag
    if linear == 'direct':
        y = spla.spsolve(Ks.tocsc(), bs)
        its = 0
    else:
        y, its = conjugate_gradient(Ks, bs, x0 * diag, tol=tol, max_iter=max_iter)
    return y / diag, its


def _system(model):
    """(energy(GC), residual_tangent(GC), residual(GC), n_gc) of a model: the plain free-DOF
    system, or the Couple-reduced one when the model has reference points."""
    if model.get('couples') or model.get('rps'):
        from . import couple as cp
        return (lambda GC: cp.total_energy(model, GC), lambda GC: cp.assemble_reduced(model, GC),
                lambda GC: cp.assemble_force(model, GC), cp.n_gc(model))
    return (lambda GC: asm.total_energy(model, asm.gc_to_u(model, GC)),
            lambda GC: asm.assemble_reduced(model, asm.gc_to_u(model, GC)),
            lambda GC: asm.assemble_force(model, asm.gc_to_u(model, GC)), int(asm.free_mask(model).sum()))


def line_search(model, GC0, dGC, R, energy0, armijo_slack=0.0):
    """_line_search (static_implicit.py:98-177): Armijo c1=0.3, alpha halving to 1e-12.

    ``armijo_slack`` (absolute energy) is 0 in the reference.
    """
    energy_fn = _system(model)[0]
    alpha, c1 = 1.0, 0.3
    deltaE = float(dGC @ R)
    if deltaE > 0:
        dGC, deltaE = -dGC, -deltaE
    while True:
        GCnew = GC0 + alpha * dGC
        e_new = energy_fn(GCnew)
        if (not np.isfinite(e_new)) or e_new > energy0 + c1 * deltaE * alpha + armijo_slack:
            alpha *= 0.5
            if alpha < 1e-12:
                return 0.0, GC0.copy(), energy0
        else:
            return alpha, GCnew, e_new


def scale_loads(model, factor):
    """Copy of the model with pressure, traction and reference-point loads scaled (load increments)."""
    m = dict(model)
    m['pressures'] = [dict(p, p=p['p'] * factor) for p in model.get('pressures', [])]
    m['tractions'] = []
    for t in model.get('tractions', []):
        asm.prepare(model)
        m['tractions'].append(dict(t, t=tuple(np.asarray(t['t']) * factor), f=t['f'] * factor))
    if model.get('rps'):
        m['rps'] = [dict(rp, **{k: (None if rp.get(k) is None else np.asarray(rp[k], dtype=np.float64) * factor)
                                for k in ('force', 'moment')}) for rp in model['rps']]
    return m


def newton(model, GC, tol_error=1e-5, rel_tol=None, max_iter=200, linear='cg',
           cg_tol=1e-5, armijo_slack_rel=0.0, r_ref=None, step_tol=1e-12):
    """_solve_iteration (static_implicit.py:179-295).

    Reference convergence: (|dGC|inf < tol and |R|inf < tol) or |R|inf < 1e-6 (:290).
    With ``rel_tol`` the test is ||R||2 <= rel_tol * r_ref instead (r_ref = the
    load's residual norm at u = 0, default: first residual of this call), or a
    full step with max|dGC| <= step_tol max|GC| (correction at round-off).
    Returns (GC or False, iterations).
    """
    energy_fn, rk_fn = _system(model)[:2]
    energy = energy_fn(GC)
    dGC = np.zeros_like(GC)
    alpha, it, r0 = 0.0, 0, None
    while True:
        if it > max_iter:
            return False, it
        R, K = rk_fn(GC)
        it += 1
        nR = np.linalg.norm(R)
        if r0 is None:
            r0 = nR if r_ref is None else r_ref
        if rel_tol is not None and nR <= rel_tol * r0:
            return GC, it
        dGC, _ = jacobi_scaled_solve(K, -R, dGC, linear=linear, tol=cg_tol,
                                     max_iter=6000 if it % 20 == 0 else 1500)
        alpha, GCnew, e_new = line_search(model, GC, dGC, R, energy,
                                          armijo_slack=armijo_slack_rel * abs(energy))
        if alpha == 0 and np.abs(R).max() > tol_error and rel_tol is None:
            return False, it
        if alpha == 0:
            if rel_tol is not None:
                return False, it
            break
        GC = GCnew
        energy = energy_fn(GC)
        if rel_tol is not None and alpha == 1.0 and np.abs(dGC).max() <= step_tol * np.abs(GC).max():
            return GC, it
        if rel_tol is None and ((np.abs(dGC).max() < tol_error and np.abs(R).max() < tol_error)
                                or np.abs(R).max() < 1e-6):
            break
    return GC, it


def extrapolate(hist, order):
    """Equal-step Lagrange extrapolation of degree m = min(order, len(hist) - 1) through
    the states hist[0..m] (newest first) to the next level: sum_j (-1)^j C(m+1, j+1) hist[j]."""
    m = min(order, len(hist) - 1)
    return sum((-1.0) ** j * comb(m + 1, j + 1) * hist[j] for j in range(m + 1))


def auto_order(hist, max_order):
    """The order m <= max_order whose extrapolation from hist[1:] misses hist[0] least in the
    2-norm; one higher when that is the highest order hist can test."""
    top = min(max_order, len(hist) - 2)
    miss = [np.linalg.norm(extrapolate(hist[1:], m) - hist[0]) for m in range(top + 1)]
    best = int(np.argmin(miss))
    return best + 1 if best == top and top + 1 <= min(max_order, len(hist) - 1) else best


def static_solve(model, increments=1, predictor=True, GC0=None, orders=None, **kw):
    """Load-incremented static solve: Newton per increment, started from the equal-step
    Lagrange extrapolation of the previous converged states (``predictor`` True / 'linear':
    2 g1 - g2; 'quadratic': 3 g1 - 3 g2 + g3; 'cubic': 4 g1 - 6 g2 + 4 g3 - g4; an int:
    that order; lower orders while fewer states exist; 'auto': ``auto_order`` up to 6 per
    increment, replaced by g1 when the prediction's potential energy at the new load
    exceeds g1's; False: from g1).  ``orders`` (a list) receives the order used per increment."""
    order = predictor if isinstance(predictor, int) and not isinstance(predictor, bool) else \
        {False: 0, True: 1, 'linear': 1, 'quadratic': 2, 'cubic': 3, 'auto': 6}[predictor]
    keep = order + 2 if predictor == 'auto' else order + 1
    _, _, force_fn, n = _system(model)
    GC = np.zeros(n) if GC0 is None else np.asarray(GC0, dtype=np.float64).copy()
    total = 0
    kw.setdefault('r_ref', float(np.linalg.norm(force_fn(np.zeros(n)))))
    hist = [GC.copy()]                     # converged states, newest first
    for k in range(1, increments + 1):
        if order and k >= 2:
            m = auto_order(hist, order) if predictor == 'auto' else order
            GC = extrapolate(hist, m)
            if predictor == 'auto' and m > 0:
                # safeguard: the last equilibrium when its energy at the new load is lower
                e = _system(scale_loads(model, k / increments))[0]
                if not e(GC) <= e(hist[0]):
                    m, GC = 0, hist[0].copy()
            if orders is not None:
                orders.append(min(m, len(hist) - 1))
        GC, its = newton(scale_loads(model, k / increments), GC, **kw)
        total += its
        if GC is False:
            return False, total
        hist = [GC.copy()] + hist[:keep - 1]
    return GC, total


# --- explicit dynamics ---------------------------------------------------------

def lumped_mass(model):
    """Row-sum lumped mass on free DOFs (dynamic_explicit.py:147-166, C3base.py:348-384).

    The rows are summed over the free-DOF columns only: assemble_mass_matrix
    drops every entry with a constrained row or column (assembly.py:434-437)
    before get_lumped_mass_inv sums the rows."""
    from . import elements as el
    asm.prepare(model)
    N = model['nodes'].shape[0]
    M = np.zeros(3 * N)
    free = asm.free_mask(model)
    for g in model['groups']:
        vals = el.mass_values(g['N'], g['wdet'], g.get('rho', 1.0))
        r, c = el.coo_indices(g['conn'])
        keep = free[c]                      # reduced matrix keeps free x free entries only
        np.add.at(M, r[keep], vals[keep])
    m = M[free]
    m[np.abs(m) < 1e-12] = 1.0
    return m


def central_difference(residual, minv, GC0, GV0, dt, nsteps):
    """The central-difference loop of DynamicExplicitSolver.solve (dynamic_explicit.py:80-134)
    for a generic residual R(GC) (a = -M^-1 R).  Returns (GC, GV, GA) after nsteps."""
    GA = minv * (-residual(GC0))
    GVh = GV0 - 0.5 * dt * GA
    GC = np.array(GC0, dtype=float, copy=True)
    GV = np.array(GV0, dtype=float, copy=True)
    for _ in range(nsteps):
        GVh_new = GVh + dt * GA
        GC = GC + dt * GVh_new
        GA = minv * (-residual(GC))
        GV = 0.5 * (GVh_new + GVh)
        GVh = GVh_new
    return GC, GV, GA


def explicit_steps(model, GC0, GV0, dt, nsteps, extra_force=None):
    """Explicit dynamics of a model with the row-sum lumped mass.

    ``extra_force(u_nodes) -> [3N]`` adds an external force (e.g. contact).
    """
    free = asm.free_mask(model)
    minv = 1.0 / lumped_mass(model)

    def residual(GC):
        u = asm.gc_to_u(model, GC)
        R = asm.assemble_force(model, u)
        if extra_force is not None:
            R = R - extra_force(u)[free]
        return R

    return central_difference(residual, minv, GC0, GV0, dt, nsteps)


def consistent_mass(model):
    """Reduced consistent mass, free x free CSR (Element_3D.get_mass_matrix,
    C3base.py:348-384; Assembly.assemble_mass_matrix, assembly.py:421-443)."""
    from . import elements as el
    asm.prepare(model)
    N = model['nodes'].shape[0]
    free = asm.free_mask(model)
    gc = -np.ones(3 * N, dtype=np.int64)
    gc[free] = np.arange(int(free.sum()))
    rows, cols, vals = [], [], []
    for g in model['groups']:
        v = el.mass_values(g['N'], g['wdet'], g.get('rho', 1.0))
        r, c = el.coo_indices(g['conn'])
        keep = free[r] & free[c]
        rows.append(gc[r[keep]])
        cols.append(gc[c[keep]])
        vals.append(v[keep])
    n = int(free.sum())
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def newmark_steps(residual_tangent, energy, M, GC0, GV0, dt, nsteps, beta=0.25, gamma=0.5, tol_error=1e-10,
                  max_iter=200):
    """Newmark-beta implicit dynamics (DynamicImplicitSolver, dynamic_implicit.py:146-281, :339-472)
    for generic callables: residual_tangent(GC) -> (R, K) with R = dPi/dGC, energy(GC) -> Pi.

    Per step: predictor GC_pre = GC + dt GV + (1/2 - beta) dt^2 GA (:181 has the
    literal 0.25, the same at beta = 1/4); Newton with direct solves on the incremental functional
    Pi(GC) + (GC - GC_pre).M.(GC - GC_pre) / (2 beta dt^2) (:100-144), Armijo
    c1 = 0.3 with halving (:285-330), stop when max|dGC| and max|R| < tol_error
    (:468); then the Newmark velocity / acceleration update (:261-281).  The
    initial acceleration is taken from M GA0 = -R(GC0) (the equation of motion;
    the reference solves M GA0 = +R, :266-269).  Returns (GC, GV, GA) arrays of
    every step (nsteps + 1 rows)."""
    c = 1.0 / (beta * dt * dt)
    GC, GV = np.array(GC0, dtype=np.float64), np.array(GV0, dtype=np.float64)
    R0, _ = residual_tangent(GC)
    GA = spla.spsolve(M.tocsc(), -R0) if M.shape[0] > 1 else -R0 / M.toarray()[0, 0]
    out = [(GC.copy(), GV.copy(), np.atleast_1d(GA).copy())]
    for _ in range(nsteps):
        pre = GC + dt * GV + (0.5 - beta) * dt * dt * GA

        def inc_energy(x):
            d = x - pre
            return energy(x) + 0.5 * c * float(d @ (M @ d))

        x = GC.copy()
        e0 = inc_energy(x)
        for it in range(max_iter):
            R, K = residual_tangent(x)
            R = R + c * (M @ (x - pre))
            Ke = (K + c * M).tocsc()
            d = spla.spsolve(Ke, -R) if Ke.shape[0] > 1 else -R / Ke.toarray()[0, 0]
            d = np.atleast_1d(d)
            alpha, dE = 1.0, float(d @ R)
            if dE > 0:
                d, dE = -d, -dE
            while True:
                e1 = inc_energy(x + alpha * d)
                if np.isfinite(e1) and e1 <= e0 + 0.3 * dE * alpha:
                    break
                alpha *= 0.5
                if alpha < 1e-12:
                    alpha = 0.0
                    break
            if alpha == 0.0:
                break
            x, e0 = x + alpha * d, e1
            if np.abs(alpha * d).max() < tol_error and np.abs(R).max() < tol_error:
                break
        GV_new = gamma / (beta * dt) * (x - GC) + (1 - gamma / beta) * GV + dt * (1 - gamma / (2 * beta)) * GA
        GA = c * (x - GC) - GV / (beta * dt) - (1 / (2 * beta) - 1) * GA
        GC, GV = x, GV_new
        out.append((GC.copy(), GV.copy(), np.atleast_1d(GA).copy()))
    return out


def newmark(model, GC0, GV0, dt, nsteps, **kw):
    """newmark_steps on an oracle model (free-DOF vectors)."""
    M = consistent_mass(model)

    def rt(GC):
        return asm.assemble_reduced(model, asm.gc_to_u(model, GC))

    def en(GC):
        return asm.total_energy(model, asm.gc_to_u(model, GC))

    return newmark_steps(rt, en, M, GC0, GV0, dt, nsteps, **kw)
