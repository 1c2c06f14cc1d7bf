"""Smoothed-aggregation multigrid V-cycle and the PCG it preconditions (numpy, float64).

Test infrastructure only (see oracle/__init__.py).  The multigrid preconditioner is an
addition of this build (the reference's CG is point-Jacobi scaled,
FEA/solver/static_implicit.py:314-321); this module restates the V-cycle the GPU applies
on a given hierarchy (levels as exported by fea_amd.solver.amg.host_hierarchy), so the
GPU's preconditioner and PCG can be checked against it:

* Chebyshev smoothing of D^-1 A on [rho / 30, 1.1 rho] with the three-term recurrence
  (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1), the same
  polynomial before and after the coarse correction;
* residual restriction r_c = P^T r, prolongation x += P x_c, a dense solve on the
  coarsest level;
* level 0 smooths with the current matrix and its node-block Jacobi inverses; the device
  reads a binary16 copy of that matrix there (``half_values``), and so may the oracle.
"""
import numpy as np
import scipy.sparse as sp


def half_values(v):
    """The binary16 copy of matrix values the device's V-cycle uses on level 0
    (fea_amg_half_values): with max |v| = f 2^e (f in [1/2, 1), taken in float32), each value
    times 2^k, k = 15 - e, rounded to binary16, returned in float64 and divided by 2^k
    again.  Pinned in tests/test_oracle_solvers.py by the binary16 range and rounding bound."""
    v = np.asarray(v, dtype=np.float64)
    m = float(np.abs(v.astype(np.float32)).max()) if v.size else 0.0
    if m == 0.0:
        return v.copy()
    k = 15 - int(np.frexp(np.float32(m))[1])
    return np.ldexp(np.ldexp(v, k).astype(np.float16).astype(np.float64), -k)


def cheb_coeffs(rho, deg):
    """(c1, c2) of each smoothing step: d = c1 d + c2 D^-1 r, x += d."""
    lmax, lmin = 1.1 * rho, rho / 30.0
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    rk = 1.0 / sigma
    c = [(0.0, 1.0 / theta)]
    for _ in range(1, deg):
        rn = 1.0 / (2.0 * sigma - rk)
        c.append((rn * rk, 2.0 * rn / delta))
        rk = rn
    return c


def chebyshev(A, Dinv, rho, b, x, deg):
    """deg smoothing steps from x (None: x = 0)."""
    d = None
    x = np.zeros_like(b) if x is None else x.copy()
    for c1, c2 in cheb_coeffs(rho, deg):
        w = Dinv @ (b - A @ x)
        d = c2 * w if d is None or c1 == 0.0 else c1 * d + c2 * w
        x = x + d
    return x


def group_block_inverse(A, grp):
    """Block-diagonal inverse of the node-group diagonal blocks (as a sparse matrix)."""
    A = sp.csr_matrix(A)
    blocks = []
    for g in range(len(grp) - 1):
        r0, r1 = grp[g], grp[g + 1]
        B = A[r0:r1, r0:r1].toarray()
        blocks.append(np.linalg.inv(0.5 * (B + B.T)))
    return sp.block_diag(blocks, format='csr')


def vcycle(levels, A0, D0inv, rho0, b, deg):
    """z = V(b): levels[l] has 'P' (all but the coarsest), 'A' and 'Dinv' / 'rho' for
    levels >= 1, 'dense_inv' on the coarsest; level 0 uses A0, D0inv, rho0."""
    def op(l):
        return A0 if l == 0 else levels[l]['A']

    def smoother(l):
        return (D0inv, rho0) if l == 0 else (levels[l]['Dinv'], levels[l]['rho'])

    def cycle(l, bl):
        if l == len(levels) - 1:
            return levels[l]['dense_inv'] @ bl
        A = op(l)
        Di, rh = smoother(l)
        x = chebyshev(A, Di, rh, bl, None, deg)
        P = levels[l]['P']
        x = x + P @ cycle(l + 1, P.T @ (bl - A @ x))
        return chebyshev(A, Di, rh, bl, x, deg)

    return cycle(0, b)


def pcg(A, b, Minv, x0=None, tol=1e-5, max_iter=1500):
    """PCG with the stopping test r.z / b.(M^-1 b) < tol of _linear_solver.py:37 (on the
    preconditioned residual); returns (x, iterations)."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - A @ x
    z = Minv(r)
    bz = float(b @ Minv(b))
    p = z.copy()
    rz = float(r @ z)
    for k in range(1, max_iter + 1):
        q = A @ p
        pq = float(p @ q)
        alpha = rz / pq if pq != 0.0 else 0.0
        x += alpha * p
        r -= alpha * q
        z = Minv(r)
        rz_new = float(r @ z)
        if rz_new == 0.0 or rz_new / bz < tol or k >= max_iter:
            return x, k
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_iter
