"""CPU FP64 oracle for the FEA nonlinear solid-mechanics hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``fea_amd``) imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may use it, and only as the checker / reported CPU
baseline.  It is a plain numpy restatement of the reference algorithm
(sssnnn7119/FEA), each function citing the reference ``file:line`` it follows.

Pinning status (what independently checks each part; see tests/test_oracle_*.py):

* ``elements``  shape tables / quadrature / dN/dX / w*detJ: closed forms
  (partition of unity, sum of derivatives = 0, sum w*detJ = exact volume,
  affine patch test F = A).                                       -> pinned
* ``elements`` residual / tangent (reference einsum pipeline): translation
  invariance, homogeneous patch test, K symmetric, K = dR/du by central FD,
  single-element energy derivative by complex step.               -> pinned
* ``materials`` NeoHookean (reference closed form) and MooneyRivlin (new):
  P = dW/dF and C = dP^T/dF by complex-step differentiation of W / P,
  MR(C01=0) == NH(mu=2C10, kappa=2/D1), objectivity W(QF)=W(F).   -> pinned
* ``pattern`` CSR (crow, col) of the reduced free-DOF matrix: brute-force
  dense construction from connectivity + free mask.                 -> pinned
* ``surfaces`` T3/T6/Q4 tables and Gauss rules: N_a(node_b) = delta_ab,
  monomial exactness of the rules; pressure volume / dV / d2V and traction
  on closed off-origin box surfaces with non-affine facets: V = box volume
  (divergence theorem), sum w*detJ0 = area, traction resultant = t*A0 and
  its moment, zero net pressure force and moment, a third of the vector
  area on an open face, central differences of V and of the force
  (tests/test_oracle_surfaces.py).                                  -> pinned
* ``solvers`` CG: vs numpy.linalg.solve; Newton: residual at solution,
  path independence vs a direct-solve Newton.                        -> pinned
* ``explicit``: lumped mass totals = rho*V, 1-DOF oscillator closed form.
                                                                    -> pinned
* ``contact`` candidate pairs: brute force vs scipy cKDTree.query_ball_point
  semantics (distance <= r, ties included).                          -> pinned
"""
