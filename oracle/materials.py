"""Hyperelastic materials in the reference's (s = P^T, C = dP^T/dF) convention.

Test infrastructure only (see oracle/__init__.py).  Works for real or complex
F (complex step differentiation in the tests).
"""
import numpy as np


def _inv_det(F):
    J = np.linalg.det(F)
    invF = np.linalg.inv(F)
    return J, invF


class NeoHookean:
    """NeoHookean (hyperelastic.py:6-84).

    W = mu/2 (I1bar - 3) + kappa/2 (J - 1)^2, I1bar = tr(F^T F) J^(-2/3)
    (hyperelastic.py:25-33).
    """

    def __init__(self, mu, kappa):
        self.mu = float(mu)
        self.kappa = float(kappa)

    def energy(self, F):
        J = np.linalg.det(F)
        I1 = (F ** 2).sum((-1, -2)) * J ** (-2.0 / 3.0)
        return self.mu / 2 * (I1 - 3) + self.kappa / 2 * (J - 1) ** 2

    def constitutive(self, F):
        """material_Constitutive_C3 (hyperelastic.py:35-84): returns s [..,3,3], C [..,3,3,3,3]."""
        J, invF = _inv_det(F)
        Jneg = J ** (-2.0 / 3.0)
        I1 = (F ** 2).sum((-1, -2)) * Jneg
        mu, kappa = self.mu, self.kappa
        J4 = J[..., None, None]
        Jn4 = Jneg[..., None, None]
        I14 = I1[..., None, None]
        muJneg = mu * Jn4
        FtMuJneg = np.swapaxes(F, -1, -2) * muJneg
        muI1invF = mu * I14 * invF
        kappaJinvF = kappa * J4 * invF
        s = FtMuJneg + (-1.0 / 3.0 * muI1invF + kappaJinvF * (J4 - 1))          # :67
        C = np.einsum('...ij,...lk->...ijkl',
                      -2.0 / 3.0 * FtMuJneg + kappaJinvF * (2 * J4 - 1) + 2.0 / 9.0 * muI1invF,
                      invF)                                                    # :69-72
        C = C + np.einsum('...ij,...kl->...ijkl', -2.0 / 3.0 * muJneg * invF, F)  # :74-75
        C = C + np.einsum('...ik,...lj->...ijkl',
                          (1.0 / 3.0 * muI1invF - kappaJinvF * (J4 - 1)), invF)  # :76-78
        for m in range(3):
            for n in range(3):
                C[..., m, n, n, m] += muJneg[..., 0, 0]                          # :80-82
        return s, C


class MooneyRivlin:
    """Mooney-Rivlin (absent in the reference; Abaqus form).

    W = C10 (I1bar - 3) + C01 (I2bar - 3) + (1/D1)(J - 1)^2,
    I2bar = J^(-4/3) (I1^2 - tr((F^T F)^2)) / 2.  With C01 = 0 it equals the
    reference NeoHookean with mu = 2 C10, kappa = 2 / D1 (inp.py:465-466 mapping).
    Same (s = P^T, C = dP^T/dF) convention as hyperelastic.py:67-82.
    """

    def __init__(self, C10, C01, D1):
        self.C10 = float(C10)
        self.C01 = float(C01)
        self.D1 = float(D1)
        self._nh = NeoHookean(2 * self.C10, 2.0 / self.D1)

    def energy(self, F):
        J = np.linalg.det(F)
        Cg = np.swapaxes(F, -1, -2) @ F
        I1 = np.trace(Cg, axis1=-2, axis2=-1)
        I2 = 0.5 * (I1 ** 2 - np.trace(Cg @ Cg, axis1=-2, axis2=-1))
        return (self.C10 * (I1 * J ** (-2.0 / 3.0) - 3) + self.C01 * (I2 * J ** (-4.0 / 3.0) - 3)
                + (J - 1) ** 2 / self.D1)

    def constitutive(self, F):
        s, C = self._nh.constitutive(F)
        if self.C01 == 0.0:
            return s, C
        c01 = self.C01
        J, invF = _inv_det(F)
        m = J ** (-4.0 / 3.0)
        Cg = np.swapaxes(F, -1, -2) @ F                  # F^T F
        b = F @ np.swapaxes(F, -1, -2)                   # F F^T
        I1 = np.trace(Cg, axis1=-2, axis2=-1)
        I2 = 0.5 * (I1 ** 2 - np.trace(Cg @ Cg, axis1=-2, axis2=-1))
        bF = b @ F
        m4, I14, I24 = m[..., None, None], I1[..., None, None], I2[..., None, None]
        invFT = np.swapaxes(invF, -1, -2)
        # P2 = C01 m [2 (I1 F - b F) - 4/3 I2 F^-T]
        P2 = c01 * m4 * (2.0 * (I14 * F - bF) - 4.0 / 3.0 * I24 * invFT)
        s = s + np.swapaxes(P2, -1, -2)
        # dP2_pq / dF_kl, assembled as D[..., p, q, k, l]
        eye = np.eye(3)
        G = I14 * F - bF
        D = (-8.0 / 3.0) * np.einsum('...lk,...pq->...pqkl', invF, G)
        D = D + 4.0 * np.einsum('...kl,...pq->...pqkl', F, F)
        D = D + 2.0 * I1[..., None, None, None, None] * np.einsum('pk,ql->pqkl', eye, eye)
        D = D - 2.0 * np.einsum('pk,...lq->...pqkl', eye, Cg)
        D = D - 2.0 * np.einsum('...pl,...kq->...pqkl', F, F)
        D = D - 2.0 * np.einsum('...pk,ql->...pqkl', b, eye)
        D = D + (16.0 / 9.0) * I2[..., None, None, None, None] * np.einsum('...lk,...qp->...pqkl', invF, invF)
        D = D - (8.0 / 3.0) * np.einsum('...kl,...qp->...pqkl', G, invF)
        D = D + (4.0 / 3.0) * I2[..., None, None, None, None] * np.einsum('...qk,...lp->...pqkl', invF, invF)
        D = c01 * m[..., None, None, None, None] * D
        # C_ijkl = dP_ji / dF_kl
        C = C + np.swapaxes(D, -3, -4)
        return s, C


def first_pk(material, F):
    s, _ = material.constitutive(F)
    return np.swapaxes(s, -1, -2)
