"""Pin the oracle's constitutive laws independently of their own formulas.

* P = dW/dF and C = dP^T/dF by complex-step differentiation (exact to ~1e-15),
* MR(C01=0) == NH(mu=2 C10, kappa=2/D1),
* objectivity W(QF) = W(F), P(QF) = Q P(F).
"""
import numpy as np
import pytest

from oracle import materials as om


def rand_F(rng, n=5, amp=0.3):
    F = np.eye(3)[None] + amp * rng.standard_normal((n, 3, 3))
    d = np.linalg.det(F)
    F[d < 0] *= -1.0
    return F


def cs_dW(mat, F, h=1e-30):
    P = np.zeros(F.shape)
    for k in range(3):
        for l in range(3):
            Fc = F.astype(complex)
            Fc[:, k, l] += 1j * h
            P[:, k, l] = mat.energy(Fc).imag / h
    return P


def cs_dPT(mat, F, h=1e-30):
    C = np.zeros(F.shape[:1] + (3, 3, 3, 3))
    for k in range(3):
        for l in range(3):
            Fc = F.astype(complex)
            Fc[:, k, l] += 1j * h
            s, _ = mat.constitutive(Fc)
            C[:, :, :, k, l] = s.imag / h
    return C


MATS = [om.NeoHookean(1.0, 100.0), om.NeoHookean(0.48, 5.0), om.MooneyRivlin(0.5, 0.1, 0.01),
        om.MooneyRivlin(0.3, 0.2, 0.5)]


@pytest.mark.parametrize('mat', MATS, ids=lambda m: type(m).__name__)
def test_stress_is_energy_gradient(mat):
    F = rand_F(np.random.default_rng(1))
    s, _ = mat.constitutive(F)
    P = np.swapaxes(s, -1, -2)
    Pcs = cs_dW(mat, F)
    assert np.abs(P - Pcs).max() <= 1e-12 * np.abs(Pcs).max()


@pytest.mark.parametrize('mat', MATS, ids=lambda m: type(m).__name__)
def test_tangent_is_stress_derivative(mat):
    F = rand_F(np.random.default_rng(2))
    _, C = mat.constitutive(F)
    Ccs = cs_dPT(mat, F)
    assert np.abs(C - Ccs).max() <= 1e-12 * np.abs(Ccs).max()


def test_tangent_major_symmetry():
    # C_ijkl = d2W/dF_ji dF_lk  => C_ijkl = C_lkji
    F = rand_F(np.random.default_rng(3))
    for mat in MATS:
        _, C = mat.constitutive(F)
        assert np.abs(C - np.transpose(C, (0, 4, 3, 2, 1))).max() <= 1e-12 * np.abs(C).max()


def test_mr_reduces_to_nh():
    F = rand_F(np.random.default_rng(4))
    mr = om.MooneyRivlin(0.7, 0.0, 0.04)
    nh = om.NeoHookean(1.4, 50.0)
    assert np.allclose(mr.energy(F), nh.energy(F), rtol=1e-14, atol=0)
    s1, C1 = mr.constitutive(F)
    s2, C2 = nh.constitutive(F)
    assert np.abs(s1 - s2).max() == 0.0 and np.abs(C1 - C2).max() == 0.0


def test_objectivity():
    rng = np.random.default_rng(5)
    F = rand_F(rng)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    for mat in MATS:
        assert np.allclose(mat.energy(q[None] @ F), mat.energy(F), rtol=1e-12)
        P = om.first_pk(mat, F)
        PQ = om.first_pk(mat, q[None] @ F)
        assert np.allclose(PQ, q[None] @ P, rtol=1e-11, atol=1e-11 * np.abs(P).max())


def test_stress_free_reference():
    for mat in MATS:
        s, _ = mat.constitutive(np.eye(3)[None])
        assert np.abs(s).max() < 1e-14


def test_small_strain_limit():
    # linearised NH at F=I: C_ijkl = mu (d_ik d_jl + d_il d_jk) + (kappa - 2mu/3) d_ij d_kl (in P^T form)
    mat = om.NeoHookean(1.3, 7.0)
    _, C = mat.constitutive(np.eye(3)[None])
    d = np.eye(3)
    mu, k = 1.3, 7.0
    # C_ijkl = dP_ji/dF_kl ; linear: P = mu (H + H^T) + (k - 2mu/3) tr H I
    ref = (mu * (np.einsum('jk,il->ijkl', d, d) + np.einsum('jl,ik->ijkl', d, d))
           + (k - 2 * mu / 3) * np.einsum('ji,kl->ijkl', d, d))
    assert np.abs(C[0] - ref).max() < 1e-13
