"""Hyperelastic materials (reference: FEA/assemble/elements/materials).

Type ids follow the reference factory (materials/__init__.py:8-21):
1 = NeoHookean(mu, kappa); 2 = LinearElastic (not part of this hot path);
3 = MooneyRivlin(C10, C01, D1) (new, absent in the reference).

Every material reduces to the kernel parameter triple {w1, w2, kv} of
W = w1 (I1bar - 3) + w2 (I2bar - 3) + kv/2 (J - 1)^2 evaluated on the GPU.
"""
import numpy as np
import torch


def _scalar(v):
    if isinstance(v, torch.Tensor):
        if v.numel() != 1:
            raise NotImplementedError('per-element material parameters are not supported; '
                                      'pass scalars (the reference factory also reads row 0 only)')
        return float(v.reshape(-1)[0].item())
    return float(np.asarray(v).reshape(-1)[0])


class Materials_Base:
    type: int = 0

    def kernel_params(self) -> np.ndarray:
        raise NotImplementedError


class NeoHookean(Materials_Base):
    """W = mu/2 (I1bar - 3) + kappa/2 (J - 1)^2 (hyperelastic.py:6-33)."""

    def __init__(self, mu, kappa) -> None:
        self.type = 1
        self.mu = torch.tensor(_scalar(mu), dtype=torch.float64)
        self.kappa = torch.tensor(_scalar(kappa), dtype=torch.float64)

    def kernel_params(self):
        return np.array([0.5 * float(self.mu), 0.0, float(self.kappa)], dtype=np.float64)


class MooneyRivlin(Materials_Base):
    """W = C10 (I1bar - 3) + C01 (I2bar - 3) + (J - 1)^2 / D1 (Abaqus form; new).

    With C01 = 0 it is NeoHookean(mu = 2 C10, kappa = 2 / D1), the mapping the
    reference's INP reader applies to *Hyperelastic (inp.py:465-466).
    """

    def __init__(self, C10, C01, D1) -> None:
        self.type = 3
        self.C10 = torch.tensor(_scalar(C10), dtype=torch.float64)
        self.C01 = torch.tensor(_scalar(C01), dtype=torch.float64)
        self.D1 = torch.tensor(_scalar(D1), dtype=torch.float64)

    def kernel_params(self):
        return np.array([float(self.C10), float(self.C01), 2.0 / float(self.D1)], dtype=np.float64)


def initialize_materials(materials_type: int, materials_params: torch.Tensor):
    """Factory of materials/__init__.py:8-21 (+ type 3 = Mooney-Rivlin)."""
    p = torch.as_tensor(materials_params)
    if p.dim() == 1:
        p = p.unsqueeze(0)
    if materials_type == 1:
        return NeoHookean(p[0, 0], p[0, 1])
    if materials_type == 3:
        return MooneyRivlin(p[0, 0], p[0, 1], p[0, 2])
    if materials_type == 2:
        raise NotImplementedError('LinearElastic is outside the accelerated hot path')
    raise Exception('Unknown material type')


__all__ = ['initialize_materials', 'Materials_Base', 'NeoHookean', 'MooneyRivlin']
