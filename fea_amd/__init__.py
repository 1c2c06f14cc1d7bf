"""fea_amd — MI355X-native (gfx950) nonlinear solid-mechanics hot path with the
Python API of sssnnn7119/FEA (FEA/__init__.py): Part / Instance / Assembly,
elements (C3D4, C3D8, C3D10), materials (NeoHookean, MooneyRivlin), loads
(Pressure, Traction, Concentrate_Force, Moment), boundary conditions, reference
points with Couple constraints, the static / explicit / Newmark solvers and INP
ingestion (FEA.inp / FEA.from_inp).

All numerical work runs in the HIP library fea_amd/_lib/libfea_amd.so (C ABI in
include/fea_amd.h); importing works without a GPU, calling the hot path does not.
"""
from .controller import FEAController
from .assemble import Part, Instance, Assembly, ReferencePoint
from .assemble import materials, elements, loads, boundarys, constraints, surfaces
from . import solver, meshgen
from .inp import FEA_INP, from_inp

__all__ = ['FEAController', 'Part', 'Instance', 'Assembly', 'ReferencePoint', 'materials', 'elements', 'loads',
           'boundarys', 'constraints', 'surfaces', 'solver', 'meshgen', 'FEA_INP', 'from_inp']
