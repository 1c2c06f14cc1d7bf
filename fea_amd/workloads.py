"""Benchmark / test models built through the public API.

* ``cube_pressure``   unit cube of Kuhn-split C3D4 (or C3D10), NeoHookean, bottom
                      face z=0 fixed, follower pressure on the top face z=1
                      (the 10x10x10 / mu=1 / kappa=100 / p=0.2 / jitter 5%
                      static workload and its smaller test versions).
* ``beam_traction``   4x1x1 cantilever of C3D8 bricks, Mooney-Rivlin
                      C10=0.5 C01=0.1 D1=0.01, clamped at x=0, uniform dead
                      traction on x=4 (the 128x32x32 mixed-precision workload).
"""
import numpy as np
import torch

from . import meshgen
from .assemble import Assembly, Instance, Part, boundarys, elements, loads, materials
from .controller import FEAController


def _controller_from_mesh(mesh, material, *, fixed_set, load=None, name='block', rho=1.0):
    part = Part(torch.as_tensor(mesh['nodes'], dtype=torch.float64))
    for k, v in mesh['node_sets'].items():
        part.set_nodes[k] = np.asarray(v)
    for k, v in mesh['face_sets'].items():
        part.add_surface_set(k, v)
    E = mesh['conn'].shape[0]
    el = elements.initialize_element(mesh['etype'], torch.arange(E), torch.as_tensor(mesh['conn']), part)
    el.set_materials(material)
    el.set_density(rho)
    part.add_element(el, 'element-0')
    asm = Assembly()
    asm.add_part(part, name)
    asm.add_instance(Instance(part), name)
    asm.add_boundary(boundarys.Boundary_Condition(instance_name=name, set_nodes_name=fixed_set))
    if load is not None:
        asm.add_load(load(name), 'load-0')
    fe = FEAController()
    fe.assembly = asm
    return fe


def cube_pressure(n=10, *, mu=1.0, kappa=100.0, pressure=0.2, jitter=0.05, seed=0, quadratic=False):
    mesh = meshgen.kuhn_tet_block(n, n, n, jitter=jitter, seed=seed)
    if quadratic:
        mesh = meshgen.tet10_from_tet4(mesh)
    mat = materials.NeoHookean(mu, kappa)
    return _controller_from_mesh(mesh, mat, fixed_set='z0',
                                 load=lambda nm: loads.Pressure(nm, 'z1', pressure)), mesh


def beam_traction(nx=128, ny=32, nz=32, *, C10=0.5, C01=0.1, D1=0.01, traction=(0.0, 0.0, -0.01),
                  length=4.0, jitter=0.0, seed=0):
    mesh = meshgen.hex_block(nx, ny, nz, size=(length, 1.0, 1.0), jitter=jitter, seed=seed)
    mat = materials.MooneyRivlin(C10, C01, D1)
    return _controller_from_mesh(mesh, mat, fixed_set='x0',
                                 load=lambda nm: loads.Traction(nm, 'x1', traction)), mesh
