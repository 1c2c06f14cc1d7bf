"""Benchmark / test models built through the public API.

* ``cube_pressure``   unit cube of Kuhn-split C3D4 (or C3D10), NeoHookean, bottom
                      face z=0 fixed, follower pressure on the top face z=1
                      (the 10x10x10 / mu=1 / kappa=100 / p=0.2 / jitter 5%
                      static workload and its smaller test versions).
* ``beam_traction``   4x1x1 cantilever of C3D8 bricks, Mooney-Rivlin
                      C10=0.5 C01=0.1 D1=0.01, clamped at x=0, uniform dead
                      traction on x=4 (the 128x32x32 mixed-precision workload).
* ``inp_couple``      an INP model with the reference's pressure-test set-up:
                      pressure on surface_1_All, surface_0_Bottom fixed,
                      surface_0_Head coupled to a reference point.
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


def two_blocks_contact(n=40, *, mu=1.0, kappa=50.0, rho=1.0, penalty=1e3, velocity=0.1, gap=None,
                       jitter=0.0, seed=0, upper_width=1.0, offset=(0.0, 0.0)):
    """Two blocks of n^3 Kuhn-split C3D4 cells, a unit cube and above it, by
    ``gap`` (default a quarter cell), an upper_width x upper_width x 1 block, centred and
    then shifted by ``offset`` in x and y, moving down at ``velocity``; the lower block's
    base z=0 is fixed.  Node-to-surface penalty contact: slave nodes the upper block's
    bottom face, master facets the lower block's top face.

    With the default offset and width the slave nodes sit exactly above master vertices and
    edges, and stay there (to round-off) while the blocks first meet: which facet takes such
    a node is then decided by round-off.  A trajectory compared with another implementation
    step by step needs an offset that keeps the slave nodes off the master edges.

    Returns (controller, info) with info = {'meshes': (lower, upper),
    'GV0': initial velocity in free-DOF order (after fe.initialize()), ...}."""
    h = 1.0 / n
    gap = 0.25 * h if gap is None else gap
    lower = meshgen.kuhn_tet_block(n, n, n, jitter=jitter, seed=seed)
    w = float(upper_width)
    upper = meshgen.kuhn_tet_block(n, n, n, size=(w, w, 1.0),
                                   origin=(0.5 - 0.5 * w + offset[0], 0.5 - 0.5 * w + offset[1], 1.0 + gap),
                                   jitter=jitter, seed=seed + 1)
    mat = materials.NeoHookean(mu, kappa)
    asm = Assembly()
    for name, mesh in (('lower', lower), ('upper', upper)):
        part = Part(torch.as_tensor(mesh['nodes'], dtype=torch.float64))
        for k, v in mesh['node_sets'].items():
            part.set_nodes[k] = np.asarray(v)
        for k, v in mesh['face_sets'].items():
            part.add_surface_set(k, v)
        E = mesh['conn'].shape[0]
        el = elements.initialize_element('C3D4', torch.arange(E), torch.as_tensor(mesh['conn']), part)
        el.set_materials(mat)
        el.set_density(rho)
        part.add_element(el, 'element-0')
        asm.add_part(part, name)
        asm.add_instance(Instance(part), name)
    asm.add_boundary(boundarys.Boundary_Condition(instance_name='lower', set_nodes_name='z0'))
    asm.add_load(loads.NodeToSurfaceContact('upper', 'z0', 'lower', 'z1', penalty=penalty), 'contact')
    fe = FEAController()
    fe.assembly = asm
    info = {'meshes': (lower, upper), 'velocity': velocity, 'gap': gap}
    return fe, info


def initial_velocity(fe, info):
    """Free-DOF initial velocity of ``two_blocks_contact``: -velocity along z on the upper block."""
    p = fe.assembly.plan
    nl = info['meshes'][0]['nodes'].shape[0]
    V = torch.zeros(p.n_nodes, 3, dtype=torch.float64, device=p.device)
    V[nl:, 2] = -info['velocity']
    return p.u_to_gc(V.reshape(-1).contiguous())


def void_block_pressure(m=10, k=5, ns=5, *, mu=1.0, kappa=100.0, pressure=0.05, amp=0.02, seed=0,
                        quadratic=True):
    """2x1x1 block of C3D10 (default 48000 elements) around a perturbed spherical
    void of radius 0.3, NeoHookean, face x=0 fixed, follower pressure inside
    the void (the shape-sensitivity workload)."""
    mesh = meshgen.block_with_void(m, k, ns, amp=amp, seed=seed, quadratic=quadratic)
    mat = materials.NeoHookean(mu, kappa)
    return _controller_from_mesh(mesh, mat, fixed_set='x0',
                                 load=lambda nm: loads.Pressure(nm, 'void', pressure)), mesh


def inp_couple(path, *, pressure=0.06, rp=(0.0, 0.0, 80.0), instance='final_model', contact_surfaces=()):
    """The reference's Couple set-up (tests/pressure_test/test_benchmark.py:24-41,
    tests/gradient_test/test_grad.py:27-52) on the INP mesh at ``path``: follower pressure on
    surface_1_All, surface_0_Bottom fixed, the surface_0_Head node set coupled to a reference
    point at ``rp``, and a ContactSelf load on each of ``contact_surfaces``."""
    from .assemble import ReferencePoint, constraints
    from .inp import FEA_INP, from_inp
    inp = FEA_INP().read_inp(path)
    fe = from_inp(inp)
    a = fe.assembly
    a.add_load(loads.Pressure(instance_name=instance, surface_set='surface_1_All', pressure=pressure),
               name='pressure-1')
    a.add_boundary(boundarys.Boundary_Condition(instance_name=instance, set_nodes_name='surface_0_Bottom'))
    for s in contact_surfaces:
        a.add_load(loads.ContactSelf(instance_name=instance, surface_name=s))
    name = a.add_reference_point(ReferencePoint(list(rp)))
    a.add_constraint(constraints.Couple(instance_name=instance, set_nodes_name='surface_0_Head', rp_name=name))
    return fe, inp
