"""GPU parity on the reference's own element-test meshes imported through
fea_amd.inp (tests/golden/inp/): the Neo-Hooke material of the file, the
assembly-level node set 'bottom' fixed and a dead traction on the surface
'head'; tangent / residual against the oracle (CSR bit-exact, values 1e-10)
and the static Newton solve against the oracle's direct-solve Newton
(1e-8 relative displacement)."""
import os

import numpy as np
import pytest
import torch

from fea_amd import inp as finp
from fea_amd import solver
from fea_amd.assemble import boundarys, loads
from oracle import assembly as oasm
from oracle import solvers as osol
from tests.helpers import FACE_NODES, nh, rel_err

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'inp')


def models(etype, traction=(0.0, 0.0, -0.002)):
    m = finp.FEA_INP().read_inp(os.path.join(GOLDEN, etype + '.inp.gz'))
    fe = finp.from_inp(m)
    part = fe.assembly._parts['Part-1']
    bottom = np.array(sorted(m.assemble.sets_nodes['Part-1-1.bottom']))
    part.set_nodes['bottom'] = bottom
    part.add_surface_set('head', m.assemble.surfaces['Part-1-1.head'])
    fe.assembly.add_boundary(boundarys.Boundary_Condition(instance_name='Part-1', set_nodes_name='bottom'))
    fe.assembly.add_load(loads.Traction('Part-1', 'head', traction), 'load-0')
    X = m.part['Part-1'].nodes.numpy()[:, 1:]
    conn = m.part['Part-1'].elems[etype][:, 1:]
    fixed = np.zeros(X.shape, dtype=bool)
    fixed[bottom] = True
    tri = np.concatenate([conn[ids][:, FACE_NODES[etype][face]] for ids, face in m.assemble.surfaces['Part-1-1.head']])
    stype = {3: 'T3', 6: 'T6', 4: 'Q4'}[tri.shape[1]]
    om = {'nodes': X, 'fixed': fixed, 'pressures': [], 'tractions': [{'stype': stype, 'conn': tri, 't': traction}],
          'groups': [{'etype': etype, 'conn': conn, 'material': nh(0.48, 5.0)}]}
    return fe, om


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8', 'C3D10'])
def test_imported_tangent_matches_oracle(gpu, etype):
    fe, om = models(etype)
    fe.assembly.initialize()
    asm = fe.assembly
    p = asm.plan
    rng = np.random.default_rng(3)
    GC = 0.05 * rng.standard_normal(p.n_free)
    R, Kidx, Kval = asm.assemble_Stiffness_Matrix(GC=torch.as_tensor(GC, device='cuda'))
    R_o, K_o = oasm.assemble_reduced(om, oasm.gc_to_u(om, GC))
    assert np.array_equal(Kidx[1].cpu().numpy(), K_o.indices)
    assert np.array_equal(np.diff(p.rowptr_np), np.diff(K_o.indptr))
    assert rel_err(Kval.cpu().numpy(), K_o.data) < 1e-10
    assert rel_err(R.cpu().numpy(), R_o) < 1e-10


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8'])
def test_imported_static_solve_matches_oracle(gpu, etype):
    fe, om = models(etype)
    fe.solver = solver.StaticImplicitSolver(precision='fp64', rel_tol=1e-10, load_steps=2, cg_tol=1e-20,
                                            cg_max_iter=20000, armijo_slack=1e-12)
    GC = fe.solve()
    assert GC is not False
    GC_o, _ = osol.static_solve(om, increments=2, rel_tol=1e-10, linear='direct', armijo_slack_rel=1e-12)
    assert rel_err(GC.cpu().numpy(), GC_o) < 1e-8
    assert np.abs(GC_o).max() > 1e-3                                  # the load deforms the block
