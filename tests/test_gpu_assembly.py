"""GPU parity of the global assembly (Assembly.assemble_Stiffness_Matrix etc.)
against the oracle's restatement of assembly.py:283-415 on the same seeded mesh:
CSR (indptr, indices) bit-exact, residual and tangent within 1e-10 relative in
FP64, energy within 1e-12 relative."""
import numpy as np
import pytest
import torch

from fea_amd import workloads
from oracle import assembly as oasm
from tests.helpers import mr, nh, oracle_model, rel_err

pytestmark = pytest.mark.gpu


def cube(n=3, quadratic=False, p=0.2):
    fe, mesh = workloads.cube_pressure(n, pressure=p, quadratic=quadratic, jitter=0.05, seed=2)
    fe.assembly.initialize()
    om = oracle_model(mesh, nh(1.0, 100.0), 'z0', pressure=p, load_set='z1')
    return fe, om


def beam(nx=6, ny=2, nz=2):
    fe, mesh = workloads.beam_traction(nx, ny, nz, traction=(0.0, 0.0, -0.01))
    fe.assembly.initialize()
    om = oracle_model(mesh, mr(0.5, 0.1, 0.01), 'x0', traction=(0.0, 0.0, -0.01), load_set='x1')
    return fe, om


CASES = {'cube_tet4': lambda: cube(), 'cube_tet10': lambda: cube(2, True), 'beam_hex8': lambda: beam()}


@pytest.mark.parametrize('case', list(CASES))
def test_assemble_stiffness_matches_oracle(gpu, case):
    fe, om = CASES[case]()
    asm = fe.assembly
    rng = np.random.default_rng(4)
    GC = torch.as_tensor(0.02 * rng.standard_normal(asm.GC.shape[0]), device='cuda')
    R, Kidx, Kval = asm.assemble_Stiffness_Matrix(GC=GC)
    U = oasm.gc_to_u(om, GC.cpu().numpy())
    R_o, K_o = oasm.assemble_reduced(om, U)
    rows = np.repeat(np.arange(K_o.shape[0]), np.diff(K_o.indptr))
    assert np.array_equal(Kidx[0].cpu().numpy(), rows)
    assert np.array_equal(Kidx[1].cpu().numpy(), K_o.indices)
    assert rel_err(Kval.cpu().numpy(), K_o.data) < 1e-10
    assert rel_err(R.cpu().numpy(), R_o) < 1e-10
    Rf = asm.assemble_force(GC=GC)
    assert rel_err(Rf.cpu().numpy(), R_o) < 1e-10
    E = float(asm._total_Potential_Energy(GC=GC))
    E_o = oasm.total_energy(om, U)
    assert abs(E - E_o) <= 1e-12 * max(1.0, abs(E_o))


def test_gc_rgc_roundtrip(gpu):
    fe, om = cube()
    asm = fe.assembly
    GC = torch.arange(asm.GC.shape[0], dtype=torch.float64, device='cuda')
    RGC = asm._GC2RGC(GC)
    assert torch.equal(asm._RGC2GC(RGC), GC)
    free = oasm.free_mask(om)
    assert int(free.sum()) == asm.GC.shape[0]
    # reference initial-guess quirk: last row of the last instance block = 1e-5 (assembly.py:113)
    assert float(asm.RGC[0][-1, 0]) == 1e-5


@pytest.mark.parametrize('etype', ['C3D4', 'C3D8'])
def test_pressure_on_face_through_origin(gpu, etype):
    """A loaded face in a plane through the origin makes [x, x_xi, x_eta] singular,
    where the reference's det * inverse form (pressure.py:112) divides by zero;
    the kernels' cofactor form stays finite and equals the autograd derivatives
    of the enclosed-volume functional V (an oracle with no inverse)."""
    from fea_amd import groupops, meshgen
    from oracle import autodiff as oad
    from oracle import surfaces as osurf
    from tests.helpers import facets
    mesh = meshgen.kuhn_tet_block(3, 2, 2) if etype == 'C3D4' else meshgen.hex_block(3, 2, 2)
    conn, stype = facets(mesh, 'y0')
    X = mesh['nodes']
    n = conn.shape[1]
    p = 0.3
    Xd = torch.as_tensor(X, device='cuda').reshape(-1).contiguous()
    Ud = torch.zeros_like(Xd)
    Fv, Kv = groupops.surface_pressure_per_facet(n, torch.as_tensor(conn), Xd, Ud, p, False)
    Fv, Kv = Fv.cpu().numpy(), Kv.cpu().numpy()
    assert np.all(np.isfinite(Fv)) and np.all(np.isfinite(Kv))
    N = X.shape[0]
    F = np.zeros(3 * N)
    np.add.at(F, osurf.pressure_force_idx(conn), Fv)
    K = np.zeros((3 * N, 3 * N))
    idx = osurf.pressure_stiff_idx(conn)
    np.add.at(K, (idx[0], idx[1]), Kv)
    x = torch.as_tensor(X, dtype=torch.float64).requires_grad_(True)
    V = oad.pressure_volume(stype, conn, x)
    g = torch.autograd.grad(V, x)[0].numpy().reshape(-1)
    H = torch.autograd.functional.hessian(lambda y: oad.pressure_volume(stype, conn, y.view(-1, 3)),
                                          torch.as_tensor(X, dtype=torch.float64).reshape(-1)).numpy()
    assert np.abs(F - (-p) * g).max() <= 1e-13 * max(1.0, np.abs(g).max())
    assert np.abs(K - (-p) * H).max() <= 1e-13 * max(1.0, np.abs(H).max())
    assert np.abs(g).max() > 0 and np.abs(H).max() > 0


def _partial_cube():
    from fea_amd import boundarys
    fe, mesh = workloads.cube_pressure(3, pressure=0.2, jitter=0.05, seed=2)
    fe.assembly.add_boundary(boundarys.Boundary_Condition('block', 'x0', indexDoF=[0]))
    fe.assembly.add_boundary(boundarys.Boundary_Condition('block', 'y1', indexDoF=[0, 1]))
    fe.assembly.initialize()
    return fe, None


@pytest.mark.parametrize('case', ['cube_tet4', 'cube_tet10', 'beam_hex8', 'partial'])
def test_direct_assembly_matches_staged(gpu, case):
    """Element stiffness added into the CSR colour by colour (fea_elem_tangent_direct)
    against the staged blocks summed by fea_gather_csr, on the same plan and state: the
    same sums in another order (float64 within 1e-13; float32 values within their
    rounding), the residual bit-identical; with pressure facets staged beside it and with
    1- and 2-row node groups."""
    from fea_amd import device
    from fea_amd._native import F32, F64
    device.DIRECT_ASSEMBLY = True
    try:
        fe, _ = _partial_cube() if case == 'partial' else CASES[case]()
    finally:
        device.DIRECT_ASSEMBLY = False
    p = fe.assembly.plan
    assert p.direct_assembly
    for g in p.groups:
        if g.kind == 'elem':
            conn = g.conn_np
            order = g.order.cpu().numpy()
            for c in range(len(g.color_ptr) - 1):
                nodes = conn[order[g.color_ptr[c]:g.color_ptr[c + 1]]].reshape(-1)
                assert np.unique(nodes).size == nodes.size
    U = p.gc_to_u(torch.as_tensor(0.02 * np.random.default_rng(6).standard_normal(p.n_free), device='cuda'))
    for dt, vt in ((F64, F64), (F64, F32), (F32, F32)):
        p.direct_assembly = True
        R1, K1 = p.tangent(U, 1.0, dt, val_dtype=vt)
        R1, K1 = R1.clone(), K1.double().clone()
        p.direct_assembly = False
        R2, K2 = p.tangent(U, 1.0, dt, val_dtype=vt)
        K2 = K2.double()
        p.direct_assembly = True
        assert torch.equal(R1, R2)
        tol = 1e-13 if vt == F64 else 3e-6
        err = float((K1 - K2).abs().max() / K2.abs().max())
        print('%s %s/%s: direct vs staged %.2e' % (case, dt, vt, err))
        assert err <= tol
