"""GPU parity of the hex8 (3D solid) path: the device element matrix
against the reference's sample fixture and the oracle, the matrix-free and
assembled operators, the solve of the reference's own hex8 sample
(sample/solid), the mixed-precision solve, the 3D filter, and a short 3D
SIMP loop against the oracle's loop."""
import math

import numpy as np
import pytest

from tests import golden_io as gi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import pansfem2_amd
    return pansfem2_amd


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_hex8_element_matrix(pa):
    from oracle import pyoracle as po
    ctx = pa.Context(0)
    ctx.grid3d(50, 5, 5, 50.0, 5.0, 5.0)
    Ke = ctx.element_stiffness(0, 210000.0, 0.3)
    # the reference's sample (sample/solid) has element 0 on the unit cube too
    assert gi.sha(Ke.reshape(-1)) == gi.digest("solid_sample", "ke_e0")
    xyz, conn = po.grid3d(7, 3, 4)
    ctx.grid3d(7, 3, 4)
    for e in (0, 17, 83):
        assert np.array_equal(ctx.element_stiffness(e, 2.5, 0.3), po.ke_h8(xyz[conn[e]], 2.5, 0.3))


def test_hex8_solid_sample_solve(pa):
    """The reference's hex8 sample: 50x5x5 beam, x=0 face clamped, four -100
    loads at the centre of the x=50 face, E = 210000, ScalingCG to 1e-10."""
    g = gi.load("solid_sample")
    ref_u = g["u_nodal"].reshape(-1, 3) if "u_nodal" in g else None
    nodes_csv = None
    ctx = pa.Context(0)
    nx, ny, nz = 50, 5, 5
    ctx.grid3d(nx, ny, nz, 50.0, 5.0, 5.0)
    face = np.arange((ny + 1) * (nz + 1), dtype=np.int32)
    ctx.set_dirichlet(np.repeat(face, 3), np.tile(np.arange(3, dtype=np.int32), len(face)))
    node = lambda i, j, k: (ny + 1) * (nz + 1) * i + (nz + 1) * j + k
    load_nodes = [node(50, 2, 2), node(50, 3, 2), node(50, 2, 3), node(50, 3, 3)]
    ctx.set_point_loads(load_nodes, [1] * 4, [-100.0] * 4)
    ctx.set_element_matrix(ctx.element_stiffness(0, 1.0, 0.3))
    ctx.set_moduli(np.full(nx * ny * nz, 210000.0))
    n2g = ctx.dofmap()
    # CSV node id = i + 51 j + 306 k (x fastest); map the reference's nodal u
    u_csv = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    free_csv = g["dofmap"].reshape(-1, 3)
    uref_free = g["u_free"]
    for d in range(3):
        m = free_csv[:, d] >= 0
        u_csv[m, d] = uref_free[free_csv[m, d]]
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    csv_id = (I + 51 * J + 306 * K).reshape(-1)
    u_expect = u_csv[csv_id]
    for op in (pa.PF_OP_EBE64, pa.PF_OP_CSR64):
        if op == pa.PF_OP_CSR64:
            ctx.assemble_csr()
        u, st = ctx.pcg(ctx.load(), op=op, eps=1e-10)
        assert st["status"] == 1
        un = np.zeros(n2g.size)
        un[n2g >= 0] = u[n2g[n2g >= 0]]
        assert rel(un.reshape(-1, 3), u_expect) < 1e-8, (op, rel(un.reshape(-1, 3), u_expect))
    assert abs(un.reshape(-1, 3)[:, 1].min() - (-1.48963)) < 1e-5


def test_hex8_csr_bit_exact_and_operators(pa):
    from oracle import pyoracle as po
    nx, ny, nz = 6, 4, 3
    ctx = pa.cantilever3d(nx, ny, nz)
    xyz, conn = po.grid3d(nx, ny, nz)
    rng = np.random.default_rng(7)
    rho = rng.uniform(1e-3, 1.0, nx * ny * nz)
    E = np.array([math.pow(r, 3.0) + 1e-9 * (1.0 - math.pow(r, 3.0)) for r in rho])
    ctx.set_moduli(E)
    ctx.assemble_csr()
    ip, ix, dv = ctx.csr()
    n2g = ctx.dofmap()
    ke = np.stack([po.ke_h8(xyz[c], e, 0.3) for c, e in zip(conn, E)])
    rip, rix, rdv, _ = po.assemble_csr(ctx.N, conn, n2g, ke, 3)
    assert np.array_equal(ip, rip) and np.array_equal(ix, rix) and np.array_equal(dv, rdv)
    x = rng.uniform(-1, 1, ctx.N)
    y_ref = po.spmv(rip, rix, rdv, x)
    assert np.array_equal(ctx.apply_K(x, pa.PF_OP_CSR64), y_ref)
    assert rel(ctx.apply_K(x, pa.PF_OP_EBE64), y_ref) < 1e-13
    assert rel(ctx.apply_K(x, pa.PF_OP_EBE32), y_ref) < 1e-5


@pytest.mark.parametrize("nx,ny,nz", [(9, 5, 3), (13, 20, 20), (6, 3, 90), (4, 2, 230)])
def test_hex8_pcg_plane_kernel_shapes(pa, nx, ny, nz):
    """The matrix-free hex8 PCG (one operator launch per iteration that marches
    over chunks of node planes, 256-node slabs of a plane per workgroup) against
    the PCG over the assembled CSR matrix: partial plane chunks (9x5x3),
    several slabs per plane (13x20x20), long planes that take the kernel's
    wider register set (nz = 90), and planes too long for it (nz = 230, the
    three-launch form).  FP64 to 1e-10, and the FP32 inner solve of the mixed
    solver to a 1e-8 true residual."""
    print(f"hex8 pcg {nx}x{ny}x{nz}", flush=True)
    ctx = pa.cantilever3d(nx, ny, nz)
    rng = np.random.default_rng(nx * 10000 + ny * 100 + nz)
    ctx.set_density(rng.uniform(1e-2, 1.0, nx * ny * nz))
    ctx.assemble_csr()
    b = ctx.load()
    xc, sc = ctx.pcg(b, op=pa.PF_OP_CSR64, eps=1e-10)
    xe, se = ctx.pcg(b, op=pa.PF_OP_EBE64, eps=1e-10)
    assert sc["status"] == 1 and se["status"] == 1
    assert rel(xe, xc) < 1e-8, rel(xe, xc)
    assert abs(se["iterations"] - sc["iterations"]) <= max(2, sc["iterations"] // 100)
    x32, s32 = ctx.pcg(b, op=pa.PF_OP_EBE32, eps=1e-8)
    assert s32["status"] == 1 and s32["true_residual"] <= 1.0000001e-8
    assert rel(x32, xc) < 1e-6
    ctx.close()


def test_hex8_mixed_precision_solve(pa):
    """FP32 matrix-free hex8 inner PCG inside FP64 iterative refinement: the
    solve ends at a 1e-8 FP64 true residual and agrees with the FP64
    Jacobi-PCG solution to what the residuals allow."""
    nx, ny, nz = 32, 16, 16
    ctx = pa.cantilever3d(nx, ny, nz)
    rng = np.random.default_rng(31)
    ctx.set_density(rng.uniform(0.2, 1.0, nx * ny * nz))
    b = ctx.load()
    x64, s64 = ctx.pcg(b, op=pa.PF_OP_EBE64, eps=1e-10)
    x32, s32 = ctx.pcg(b, op=pa.PF_OP_EBE32, eps=1e-8)
    print(f"hex8 mixed: {s32['iterations']} inner iterations, {s32['outer']} outer, "
          f"{s32['fallbacks']} fallbacks; fp64 {s64['iterations']}", flush=True)
    assert s64["status"] == 1 and s32["status"] == 1
    assert s32["true_residual"] <= 1.0000001e-8
    assert rel(x32, x64) < 1e-5
    assert abs(x32 @ b - x64 @ b) / abs(x64 @ b) < 1e-8
    ctx.close()


def test_hex8_filter_bit_exact(pa):
    from oracle import pyoracle as po
    nx, ny, nz = 6, 4, 3
    ctx = pa.cantilever3d(nx, ny, nz)
    xyz, conn = po.grid3d(nx, ny, nz)
    fp, fi, fw = po.filter_build_brute(po.centroids(conn, xyz), 1.5)
    assert np.diff(fp).max() == 19
    s = np.random.default_rng(3).uniform(0, 1, nx * ny * nz)
    assert np.array_equal(ctx.filter(s), po.filter_apply(fp, fi, fw, s))


def test_hex8_simp_loop_against_oracle(pa):
    from oracle import pyoracle as po
    nx, ny, nz = 8, 4, 4
    nodes, dofs, vals = pa.line_load_nodes3d(nx, ny, nz)
    ref = po.simp3d(nx, ny, nz, 5, list(zip(nodes.tolist(), dofs.tolist(), vals.tolist())), vf=0.3)
    ctx = pa.cantilever3d(nx, ny, nz)
    prm = pa.simp_default_params(vf=0.3)
    r = ctx.simp_run(5, prm, s0=0.3)
    f_ref = np.array([h[0] for h in ref["hist"]])
    assert np.max(np.abs(r["f"] - f_ref) / f_ref) < 1e-9
    assert np.max(np.abs(r["rho"] - ref["rho"])) < 1e-6
    assert abs(int(r["cg_iters"][0]) - int(ref["hist"][0][2])) <= 2


def test_hex8_simp_loop_multigrid_against_oracle(pa):
    """The same 5 OC iterations with every solve multigrid-preconditioned
    (PF_OP_MG, to the true residual 1e-10): the same compliance history and
    densities as the oracle's Jacobi-PCG loop."""
    from oracle import pyoracle as po
    nx, ny, nz = 8, 4, 4
    nodes, dofs, vals = pa.line_load_nodes3d(nx, ny, nz)
    ref = po.simp3d(nx, ny, nz, 5, list(zip(nodes.tolist(), dofs.tolist(), vals.tolist())), vf=0.3)
    ctx = pa.cantilever3d(nx, ny, nz)
    ctx.mg_setup(1)
    r = ctx.simp_run(5, pa.simp_default_params(vf=0.3, solver=pa.PF_OP_MG), s0=0.3)
    f_ref = np.array([h[0] for h in ref["hist"]])
    assert np.max(np.abs(r["f"] - f_ref) / f_ref) < 1e-9
    assert np.max(np.abs(r["rho"] - ref["rho"])) < 1e-6
    assert np.max(r["cg_iters"]) < 0.5 * int(ref["hist"][0][2])
