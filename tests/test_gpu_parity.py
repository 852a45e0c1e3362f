"""GPU parity of the HIP path against the reference's golden vectors and the
CPU oracle, called through the C ABI.

Bars: bit-exact for DOF numbering, CSR structure, CSR values, the CSR
product and the device element matrix; relative 1e-14 for the matrix-free
FP64 operator against the assembled one; the reference's own tolerances for
solves (displacement 1e-8, compliance 1e-9) and for the loop (density after
50 OC iterations 1e-6 max-abs).
"""
import math

import numpy as np
import pytest

from tests import golden_io as gi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import pansfem2_amd
    return pansfem2_amd


def host_E(rho, E1=1.0, E0=1e-9, p=3.0):
    # the driver's E (sample_optimize_density_oc.cpp:123), evaluated with the host libm
    return np.array([E1 * math.pow(r, p) + E0 * (1.0 - math.pow(r, p)) for r in rho])


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("case,nx,ny", [("simp_4x2_oc", 4, 2), ("simp_8x4_oc", 8, 4), ("simp_120x40_oc", 120, 40)])
def test_dofmap_csr_bit_exact(pa, case, nx, ny):
    g = gi.load(case)
    ctx = pa.cantilever2d(nx, ny)
    assert gi.sha(ctx.dofmap()) == gi.digest(case, "dofmap")
    assert gi.sha(ctx.element_stiffness(0, 1.0).reshape(-1)) == gi.digest(case, "ke_e0")
    ctx.set_moduli(host_E(g["rho0"]))
    ctx.assemble_csr()
    ip, ix, dv = ctx.csr()
    assert gi.sha(ip) == gi.digest(case, "csr_indptr")
    assert gi.sha(ix) == gi.digest(case, "csr_indices")
    assert gi.sha(dv) == gi.digest(case, "csr_data")
    assert np.array_equal(ctx.load(), g["F"])


def test_element_matrix_bit_exact(pa):
    ctx = pa.cantilever2d(120, 40)
    g = gi.manifest()["simp_120x40_oc"]["arrays"]["ke_e0"]["sha256"]
    assert gi.sha(ctx.element_stiffness(0, 1.0).reshape(-1)) == g


@pytest.mark.parametrize("case", ["operator_16x8", "operator_64x32"])
def test_operator_and_solve(pa, case):
    g = gi.load(case)
    m = gi.meta(case)
    ctx = pa.cantilever2d(m["nx"], m["ny"])
    ctx.set_moduli(host_E(g["rho"]))
    ctx.assemble_csr()
    ip, ix, dv = ctx.csr()
    assert gi.sha(dv) == gi.digest(case, "csr_data")
    # CSR<T>::operator* bit for bit
    y_csr = ctx.apply_K(g["x"], pa.PF_OP_CSR64)
    assert np.array_equal(y_csr, g["y"])
    # matrix-free FP64 operator (device moduli from rho) against the assembled one
    ctx.set_density(g["rho"])
    y_ebe = ctx.apply_K(g["x"], pa.PF_OP_EBE64)
    assert rel(y_ebe, g["y"]) < 1e-14
    y32 = ctx.apply_K(g["x"], pa.PF_OP_EBE32)
    assert rel(y32, g["y"]) < 1e-5
    # ScalingCG to 1e-10: displacement within 1e-8 of the reference's
    for op in (pa.PF_OP_EBE64, pa.PF_OP_CSR64):
        if op == pa.PF_OP_CSR64:
            ctx.set_moduli(host_E(g["rho"]))
            ctx.assemble_csr()
        u, st = ctx.pcg(g["F"], op=op, eps=1e-10)
        assert st["status"] == 1
        assert rel(u, g["u"]) < 1e-8, (op, rel(u, g["u"]), st)
        assert st["true_residual"] < 1e-9


def test_filter_bit_exact(pa):
    from oracle import pyoracle as po
    nx, ny = 120, 40
    ctx = pa.cantilever2d(nx, ny)
    xy, conn = po.grid2d(nx, ny)
    fp, fi, fw = po.filter_build(nx, ny, po.centroids(conn, xy))
    rng = np.random.default_rng(11)
    s = rng.uniform(0, 1, nx * ny)
    assert np.array_equal(ctx.filter(s, pa.PF_FILTER_VARIABLES), po.filter_apply(fp, fi, fw, s))
    d = rng.standard_normal(nx * ny)
    assert np.array_equal(ctx.filter(d, pa.PF_FILTER_SENSITIVITIES), po.filter_sens(fp, fi, fw, d))


def test_compliance_sensitivity(pa):
    case = "simp_120x40_oc"
    g = gi.load(case)
    ctx = pa.cantilever2d(120, 40)
    ctx.set_density(g["rho0"])
    f, dfdrho = ctx.compliance_sens(g["u0_free"])
    f0 = gi_f0(case)
    # compliance tolerance 1e-9 relative; u^T K u carries ~1e-11 of rounding
    # from the rigid-body part of u_e in either summation order
    assert abs(f - f0) / f0 < 1e-9
    assert np.max(np.abs(dfdrho - g["dfdrho0"])) < 1e-9 * np.max(np.abs(g["dfdrho0"]))


def gi_f0(case):
    from oracle import pyoracle as po
    return float(po.simp2d(120, 40, 1)["f"][0])


def test_oc_update_matches_oracle(pa):
    from oracle import pyoracle as po
    case = "simp_120x40_oc"
    g = gi.load(case)
    nx, ny = 120, 40
    ctx = pa.cantilever2d(nx, ny)
    xy, conn = po.grid2d(nx, ny)
    fp, fi, fw = po.filter_build(nx, ny, po.centroids(conn, xy))
    s0 = np.full(nx * ny, 0.5)
    dgds = po.filter_sens(fp, fi, fw, np.full(nx * ny, 1.0 / (0.5 * nx * ny)))
    x_ref, lam_ref, dec = po.oc_update(s0, g["dfds0"], dgds, fp, fi, fw)
    x, lam, steps = ctx.oc_update(s0, g["dfds0"], dgds)
    assert steps == len(dec)
    assert lam == lam_ref
    assert np.max(np.abs(x - x_ref)) < 1e-12


def test_simp_loop_c1_against_reference(pa):
    """50 OC iterations at 120x40: compliance history and final density."""
    case = "simp_120x40_oc"
    g = gi.load(case)
    ctx = pa.cantilever2d(120, 40)
    r = ctx.simp_run(50)
    from oracle import pyoracle as po
    ref = po.simp2d(120, 40, 50)
    assert gi.sha(ref["f"]) == gi.digest(case, "hist_f")
    assert np.max(np.abs(r["f"] - ref["f"]) / ref["f"]) < 1e-9
    assert np.max(np.abs(r["rho"] - g["rho_final"])) < 1e-6
    assert np.max(np.abs(r["s"] - g["s_final"])) < 1e-6
    # the device sums dot products in a different (fixed) order than the
    # reference's serial loops, so CG stops within a few iterations of it
    assert np.all(np.abs(r["cg_iters"] - ref["cg_iters"]) <= np.maximum(3, 0.005 * ref["cg_iters"]))


def test_operator_symmetry_and_residual_large(pa):
    """1024x512: size-independent properties (x^T K y = y^T K x, PCG true residual)."""
    nx, ny = 1024, 512
    ctx = pa.cantilever2d(nx, ny)
    rng = np.random.default_rng(20261201)
    rho = rng.uniform(1e-3, 1.0, nx * ny)
    ctx.set_density(rho)
    x = rng.uniform(-1, 1, ctx.N)
    y = rng.uniform(-1, 1, ctx.N)
    Kx = ctx.apply_K(x)
    Ky = ctx.apply_K(y)
    assert abs(x @ Ky - y @ Kx) / abs(x @ Ky) < 1e-12
    assert x @ Kx > 0
    ctx.set_moduli(host_E(rho))
    ctx.assemble_csr()
    assert ctx.nnz == 18874360
    Kc = ctx.apply_K(x, pa.PF_OP_CSR64)
    ctx.set_density(rho)
    # element matrices of the assembled path carry the grid's coordinate rounding (up to
    # ~1e-16 relative per entry); with cancellation in K x this stays below 1e-13
    assert rel(Kx, Kc) < 1e-13


def test_mma_cantilever_against_reference(pa):
    """src/Optimize/Solver/test_MMA.cpp through the device MMA (n = 5 as a 5x1 grid)."""
    g = gi.load("mma_cantilever")["traj"].reshape(-1, 7)
    ctx = pa.Context(0)
    ctx.grid2d(5, 1)
    ctx.mma_init(m=1, a0=1.0, a=0.0, c=1000.0, d=1.0, xmin=1.0, xmax=10.0, raa0=1e-5, albefa=0.1, move=0.5,
                 asyinit=0.5, asydecr=0.7, asyincr=1.2, epsvalue=1e-5)
    C1, cc = 0.0624, np.array([61.0, 37.0, 19.0, 7.0, 1.0])
    s = np.full(5, 5.0)
    rows = []
    for k in range(len(g)):
        obj = C1 * (s[0] + s[1] + s[2] + s[3] + s[4])
        con = sum(c * math.pow(v, -3.0) for c, v in zip(cc, s)) - 1.0
        rows.append([obj, con] + list(s))
        dcon = np.array([-3.0 * c * math.pow(v, -4.0) for c, v in zip(cc, s)])
        s, steps = ctx.mma_update(s, obj, np.full(5, C1), con, dcon)
    rows = np.array(rows)
    # n = 5 sums are evaluated in a different order than the reference's serial
    # loops; the trajectory agrees to rounding
    assert np.max(np.abs(rows - g) / np.maximum(np.abs(g), 1.0)) < 1e-10


def test_simp_mma_loop_c1_against_reference(pa):
    """50 MMA iterations at 120x40 (sample_optimize_density_mma.cpp loop)."""
    case = "simp_120x40_mma"
    g = gi.load(case)
    ctx = pa.cantilever2d(120, 40)
    r = ctx.simp_run(50, pa.simp_default_params(use_mma=1))
    from oracle import pyoracle as po
    ref = po.simp2d(120, 40, 50, po.simp_params(use_mma=1))
    assert gi.sha(ref["f"]) == gi.digest(case, "hist_f")
    assert np.max(np.abs(r["f"] - ref["f"]) / ref["f"]) < 1e-9
    assert np.max(np.abs(r["s"] - g["s_final"])) < 1e-6
    assert np.max(np.abs(r["rho"] - g["rho_final"])) < 1e-6


def test_facade_entry_points_against_reference(pa):
    """The entry points the C++ facade uses: CSR from caller-given element
    matrices (bit-exact CSR), the density filter over caller-given lists
    (bit-exact filtered densities) and one OC candidate per lambda."""
    case = "simp_120x40_oc"
    g = gi.load(case)
    nx, ny = 120, 40
    ctx = pa.cantilever2d(nx, ny)
    from oracle import pyoracle as po
    # element matrices at the first iteration's moduli, computed on the host
    # in the reference's arithmetic (the oracle's restatement)
    xy, conn = po.grid2d(nx, ny)
    E = host_E(g["rho0"])
    ke_all = np.stack([po.ke_q4(xy[conn[e]], E[e], 0.3, 1.0) for e in range(nx * ny)])
    ctx.assemble_csr_elements(ke_all)
    ip, ix, dv = ctx.csr()
    assert gi.sha(ip) == gi.digest(case, "csr_indptr")
    assert gi.sha(ix) == gi.digest(case, "csr_indices")
    assert gi.sha(dv) == gi.digest(case, "csr_data")
    # filter lists
    ctx.filter_set_lists(g["filter_ptr"], g["filter_idx"], g["filter_w"])
    s0 = np.full(nx * ny, 0.5)
    assert np.array_equal(ctx.filter(s0), g["rho0"])
    dfds = ctx.filter(g["dfdrho0"], mode=pa.PF_FILTER_SENSITIVITIES)
    assert np.array_equal(dfds, g["dfds0"])
    # OC candidate at the first lambda of the bisection (OC.h:84-91)
    lam = 0.5 * (0.0 + 1e9)
    xp = ctx.oc_candidate(s0, g["dfds0"], g["dgds0"], lam)
    base = -g["dfds0"] / (g["dgds0"] * lam)
    want = np.clip(np.sqrt(base) * s0, np.maximum(0.0, 0.8 * s0), np.minimum(1.0, 1.2 * s0))
    assert np.array_equal(xp, want)


def test_heaviside_entry_point_against_oracle(pa):
    """pf_heaviside over the reference's filter lists: projected variables
    and chain-rule sensitivities (HeavisideFilter.h:62-99).  The device tanh
    and glibc's may differ in the last bit, so the bar is 1e-14 relative."""
    from oracle import pyoracle as po
    g = gi.load("simp_120x40_oc")
    nx, ny = 120, 40
    ctx = pa.cantilever2d(nx, ny)
    ctx.filter_set_lists(g["filter_ptr"], g["filter_idx"], g["filter_w"])
    rng = np.random.default_rng(8)
    s = rng.uniform(0.0, 1.0, nx * ny)
    d = rng.uniform(-1.0, 0.0, nx * ny)
    fp, fi, fw = g["filter_ptr"], g["filter_idx"], g["filter_w"]
    for beta in (1.0, 2.0, 16.0):
        r = ctx.heaviside(beta, s)
        assert np.max(np.abs(r - po.heaviside_vars(fp, fi, fw, beta, s))) < 1e-14
        ds = ctx.heaviside(beta, s, d)
        want = po.heaviside_sens(fp, fi, fw, beta, s, d)
        assert np.max(np.abs(ds - want)) < 1e-14 * np.max(np.abs(want)) * 10
    # the grid rule gives the same values as the reference's lists
    ctx.filter_setup(1.5)
    assert np.max(np.abs(ctx.heaviside(2.0, s) - po.heaviside_vars(fp, fi, fw, 2.0, s))) < 1e-14


def test_simp_heaviside_loop_c1_against_reference(pa):
    """50 OC iterations at 120x40 with HeavisideFilter and the beta
    continuation (beta 1, then 2 from iteration 40) against the reference."""
    case = "simp_120x40_ocH"
    g = gi.load(case)
    ctx = pa.cantilever2d(120, 40)
    r = ctx.simp_run(50, pa.simp_default_params(heaviside=1))
    from oracle import pyoracle as po
    ref = po.simp2d(120, 40, 50, po.simp_params(heaviside=1))
    assert gi.sha(ref["f"]) == gi.digest(case, "hist_f")
    assert np.max(np.abs(r["f"] - ref["f"]) / ref["f"]) < 1e-9
    assert np.max(np.abs(r["rho"] - g["rho_final"])) < 1e-6
    assert np.max(np.abs(r["s"] - g["s_final"])) < 1e-6
    # continuation across calls: 40 iterations, then 10 more from iteration 40
    r1 = ctx.simp_run(40, pa.simp_default_params(heaviside=1))
    r2 = ctx.simp_run(10, pa.simp_default_params(heaviside=1, iter_offset=40), s0=r1["s"])
    assert np.max(np.abs(r2["s"] - g["s_final"])) < 1e-6


def test_sensitivity_filters_against_reference(pa):
    """SensitivityFilter and SensitivityFilter2 on the device (reference lists
    and the grid rule) against the reference's own classes, bit for bit."""
    from oracle import pyoracle as po
    g = gi.load("sensfilter_24x12")
    nx, ny = 24, 12
    ctx = pa.cantilever2d(nx, ny)
    xy, conn = po.grid2d(nx, ny)
    fp, fi, fw = po.filter_build(nx, ny, po.centroids(conn, xy))
    for lists in (False, True):
        if lists:
            ctx.filter_set_lists(fp, fi, fw)
        assert np.array_equal(ctx.sensitivity_filter(g["s"], g["dfds"], 1), g["out1"])
        assert np.array_equal(ctx.sensitivity_filter(g["s"], g["dfds"], 2), g["out2"])


def test_simp_conlin_loop_c1_against_reference(pa):
    """50 CONLIN iterations at 120x40 (the CONLIN sample's constants) against
    the reference's own CONLIN."""
    case = "simp_120x40_conlin"
    g = gi.load(case)
    ctx = pa.cantilever2d(120, 40)
    r = ctx.simp_run(50, pa.simp_default_params(use_mma=2))
    from oracle import pyoracle as po
    ref = po.simp2d(120, 40, 50, po.simp_params(use_mma=2))
    assert gi.sha(ref["f"]) == gi.digest(case, "hist_f")
    assert np.max(np.abs(r["f"] - ref["f"]) / ref["f"]) < 1e-9
    assert np.max(np.abs(r["s"] - g["s_final"])) < 1e-6
    assert np.max(np.abs(r["rho"] - g["rho_final"])) < 1e-6
