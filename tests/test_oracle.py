"""Pins of the CPU oracle (oracle/pf_oracle.c).

Two kinds of check:
  * bit-exact agreement with the golden vectors written by the reference
    itself (tests/golden/, produced by oracle/_ref/ref_harness, which is
    compiled from the reference headers);
  * closed forms and invariants that a plausible restatement mistake (a
    dropped term, a wrong sign or index, a transposed operand) would break:
    exact symbolic element matrices, rigid-body null spaces, closed-form DOF
    numbering and sparsity, dense solves, filter stencil weights, the
    closed-form optimum of the reference's MMA cantilever test.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import golden_io as gi


# --------------------------------------------------------------- helpers ---
def simp_setup(nx, ny):
    xy, conn = po.grid2d(nx, ny)
    fn, fd = po.left_clamp_fixed(nx, ny)
    N, n2g = po.dofmap(xy.shape[0], 2, fn, fd)
    return xy, conn, N, n2g


def element_matrices(xy, conn, E):
    return np.stack([po.ke_q4(xy[c], e) for c, e in zip(conn, E)])


def simp_E(rho, E1=1.0, E0=1e-9, p=3.0):
    return E1 * np.power(rho, p) + E0 * (1.0 - np.power(rho, p))


# ------------------------------------------------ golden: reference bits ---
@pytest.mark.parametrize("case,nx,ny", [("simp_4x2_oc", 4, 2), ("simp_8x4_oc", 8, 4), ("simp_120x40_oc", 120, 40)])
def test_iteration0_against_reference(case, nx, ny):
    g = gi.load(case)
    xy, conn, N, n2g = simp_setup(nx, ny)
    assert gi.sha(n2g) == gi.digest(case, "dofmap")
    rho0 = g["rho0"]
    Ke = element_matrices(xy, conn, [math.pow(r, 3.0) * 1.0 + 1e-9 * (1.0 - math.pow(r, 3.0)) for r in rho0])
    indptr, indices, data, F = po.assemble_csr(N, conn, n2g, Ke, 2)
    F[n2g[2 * (ny + 1) * nx + 1]] += -1.0
    assert gi.sha(indptr) == gi.digest(case, "csr_indptr")
    assert gi.sha(indices) == gi.digest(case, "csr_indices")
    assert gi.sha(data) == gi.digest(case, "csr_data")
    assert np.array_equal(F, g["F"])
    assert gi.sha(po.csr_diag(indptr, indices, data)) == gi.digest(case, "diag0")
    u, it = po.scaling_cg(indptr, indices, data, F)
    assert it > 0
    assert np.array_equal(u, g["u0_free"])
    cg = po.centroids(conn, xy)
    fp, fi, fw = po.filter_build(nx, ny, cg)
    assert gi.sha(fp) == gi.digest(case, "filter_ptr")
    assert gi.sha(fi) == gi.digest(case, "filter_idx")
    assert gi.sha(fw) == gi.digest(case, "filter_w")
    assert np.array_equal(po.filter_sens(fp, fi, fw, g["dfdrho0"]), g["dfds0"])


def optimizer_of(case):
    return 2 if "conlin" in case else (1 if "mma" in case else 0)


@pytest.mark.parametrize("case,nx,ny,iters", [("simp_4x2_oc", 4, 2, 3), ("simp_8x4_oc", 8, 4, 5),
                                              ("simp_16x8_mma", 16, 8, 5), ("simp_16x8_ocH", 16, 8, 5),
                                              ("simp_16x8_conlin", 16, 8, 5)])
def test_simp_loop_small_against_reference(case, nx, ny, iters):
    g = gi.load(case)
    prm = po.simp_params(use_mma=optimizer_of(case), heaviside=1 if case.endswith("H") else 0)
    r = po.simp2d(nx, ny, iters, prm)
    assert gi.sha(r["f"]) == gi.digest(case, "hist_f")
    assert gi.sha(r["g"]) == gi.digest(case, "hist_g")
    if "oc" in case:
        assert gi.sha(r["lam"]) == gi.digest(case, "hist_lambda")
    assert np.array_equal(r["s"], g["s_final"])
    assert np.array_equal(r["rho"], g["rho_final"])
    assert np.array_equal(r["u0"], g["u0_free"])


@pytest.mark.slow
@pytest.mark.parametrize("case", ["simp_120x40_oc", "simp_120x40_mma", "simp_120x40_ocH", "simp_120x40_conlin"])
def test_simp_loop_c1_against_reference(case):
    """All 50 OC (MMA, CONLIN) iterations of the 120x40 workload, bit for bit
    (ocH: HeavisideFilter with beta continuation, beta 1 then 2 at k = 40)."""
    g = gi.load(case)
    prm = po.simp_params(use_mma=optimizer_of(case), heaviside=1 if case.endswith("H") else 0)
    r = po.simp2d(120, 40, 50, prm)
    assert gi.sha(r["f"]) == gi.digest(case, "hist_f")
    assert gi.sha(r["g"]) == gi.digest(case, "hist_g")
    if "oc" in case:
        assert gi.sha(r["lam"]) == gi.digest(case, "hist_lambda")
    assert np.array_equal(r["s"], g["s_final"])
    assert np.array_equal(r["rho"], g["rho_final"])


@pytest.mark.parametrize("case", ["operator_16x8", "operator_64x32"])
def test_operator_against_reference(case):
    g = gi.load(case)
    m = gi.meta(case)
    nx, ny = m["nx"], m["ny"]
    xy, conn, N, n2g = simp_setup(nx, ny)
    Ke = element_matrices(xy, conn, [1.0 * math.pow(r, 3.0) + 1e-9 * (1.0 - math.pow(r, 3.0)) for r in g["rho"]])
    indptr, indices, data, F = po.assemble_csr(N, conn, n2g, Ke, 2)
    assert gi.sha(indptr) == gi.digest(case, "csr_indptr")
    assert gi.sha(indices) == gi.digest(case, "csr_indices")
    assert gi.sha(data) == gi.digest(case, "csr_data")
    assert np.array_equal(po.spmv(indptr, indices, data, g["x"]), g["y"])
    F[n2g[2 * (ny + 1) * nx + 1]] += -1.0
    u, it = po.scaling_cg(indptr, indices, data, F)
    assert np.array_equal(u, g["u"])


def test_solid_sample_against_reference():
    """Hex8 path on the reference's own sample mesh (sample/solid/*.csv)."""
    g = gi.load("solid_sample")
    nodes = g["nodes"].reshape(-1, 3)
    conn = g["conn"].reshape(-1, 8)
    n2g = g["dofmap"]
    N = int(n2g.max()) + 1
    assert np.array_equal(po.ke_h8(nodes[conn[0]], 210000.0, 0.3).reshape(-1), g["ke_e0"])
    Ke = np.stack([po.ke_h8(nodes[c], 210000.0, 0.3) for c in conn])
    indptr, indices, data, F0 = po.assemble_csr(N, conn, n2g, Ke, 3)
    assert gi.sha(indptr) == gi.digest("solid_sample", "csr_indptr")
    assert gi.sha(indices) == gi.digest("solid_sample", "csr_indices")
    assert gi.sha(data) == gi.digest("solid_sample", "csr_data")
    u, it = po.scaling_cg(indptr, indices, data, g["F"])
    assert np.array_equal(u, g["u_free"])
    # the reference's shipped result (sample/solid/result_linear.vtk) quotes
    # min u_y = -1.48963 at 6 significant digits
    assert abs(u.reshape(-1)[n2g.reshape(-1, 3)[:, 1][n2g.reshape(-1, 3)[:, 1] >= 0]].min() - (-1.48963)) < 1e-5


def test_mma_cantilever_against_reference():
    """src/Optimize/Solver/test_MMA.cpp driven through the oracle MMA."""
    g = gi.load("mma_cantilever")["traj"].reshape(-1, 7)
    C1, cc = 0.0624, np.array([61.0, 37.0, 19.0, 7.0, 1.0])
    mma = po.MMA(5, 1, 1.0, [0.0], [1000.0], [1.0], [1.0] * 5, [10.0] * 5)
    mma.set_parameters(1e-5, 0.1, 0.5, 0.5, 0.7, 1.2, 1e-5)
    s = np.full(5, 5.0)
    rows = []
    for k in range(10):
        obj = C1 * (s[0] + s[1] + s[2] + s[3] + s[4])
        con = 61.0 * math.pow(s[0], -3.0) + 37.0 * math.pow(s[1], -3.0) + 19.0 * math.pow(s[2], -3.0) + \
            7.0 * math.pow(s[3], -3.0) + 1.0 * math.pow(s[4], -3.0) - 1.0
        dcon = np.array([-3.0 * c * math.pow(v, -4.0) for c, v in zip(cc, s)])
        rows.append([obj, con] + list(s))
        if mma.is_convergence(obj):
            break
        s = mma.update(s, obj, np.full(5, C1), np.array([con]), dcon[None, :])
    assert np.array_equal(np.array(rows), g)
    # closed form (KKT): x_j = k c_j^(1/4), k = (sum c_j^(1/4))^(1/3); f* = C1 (sum c_j^(1/4))^(4/3)
    q = cc ** 0.25
    xstar = q * q.sum() ** (1.0 / 3.0)
    assert np.allclose(g[-1, 2:], xstar, rtol=2e-3)
    assert abs(g[-1, 0] - C1 * q.sum() ** (4.0 / 3.0)) < 2e-3


# ------------------------------------------------ closed forms/invariants ---
def exact_ke_q4(nu):
    import sympy as sp
    x, y = sp.symbols("x y")
    N = [(1 - x) * (1 - y), x * (1 - y), x * y, (1 - x) * y]  # unit square, CCW from (0,0)
    nu = sp.Rational(nu).limit_denominator(1000)
    D = sp.Matrix([[1 - nu, nu, 0], [nu, 1 - nu, 0], [0, 0, (1 - 2 * nu) / 2]]) / ((1 - 2 * nu) * (1 + nu))
    B = sp.zeros(3, 8)
    for n in range(4):
        B[0, 2 * n] = sp.diff(N[n], x)
        B[1, 2 * n + 1] = sp.diff(N[n], y)
        B[2, 2 * n] = sp.diff(N[n], y)
        B[2, 2 * n + 1] = sp.diff(N[n], x)
    K = (B.T * D * B).applyfunc(lambda e: sp.integrate(sp.integrate(e, (x, 0, 1)), (y, 0, 1)))
    return np.array(K.evalf(30).tolist(), dtype=np.float64)


def test_ke_q4_exact_and_invariants():
    X = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    Ke = po.ke_q4(X, 1.0, 0.3, 1.0)
    assert np.allclose(Ke, exact_ke_q4(0.3), rtol=0, atol=1e-14)
    assert np.allclose(Ke, Ke.T, atol=1e-15)
    ev = np.linalg.eigvalsh(Ke)
    assert np.sum(np.abs(ev) < 1e-12) == 3 and ev.min() > -1e-12
    tx = np.tile([1.0, 0.0], 4)
    ty = np.tile([0.0, 1.0], 4)
    rot = np.array([[-p[1], p[0]] for p in X]).reshape(-1)
    for v in (tx, ty, rot):
        assert np.abs(Ke @ v).max() < 1e-14
    # linear in E and t; a rescaled element of size h keeps the same matrix (2D)
    assert np.allclose(po.ke_q4(X, 7.0, 0.3, 2.0), 14.0 * Ke, rtol=1e-14)
    assert np.allclose(po.ke_q4(2.5 * X, 1.0, 0.3, 1.0), Ke, rtol=1e-13, atol=1e-15)


def test_ke_h8_invariants():
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    Ke = po.ke_h8(X, 1.0, 0.3)
    assert np.allclose(Ke, Ke.T, atol=1e-15)
    ev = np.linalg.eigvalsh(Ke)
    assert np.sum(np.abs(ev) < 1e-12) == 6 and ev.min() > -1e-12
    for d in range(3):
        t = np.zeros(24)
        t[d::3] = 1.0
        assert np.abs(Ke @ t).max() < 1e-14
    # rotation about z: u = (-y, x, 0)
    rz = np.zeros(24)
    rz[0::3] = -X[:, 1]
    rz[1::3] = X[:, 0]
    assert np.abs(Ke @ rz).max() < 1e-14
    # uniaxial strain eps_xx = 1: strain energy = C11/2 * volume, C11 = E(1-nu)/((1+nu)(1-2nu))
    ux = np.zeros(24)
    ux[0::3] = X[:, 0]
    assert abs(0.5 * ux @ Ke @ ux - 0.5 * 0.7 / (1.3 * 0.4)) < 1e-13
    # doubling the element size doubles the 3D stiffness
    assert np.allclose(po.ke_h8(2.0 * X, 1.0, 0.3), 2.0 * Ke, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("nx,ny", [(4, 2), (37, 19), (120, 40)])
def test_dofmap_and_sparsity_closed_form(nx, ny):
    xy, conn, N, n2g = simp_setup(nx, ny)
    nn = (nx + 1) * (ny + 1)
    assert N == 2 * nx * (ny + 1)
    n = np.arange(nn)
    expect = np.where(n[:, None] >= ny + 1, 2 * (n[:, None] - (ny + 1)) + np.arange(2)[None, :], -1).reshape(-1)
    assert np.array_equal(n2g, expect)
    Ke = np.ones((nx * ny, 8, 8))
    indptr, indices, data, _ = po.assemble_csr(N, conn, n2g, Ke, 2)
    for row in [0, 1, N // 2, N // 2 + 1, N - 2, N - 1, 2 * (ny + 1) + 4]:
        node = row // 2 + (ny + 1)
        i, j = divmod(node, ny + 1)
        cols = sorted(2 * ((ny + 1) * i2 + j2 - (ny + 1)) + d for i2 in range(max(1, i - 1), min(nx, i + 1) + 1)
                      for j2 in range(max(0, j - 1), min(ny, j + 1) + 1) for d in (0, 1))
        assert list(indices[indptr[row]:indptr[row + 1]]) == cols
        # all-ones element matrices: each entry counts the elements shared by the two nodes
        assert data[indptr[row]:indptr[row + 1]].min() >= 1.0


def test_scaling_cg_dense_solve():
    nx, ny = 12, 6
    xy, conn, N, n2g = simp_setup(nx, ny)
    rng = np.random.default_rng(1)
    rho = rng.uniform(1e-3, 1.0, nx * ny)
    Ke = element_matrices(xy, conn, simp_E(rho))
    indptr, indices, data, _ = po.assemble_csr(N, conn, n2g, Ke, 2)
    A = np.zeros((N, N))
    for r in range(N):
        A[r, indices[indptr[r]:indptr[r + 1]]] = data[indptr[r]:indptr[r + 1]]
    b = rng.standard_normal(N)
    x, it = po.scaling_cg(indptr, indices, data, b, eps=1e-13)
    assert it > 0
    assert np.allclose(A, A.T)
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-12
    # itrmax exhaustion returns the last iterate and -1 (CG.h:451-452)
    x2, it2 = po.scaling_cg(indptr, indices, data, b, itrmax=3, eps=1e-13)
    assert it2 == -1 and np.linalg.norm(x2) > 0


def test_uniform_density_scales_solution():
    """K(rho = c) = E(c) K(1): u scales by 1/E(c), compliance by 1/E(c)."""
    nx, ny = 16, 8
    xy, conn, N, n2g = simp_setup(nx, ny)
    F = np.zeros(N)
    F[n2g[2 * (ny + 1) * nx + 1]] = -1.0
    sols = []
    for E in (1.0, 0.125):
        indptr, indices, data, _ = po.assemble_csr(N, conn, n2g, element_matrices(xy, conn, [E] * (nx * ny)), 2)
        sols.append(po.scaling_cg(indptr, indices, data, F, eps=1e-14)[0])
    assert np.allclose(sols[1], sols[0] / 0.125, rtol=1e-9, atol=1e-9 * np.abs(sols[1]).max())


def test_filter_stencil_closed_form():
    nx, ny = 60, 40
    xy, conn = po.grid2d(nx, ny)
    cg = po.centroids(conn, xy)
    fp, fi, fw = po.filter_build(nx, ny, cg, 1.5, brute=False)
    fp2, fi2, fw2 = po.filter_build(nx, ny, cg, 1.5, brute=True)
    assert np.array_equal(fp, fp2) and np.array_equal(fi, fi2) and np.array_equal(fw, fw2)
    counts = np.diff(fp)
    assert counts.max() == 9 and counts.min() == 4
    expect = {0.0: 1.0, 1.0: 1.0 / 3.0, math.sqrt(2.0): (1.5 - math.sqrt(2.0)) / 1.5}
    for e in [0, 5, ny * 7 + 3, nx * ny - 1]:
        for k in range(fp[e], fp[e + 1]):
            d = float(np.hypot(*(cg[e] - cg[fi[k]])))
            key = min(expect, key=lambda v: abs(v - d))
            assert abs(fw[k] - expect[key]) < 1e-14
    s = np.full(nx * ny, 0.37)
    assert np.allclose(po.filter_apply(fp, fi, fw, s), 0.37, rtol=1e-15)
    assert np.allclose(po.filter_sens(fp, fi, fw, s), 0.37, rtol=1e-15)
    # interior rows are the exact adjoint; boundary rows are row-normalised (DensityFilter.h:60-71)
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(nx * ny), rng.standard_normal(nx * ny)
    Wsum = np.add.reduceat(fw, fp[:-1])
    adj = np.zeros(nx * ny)
    np.add.at(adj, fi, np.repeat(a / Wsum, counts) * fw)
    ours = po.filter_sens(fp, fi, fw, a)
    interior = (counts == 9)
    nb_full = np.array([np.all(counts[fi[fp[i]:fp[i + 1]]] == 9) for i in range(nx * ny)])
    assert np.allclose(ours[interior & nb_full], adj[interior & nb_full], rtol=1e-13)
    assert not np.allclose(ours, adj)


def test_oc_uniform_case():
    """Uniform sensitivities keep x uniform; bisection lands on the volume target."""
    nx, ny = 20, 10
    xy, conn = po.grid2d(nx, ny)
    fp, fi, fw = po.filter_build(nx, ny, po.centroids(conn, xy))
    x0 = np.full(nx * ny, 0.5)
    x, lam, dec = po.oc_update(x0, np.full(nx * ny, -2.0), np.full(nx * ny, 1.0 / (0.5 * nx * ny)), fp, fi, fw)
    assert np.all(x == x[0])
    # x+ = x sqrt(-dfdx / (dgdx lam)) = 0.5 exactly at lam* = 2 * 0.5 * ne; the bracket is within 1e-3
    lam_star = 2.0 * 0.5 * nx * ny
    assert abs(lam - lam_star) / lam_star < 2e-3
    assert abs(x[0] - 0.5) < 1e-3
    # first decisions: lambda halves from 5e8 while the design is too light
    assert list(dec[:5]) == [0, 0, 0, 0, 0]


def test_heaviside_projection_closed_forms():
    """Pins that do not need the reference: s~ = 1/2 maps to 1/2, 0 to 0 and 1
    to 1 for any beta; large beta approaches a step; the sensitivity is the
    chain rule of the projection through the row-normalised filter
    (checked by central differences on a small grid)."""
    nx, ny = 10, 8
    xy, conn = po.grid2d(nx, ny)
    fp, fi, fw = po.filter_build(nx, ny, po.centroids(conn, xy))
    ne = nx * ny
    for beta in (1.0, 2.0, 8.0):
        for c, want in ((0.5, 0.5), (0.0, 0.0), (1.0, 1.0)):
            r = po.heaviside_vars(fp, fi, fw, beta, np.full(ne, c))
            assert np.allclose(r, want, atol=1e-15, rtol=0)
    r = po.heaviside_vars(fp, fi, fw, 64.0, np.full(ne, 0.55))
    assert np.all(r > 0.99)
    # the reference's sensitivity is row-normalised by element i's weight sum,
    # so it equals the exact gradient d(c . rho)/ds where every weight sum in
    # the neighbourhood is the same: two elements away from the boundary
    rng = np.random.default_rng(3)
    s = rng.uniform(0.2, 0.8, ne)
    c = rng.uniform(-1, 1, ne)
    beta = 4.0
    dfds = po.heaviside_sens(fp, fi, fw, beta, s, c)
    h = 1e-6
    grad = np.zeros(ne)
    for k in range(ne):
        sp, sm = s.copy(), s.copy()
        sp[k] += h
        sm[k] -= h
        grad[k] = (c @ po.heaviside_vars(fp, fi, fw, beta, sp) - c @ po.heaviside_vars(fp, fi, fw, beta, sm)) / (2 * h)
    interior = [ny * i + j for i in range(2, nx - 2) for j in range(2, ny - 2)]
    assert len(interior) == 24
    assert np.allclose(dfds[interior], grad[interior], rtol=1e-6, atol=1e-9)
    # and a dropped chain-rule factor would not pass
    assert not np.allclose(po.filter_sens(fp, fi, fw, c)[interior], grad[interior], rtol=1e-3)


def test_sensitivity_filters_against_reference():
    """Both sensitivity filters against the reference's own classes on a
    24x12 grid (golden sensfilter_24x12), and the closed form: with uniform
    s they reduce to the row-normalised density-filter sensitivity."""
    g = gi.load("sensfilter_24x12")
    nx, ny = 24, 12
    xy, conn = po.grid2d(nx, ny)
    fp, fi, fw = po.filter_build(nx, ny, po.centroids(conn, xy))
    assert np.array_equal(po.sensitivity_filter(fp, fi, fw, g["s"], g["dfds"], 1), g["out1"])
    assert np.array_equal(po.sensitivity_filter(fp, fi, fw, g["s"], g["dfds"], 2), g["out2"])
    u = np.full(nx * ny, 0.7)
    want = po.filter_sens(fp, fi, fw, g["dfds"])
    assert np.allclose(po.sensitivity_filter(fp, fi, fw, u, g["dfds"], 1), want, rtol=1e-14)
    assert np.allclose(po.sensitivity_filter(fp, fi, fw, u, g["dfds"], 2), want, rtol=1e-14)
