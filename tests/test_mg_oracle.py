"""Pins of the multigrid restatement (oracle/mg.py) by mathematics: the
multigrid has no counterpart in the reference, so no reference vector
covers it.  A wrong interpolation weight, a transposed Galerkin product or
a dropped fixed-DOF mask fails the first test; an unsymmetric smoothing or
a restriction that is not P^T fails the second; a coarse correction that
does not act fails the third."""
import numpy as np
import scipy.sparse as sp

from oracle import mg
from oracle import pyoracle as po


def unit_k0():
    xy, conn = po.grid2d(1, 1)
    return po.ke_q4(xy[conn[0]], 1.0, 0.3, 1.0)


def cantilever_fix(nx, ny):
    fix = np.zeros((nx + 1) * (ny + 1), np.uint8)
    fix[: ny + 1] = 3  # left edge, both directions
    return fix


def test_galerkin_identity():
    nx, ny = 16, 8
    rng = np.random.default_rng(11)
    E = 1e-9 + rng.uniform(0, 1, nx * ny) ** 3
    K0 = unit_k0()
    fix = cantilever_fix(nx, ny)
    lv = mg.build(nx, ny, E, K0, fix, coarse_dofs=20)
    assert len(lv) == 4
    # level 0 from the unmasked element matrices and an explicit mask
    A_raw = mg.assemble(nx, ny, E[:, None, None] * K0[None])
    M = sp.diags((~mg.fix_dofs(fix)).astype(float))
    A = M @ A_raw @ M
    for k in range(1, len(lv)):
        P = mg.prolongation(lv[k - 1]["nx"], lv[k - 1]["ny"])
        Mk = sp.diags(lv[k - 1]["free"].astype(float)) if k > 1 else M
        want = (P.T @ Mk @ A @ Mk @ P).toarray()
        got = lv[k]["A"].toarray()
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
        A = lv[k]["A"]


def test_vcycle_is_symmetric():
    nx, ny = 32, 16
    rng = np.random.default_rng(12)
    E = 1e-9 + rng.uniform(0, 1, nx * ny) ** 3
    lv = mg.build(nx, ny, E, unit_k0(), cantilever_fix(nx, ny), coarse_dofs=100)
    free = lv[0]["free"]
    a = rng.standard_normal(free.size) * free
    b = rng.standard_normal(free.size) * free
    Va, Vb = mg.vcycle(lv, a), mg.vcycle(lv, b)
    assert abs(Va @ b - a @ Vb) <= 1e-12 * np.linalg.norm(Va) * np.linalg.norm(b)
    assert Va @ a > 0  # positive definite on this sample


def test_mg_pcg_converges_where_jacobi_crawls():
    nx, ny = 64, 32
    rng = np.random.default_rng(13)
    E = 1e-9 + (1 - 1e-9) * rng.uniform(1e-3, 1, nx * ny) ** 3
    lv = mg.build(nx, ny, E, unit_k0(), cantilever_fix(nx, ny))
    l0 = lv[0]
    b = np.zeros(l0["free"].size)
    b[2 * ((ny + 1) * nx) + 1] = -1.0
    dinv = np.where(l0["free"], 1.0 / l0["D"], 0.0)
    xj, kj = mg.pcg(lambda v: mg.apply_level(l0, v), b, lambda r: dinv * r)
    xm, km = mg.pcg(lambda v: mg.apply_level(l0, v), b, lambda r: mg.vcycle(lv, r))
    assert kj > 500 and km < 100, (kj, km)
    assert np.linalg.norm(xm - xj) / np.linalg.norm(xj) < 1e-6


def unit_k0_h8():
    xy, conn = po.grid3d(1, 1, 1)
    return po.ke_h8(xy[conn[0]], 1.0, 0.3)


def test_galerkin_identity_hex8():
    nx, ny, nz = 4, 2, 2
    rng = np.random.default_rng(21)
    E = 1e-9 + rng.uniform(0, 1, nx * ny * nz) ** 3
    K0 = unit_k0_h8()
    fix = np.zeros((nx + 1) * (ny + 1) * (nz + 1), np.uint8)
    fix[: (ny + 1) * (nz + 1)] = 7  # x = 0 face
    lv = mg.build3(nx, ny, nz, E, K0, fix, coarse_dofs=10)
    assert len(lv) == 2
    A_raw = mg.assemble3(nx, ny, nz, E[:, None, None] * K0[None])
    M = sp.diags((~mg.fix_dofs3(fix)).astype(float))
    P = mg.prolongation3(nx, ny, nz)
    want = (P.T @ M @ A_raw @ M @ P).toarray()
    got = lv[1]["A"].toarray()
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))
    # and the V-cycle over it is symmetric
    free = lv[0]["free"]
    a = rng.standard_normal(free.size) * free
    b = rng.standard_normal(free.size) * free
    Va, Vb = mg.vcycle(lv, a), mg.vcycle(lv, b)
    assert abs(Va @ b - a @ Vb) <= 1e-12 * np.linalg.norm(Va) * np.linalg.norm(b)
