"""ctypes view of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke entry and the CPU-baseline leg of bench.py import this
module; it is the checker, never the thing measured or shipped.  Each wrapper
names the oracle routine and, through it, the reference file:line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_lib = None


def build():
    """Compile liboracle.so (and the reference harness where the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        c_int, c_long, c_double, c_void_p = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
        L.orc_grid2d_nodes.argtypes = [c_int, c_int, c_double, c_double, _f64p]
        L.orc_grid2d_elements.argtypes = [c_int, c_int, _i32p]
        L.orc_dofmap.argtypes = [c_int, c_int, _i32p, _i32p, c_int, _i32p]
        L.orc_dofmap.restype = c_int
        L.orc_ke_planestrain_q4.argtypes = [_f64p, c_double, c_double, c_double, _f64p]
        L.orc_ke_solid_h8.argtypes = [_f64p, c_double, c_double, _f64p]
        pp_i = ctypes.POINTER(ctypes.POINTER(ctypes.c_int))
        pp_d = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))
        L.orc_assemble_csr.argtypes = [c_int, c_int, c_int, c_int, _i32p, _i32p, _f64p, _f64p, _f64p, pp_i, pp_i, pp_d]
        L.orc_assemble_csr.restype = c_long
        L.orc_free.argtypes = [c_void_p]
        L.orc_csr_spmv.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_csr_diag.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p]
        L.orc_scaling_cg.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p, c_int, c_double, _f64p]
        L.orc_scaling_cg.restype = c_int
        L.orc_centroids.argtypes = [c_int, c_int, c_int, _i32p, _f64p, _f64p]
        L.orc_filter_build.argtypes = [c_int, c_int, _f64p, c_double, c_int, _i32p, pp_i, pp_d]
        L.orc_filter_build.restype = c_long
        L.orc_filter_build_brute.argtypes = [c_int, c_int, _f64p, c_double, _i32p, pp_i, pp_d]
        L.orc_filter_build_brute.restype = c_long
        L.orc_filter_apply.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_filter_sens.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_sensitivity_filter.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p, _f64p, c_int, _f64p]
        L.orc_heaviside_vars.argtypes = [c_int, _i32p, _i32p, _f64p, c_double, _f64p, _f64p]
        L.orc_heaviside_sens.argtypes = [c_int, _i32p, _i32p, _f64p, c_double, _f64p, _f64p, _f64p]
        L.orc_oc_update.argtypes = [c_int, _f64p, _f64p, _f64p, c_double, c_double, c_double, c_double, c_double,
                                    _i32p, _i32p, _f64p, c_double, c_double, ctypes.POINTER(c_double), _i32p, c_int]
        L.orc_oc_update.restype = c_int
        L.orc_mma_create.argtypes = [c_int, c_int, c_double, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.orc_mma_create.restype = c_void_p
        L.orc_mma_set_parameters.argtypes = [c_void_p] + [c_double] * 7
        L.orc_mma_is_convergence.argtypes = [c_void_p, c_double]
        L.orc_mma_is_convergence.restype = c_int
        L.orc_mma_update.argtypes = [c_void_p, _f64p, c_double, _f64p, _f64p, _f64p]
        L.orc_mma_destroy.argtypes = [c_void_p]
        L.orc_simp_default_params.argtypes = [c_void_p]
        L.orc_simp2d.argtypes = [c_int, c_int, c_int, c_void_p, _f64p, _f64p, _f64p, _f64p, _f64p, _i32p, _f64p]
        L.orc_simp2d.restype = c_int
        _lib = L
    return _lib


# ------------------------------------------------------------------ mesh ---
def grid2d(nx, ny, lx=None, ly=None):
    """SquareMesh nodes (nn x 2) and elements (ne x 4) (SquareMesh.h:60-106)."""
    lx = float(nx) if lx is None else lx
    ly = float(ny) if ly is None else ly
    xy = np.zeros(((nx + 1) * (ny + 1), 2))
    conn = np.zeros((nx * ny, 4), np.int32)
    lib().orc_grid2d_nodes(nx, ny, lx, ly, xy)
    lib().orc_grid2d_elements(nx, ny, conn)
    return xy, conn


def left_clamp_fixed(nx, ny):
    """Node/dof pairs of the left-edge clamp, in GenerateFixedlist order."""
    nodes = np.repeat(np.arange(ny + 1, dtype=np.int32), 2)
    dofs = np.tile(np.array([0, 1], np.int32), ny + 1)
    return nodes, dofs


def dofmap(nn, ndpn, fixed_node, fixed_dof):
    n2g = np.zeros(nn * ndpn, np.int32)
    N = lib().orc_dofmap(nn, ndpn, np.ascontiguousarray(fixed_node, np.int32),
                         np.ascontiguousarray(fixed_dof, np.int32), len(fixed_node), n2g)
    return N, n2g


def ke_q4(X, E=1.0, nu=0.3, t=1.0):
    Ke = np.zeros((8, 8))
    lib().orc_ke_planestrain_q4(np.ascontiguousarray(X, np.float64), E, nu, t, Ke)
    return Ke


def ke_h8(X, E=1.0, nu=0.3):
    Ke = np.zeros((24, 24))
    lib().orc_ke_solid_h8(np.ascontiguousarray(X, np.float64), E, nu, Ke)
    return Ke


def assemble_csr(N, conn, n2g, ke_all, ndpn, ufix=None, F=None):
    ne, nen = conn.shape
    if ufix is None:
        ufix = np.zeros(n2g.size)
    if F is None:
        F = np.zeros(N)
    ip = ctypes.POINTER(ctypes.c_int)()
    ix = ctypes.POINTER(ctypes.c_int)()
    dv = ctypes.POINTER(ctypes.c_double)()
    L = lib()
    nnz = L.orc_assemble_csr(N, ne, nen, ndpn, np.ascontiguousarray(conn, np.int32), np.ascontiguousarray(n2g, np.int32),
                             np.ascontiguousarray(ke_all, np.float64), np.ascontiguousarray(ufix, np.float64), F,
                             ctypes.byref(ip), ctypes.byref(ix), ctypes.byref(dv))
    indptr = np.ctypeslib.as_array(ip, (N + 1,)).copy()
    indices = np.ctypeslib.as_array(ix, (max(nnz, 1),))[:nnz].copy()
    data = np.ctypeslib.as_array(dv, (max(nnz, 1),))[:nnz].copy()
    for p in (ip, ix, dv):
        L.orc_free(ctypes.cast(p, ctypes.c_void_p))
    return indptr, indices, data, F


def spmv(indptr, indices, data, x):
    y = np.zeros(len(indptr) - 1)
    lib().orc_csr_spmv(len(indptr) - 1, indptr, indices, data, np.ascontiguousarray(x, np.float64), y)
    return y


def csr_diag(indptr, indices, data):
    D = np.zeros(len(indptr) - 1)
    lib().orc_csr_diag(len(indptr) - 1, indptr, indices, data, D)
    return D


def scaling_cg(indptr, indices, data, b, itrmax=100000, eps=1e-10):
    x = np.zeros(len(b))
    it = lib().orc_scaling_cg(len(b), indptr, indices, data, np.ascontiguousarray(b, np.float64), itrmax, eps, x)
    return x, it


def centroids(conn, nodes):
    ne, nen = conn.shape
    dim = nodes.shape[1]
    cg = np.zeros((ne, dim))
    lib().orc_centroids(ne, nen, dim, np.ascontiguousarray(conn, np.int32), np.ascontiguousarray(nodes, np.float64), cg)
    return cg


def filter_build(nx, ny, cg, R=1.5, brute=False):
    ne = nx * ny
    ptr = np.zeros(ne + 1, np.int32)
    ix = ctypes.POINTER(ctypes.c_int)()
    w = ctypes.POINTER(ctypes.c_double)()
    L = lib()
    nent = L.orc_filter_build(nx, ny, np.ascontiguousarray(cg, np.float64), R, 1 if brute else 0, ptr,
                              ctypes.byref(ix), ctypes.byref(w))
    idx = np.ctypeslib.as_array(ix, (max(nent, 1),))[:nent].copy()
    ww = np.ctypeslib.as_array(w, (max(nent, 1),))[:nent].copy()
    L.orc_free(ctypes.cast(ix, ctypes.c_void_p))
    L.orc_free(ctypes.cast(w, ctypes.c_void_p))
    return ptr, idx, ww


def filter_build_brute(cg, R=1.5):
    ne, dim = cg.shape
    ptr = np.zeros(ne + 1, np.int32)
    ix = ctypes.POINTER(ctypes.c_int)()
    w = ctypes.POINTER(ctypes.c_double)()
    L = lib()
    nent = L.orc_filter_build_brute(ne, dim, np.ascontiguousarray(cg, np.float64), R, ptr, ctypes.byref(ix),
                                    ctypes.byref(w))
    idx = np.ctypeslib.as_array(ix, (max(nent, 1),))[:nent].copy()
    ww = np.ctypeslib.as_array(w, (max(nent, 1),))[:nent].copy()
    L.orc_free(ctypes.cast(ix, ctypes.c_void_p))
    L.orc_free(ctypes.cast(w, ctypes.c_void_p))
    return ptr, idx, ww


def grid3d(nx, ny, nz, lx=None, ly=None, lz=None):
    """Structured hex8 grid in the library's numbering: node (i,j,k) ->
    (ny+1)(nz+1) i + (nz+1) j + k, element -> ny nz i + nz j + k, nodes in
    ShapeFunction8Cubic order; coordinates l*(i/(double)n) as SquareMesh."""
    lx = float(nx) if lx is None else lx
    ly = float(ny) if ly is None else ly
    lz = float(nz) if lz is None else lz
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    xyz = np.stack([lx * (I / float(nx)), ly * (J / float(ny)), lz * (K / float(nz))], -1).reshape(-1, 3)
    node = lambda i, j, k: (ny + 1) * (nz + 1) * i + (nz + 1) * j + k
    conn = np.zeros((nx * ny * nz, 8), np.int32)
    e = 0
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                conn[e] = [node(i, j, k), node(i + 1, j, k), node(i + 1, j + 1, k), node(i, j + 1, k),
                           node(i, j, k + 1), node(i + 1, j, k + 1), node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)]
                e += 1
    return xyz, conn


def simp3d(nx, ny, nz, iters, loads, E1=1.0, E0=1e-9, nu=0.3, p=3.0, vf=0.3, R=1.5, cg_eps=1e-10, s0=None):
    """The SIMP loop (sample_optimize_density_oc.cpp:83-208) composed from the
    oracle's routines on a hex8 grid with the x = 0 face clamped; `loads` is a
    list of (node, dof, value).  Small grids only (all-pairs filter search)."""
    xyz, conn = grid3d(nx, ny, nz)
    nn, ne = xyz.shape[0], conn.shape[0]
    fixed = np.arange((ny + 1) * (nz + 1))
    N, n2g = dofmap(nn, 3, np.repeat(fixed, 3), np.tile(np.arange(3), len(fixed)))
    fp, fi, fw = filter_build_brute(centroids(conn, xyz), R)
    ke1 = np.stack([ke_h8(xyz[c], 1.0, nu) for c in conn])
    s = np.full(ne, vf) if s0 is None else np.array(s0, np.float64)
    hist = []
    dgds = filter_sens(fp, fi, fw, np.full(ne, 1.0 / (vf * ne)))
    for k in range(iters):
        rho = filter_apply(fp, fi, fw, s)
        g = 0.0
        for r in rho:
            g += 1.0 * r / (vf * ne)
        g -= 1.0
        E = [E1 * np.power(r, p) + E0 * (1.0 - np.power(r, p)) for r in rho]
        ke = np.stack([ke_h8(xyz[c], e, nu) for c, e in zip(conn, E)])
        indptr, indices, data, F = assemble_csr(N, conn, n2g, ke, 3)
        for (nd, d, v) in loads:
            if n2g[3 * nd + d] >= 0:
                F[n2g[3 * nd + d]] += v
        u, it = scaling_cg(indptr, indices, data, F, eps=cg_eps)
        un = np.where(n2g >= 0, u[np.maximum(n2g, 0)], 0.0).reshape(-1, 3)
        f = 0.0
        dfdrho = np.zeros(ne)
        for e in range(ne):
            ue = un[conn[e]].reshape(-1)
            f += float(ue @ (ke[e] @ ue))
            dfdrho[e] = -p * (E1 - E0) * rho[e] ** (p - 1.0) * float(ue @ (ke1[e] @ ue))
        dfds = filter_sens(fp, fi, fw, dfdrho)
        hist.append((f, g, it))
        s, lam, dec = oc_update(s, dfds, dgds, fp, fi, fw, vf=vf)
    return dict(s=s, rho=filter_apply(fp, fi, fw, s), hist=hist, n2g=n2g, N=N)


def filter_apply(ptr, idx, w, s):
    out = np.zeros(len(ptr) - 1)
    lib().orc_filter_apply(len(ptr) - 1, ptr, idx, w, np.ascontiguousarray(s, np.float64), out)
    return out


def filter_sens(ptr, idx, w, dfdrho):
    out = np.zeros(len(ptr) - 1)
    lib().orc_filter_sens(len(ptr) - 1, ptr, idx, w, np.ascontiguousarray(dfdrho, np.float64), out)
    return out


def sensitivity_filter(ptr, idx, w, s, dfds, scheme=1):
    """SensitivityFilter (scheme 1) / SensitivityFilter2 (scheme 2) (SensitivityFilter.h:44-99)."""
    out = np.zeros(len(ptr) - 1)
    lib().orc_sensitivity_filter(len(ptr) - 1, ptr, idx, w, np.ascontiguousarray(s, np.float64),
                                 np.ascontiguousarray(dfds, np.float64), scheme, out)
    return out


def heaviside_vars(ptr, idx, w, beta, s):
    """HeavisideFilter::GetFilteredVariables (HeavisideFilter.h:62-75)."""
    out = np.zeros(len(ptr) - 1)
    lib().orc_heaviside_vars(len(ptr) - 1, ptr, idx, w, beta, np.ascontiguousarray(s, np.float64), out)
    return out


def heaviside_sens(ptr, idx, w, beta, s, dfdrho):
    """HeavisideFilter::GetFilteredSensitivitis (HeavisideFilter.h:78-99)."""
    out = np.zeros(len(ptr) - 1)
    lib().orc_heaviside_sens(len(ptr) - 1, ptr, idx, w, beta, np.ascontiguousarray(s, np.float64),
                             np.ascontiguousarray(dfdrho, np.float64), out)
    return out


def oc_update(x, dfdx, dgdx, fptr, fidx, fw, iota=0.5, lmin=0.0, lmax=1e9, leps=1e-3, move=0.2, vf=0.5, scale1=1.0):
    x = np.array(x, np.float64)
    lam = ctypes.c_double(0.0)
    dec = np.zeros(256, np.int32)
    n = lib().orc_oc_update(len(x), x, np.ascontiguousarray(dfdx, np.float64), np.ascontiguousarray(dgdx, np.float64),
                            iota, lmin, lmax, leps, move, fptr, fidx, fw, vf, scale1, ctypes.byref(lam), dec, 256)
    return x, lam.value, dec[:n]


class MMA:
    """orc_mma: MMA<T> (src/Optimize/Solver/MMA.h:15-509)."""

    def __init__(self, n, m, a0, a, c, d, xmin, xmax):
        f = lambda v: np.ascontiguousarray(v, np.float64)
        self._h = lib().orc_mma_create(n, m, a0, f(a), f(c), f(d), f(xmin), f(xmax))
        self.n, self.m = n, m

    def set_parameters(self, raa0, albefa, move, asyinit, asydecr, asyincr, epsvalue):
        lib().orc_mma_set_parameters(self._h, raa0, albefa, move, asyinit, asydecr, asyincr, epsvalue)

    def is_convergence(self, f):
        return bool(lib().orc_mma_is_convergence(self._h, f))

    def update(self, x, f, dfdx, g, dgdx):
        x = np.array(x, np.float64)
        lib().orc_mma_update(self._h, x, f, np.ascontiguousarray(dfdx, np.float64), np.ascontiguousarray(g, np.float64),
                             np.ascontiguousarray(dgdx, np.float64).reshape(-1))
        return x

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_mma_destroy(self._h)
            self._h = None


class SimpParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("E1", "E0", "nu", "t", "p", "vf", "R", "scale0", "scale1", "cg_eps")] + \
               [("cg_itrmax", ctypes.c_int), ("use_mma", ctypes.c_int)] + \
               [(n, ctypes.c_double) for n in ("oc_iota", "oc_lmin", "oc_lmax", "oc_leps", "oc_move", "mma_a0", "mma_c",
                                               "mma_xmin", "mma_xmax", "mma_raa0", "mma_albefa", "mma_move",
                                               "mma_asyinit", "mma_asydecr", "mma_asyincr", "mma_eps")] + \
               [("heaviside", ctypes.c_int), ("beta0", ctypes.c_double), ("beta_factor", ctypes.c_double),
                ("beta_every", ctypes.c_int)]


def simp_params(**kw):
    p = SimpParams()
    lib().orc_simp_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def simp2d(nx, ny, iters, params=None, s0=0.5):
    """orc_simp2d: the sample driver's SIMP loop (sample_optimize_density_oc.cpp:83-208)."""
    params = params or simp_params()
    ne = nx * ny
    N = 2 * nx * (ny + 1)
    s = np.full(ne, s0)
    rho = np.zeros(ne)
    fh, gh, lh = np.zeros(iters), np.zeros(iters), np.zeros(iters)
    ch = np.zeros(iters, np.int32)
    u0 = np.zeros(N)
    lib().orc_simp2d(nx, ny, iters, ctypes.byref(params), s, rho, fh, gh, lh, ch, u0)
    return dict(s=s, rho=rho, f=fh, g=gh, lam=lh, cg_iters=ch, u0=u0)


def run_ref_harness(args, outdir):
    """Run the reference-compiled harness (oracle/_ref) -- container/golden use only."""
    os.makedirs(outdir, exist_ok=True)
    subprocess.run([REF_HARNESS] + [str(a) for a in args], check=True, stdout=subprocess.DEVNULL)
