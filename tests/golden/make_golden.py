"""Write the golden fixtures under tests/golden/ from the reference itself.

Runs oracle/_ref/ref_harness (compiled from the reference headers by
oracle/Makefile with the reference README's flags, g++ -O3 -fopenmp) and
packs its raw outputs into compressed .npz files plus manifest.json.  Large
arrays are stored as SHA-256 digests of their little-endian bytes (the tests
compare digests for the bit-exact checks) with the full vectors kept for the
smaller ones.  Run from the repository root in a container that has the
reference tree:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden")
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF = os.environ.get("PANSFEM2_REFERENCE", "/root/reference")

# arrays larger than this many bytes are kept as digests only
FULL_LIMIT = 400_000


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_dir(d):
    out = {}
    for fn in sorted(os.listdir(d)):
        name, ext = fn.rsplit(".", 1)
        out[name] = np.fromfile(os.path.join(d, fn), dtype={"f64": np.float64, "i32": np.int32}[ext])
    return out


def pack(case, arrays, meta, manifest, keep=None):
    store, digests = {}, {}
    for k, v in arrays.items():
        digests[k] = {"sha256": sha(v), "dtype": str(v.dtype), "size": int(v.size)}
        if v.nbytes <= FULL_LIMIT or (keep and k in keep):
            store[k] = v
    np.savez_compressed(os.path.join(OUT, case + ".npz"), **store)
    manifest[case] = {"meta": meta, "arrays": digests}


def run(args, tmp):
    subprocess.run([HARNESS] + [str(a) for a in args], check=True, stdout=subprocess.DEVNULL)
    return load_dir(tmp)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all", "REF=" + REF], check=True)
    manifest = {}
    simp_meta = dict(E1=1.0, E0=1e-9, nu=0.3, t=1.0, p=3.0, vf=0.5, R=1.5, scale0=1.0, scale1=1.0, s0=0.5,
                     cg_itrmax=100000, cg_eps=1e-10, clamp="x=0 nodes, dofs {0,1}",
                     load="node at (nx, 0), dof 1, value -1", lx="nx", ly="ny")
    oc_meta = dict(iota=0.5, lambdamin=0.0, lambdamax=1e9, lambdaeps=1e-3, move=0.2)
    conlin_meta = dict(optimizer="CONLIN", a0=1.0, a=[0.0], c=[1e4], d=[0.0], xmin=0.01, xmax=1.0, move=0.2,
                       epsvalue=1e-6, source="sample/optimize/sample_optimize_density_CONLIN.cpp:80-85")
    heaviside_meta = dict(filter="HeavisideFilter", beta0=0.5, beta_factor=2.0, beta_every=40,
                          beta_rule="beta *= 2 at k % 40 == 0, before the filter (k = 0 gives beta = 1)")
    mma_meta = dict(a0=1.0, a=[0.0], c=[1e4], d=[0.0], xmin=0.01, xmax=1.0, raa0=1e-5, albefa=0.1, move=0.2,
                    asyinit=0.5, asydecr=0.7, asyincr=1.2, epsvalue=1e-6)
    cases = [
        ("simp_4x2_oc", ["simp", 4, 2, 3, "oc"], 3, dict(simp_meta, **oc_meta, nx=4, ny=2, iters=3)),
        ("simp_8x4_oc", ["simp", 8, 4, 5, "oc"], 5, dict(simp_meta, **oc_meta, nx=8, ny=4, iters=5)),
        ("simp_16x8_mma", ["simp", 16, 8, 5, "mma"], 5, dict(simp_meta, **mma_meta, nx=16, ny=8, iters=5)),
        ("simp_120x40_oc", ["simp", 120, 40, 50, "oc"], 0, dict(simp_meta, **oc_meta, nx=120, ny=40, iters=50)),
        ("simp_120x40_mma", ["simp", 120, 40, 50, "mma"], 0, dict(simp_meta, **mma_meta, nx=120, ny=40, iters=50)),
        ("simp_16x8_ocH", ["simp", 16, 8, 5, "ocH"], 5, dict(simp_meta, **oc_meta, **heaviside_meta, nx=16, ny=8, iters=5)),
        ("simp_16x8_conlin", ["simp", 16, 8, 5, "conlin"], 5, dict(simp_meta, **conlin_meta, nx=16, ny=8, iters=5)),
        ("simp_120x40_conlin", ["simp", 120, 40, 50, "conlin"], 0,
         dict(simp_meta, **conlin_meta, nx=120, ny=40, iters=50)),
        ("simp_120x40_ocH", ["simp", 120, 40, 50, "ocH"], 0,
         dict(simp_meta, **oc_meta, **heaviside_meta, nx=120, ny=40, iters=50)),
    ]
    for case, args, nsave, meta in cases:
        tmp = tempfile.mkdtemp()
        arrays = run(args + [tmp, nsave], tmp)
        pack(case, arrays, meta, manifest, keep={"u0_free", "F", "rho0", "dfdrho0", "dfds0", "s_final", "rho_final"})
        shutil.rmtree(tmp)
        print("wrote", case, file=sys.stderr)
    for nx, ny, seed in [(16, 8, 20261202), (64, 32, 20261203)]:
        tmp = tempfile.mkdtemp()
        arrays = run(["operator", nx, ny, seed, tmp], tmp)
        pack(f"operator_{nx}x{ny}", arrays,
             dict(nx=nx, ny=ny, seed=seed, rho="iid U[1e-3,1] SplitMix64", x="iid U[-1,1] SplitMix64 (after rho)",
                  E1=1.0, E0=1e-9, p=3.0, nu=0.3), manifest)
        shutil.rmtree(tmp)
        print("wrote operator", nx, ny, file=sys.stderr)
    subprocess.run([HARNESS, "vtk", os.path.join(OUT, "vtk_4x2.vtk")], check=True)
    manifest["vtk_4x2"] = {"meta": dict(source="src/PrePost/Export/ExportToVTK.h", file="vtk_4x2.vtk",
                                        fields="closed-form u, r (nodal vectors), rho (element scalars)"),
                           "arrays": {}}
    print("wrote vtk", file=sys.stderr)
    tmp = tempfile.mkdtemp()
    arrays = run(["sensfilter", 24, 12, 20261206, tmp], tmp)
    pack("sensfilter_24x12", arrays, dict(nx=24, ny=12, R=1.5, seed=20261206, s="iid U[0.01,1] SplitMix64",
                                          dfds="iid U[-1,0] SplitMix64 (after s)",
                                          source="src/Optimize/Filter/SensitivityFilter.h"), manifest)
    shutil.rmtree(tmp)
    print("wrote sensfilter", file=sys.stderr)
    tmp = tempfile.mkdtemp()
    arrays = run(["solid", os.path.join(REF, "sample", "solid"), tmp], tmp)
    pack("solid_sample", arrays, dict(source="sample/solid/{Node,Element,Dirichlet,Neumann}.csv", E=210000.0, nu=0.3,
                                      cg_eps=1e-10), manifest, keep={"u_free", "F", "ke_e0"})
    shutil.rmtree(tmp)
    tmp = tempfile.mkdtemp()
    arrays = run(["mmatest", tmp], tmp)
    pack("mma_cantilever", arrays, dict(source="src/Optimize/Solver/test_MMA.cpp", layout="per k: obj, con, s[0..4]"),
         manifest)
    shutil.rmtree(tmp)
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
