"""The C++ drop-in facade (include/pansfem2/pansfem2.hpp).

CPU: the facade and its sample driver compile with the host compiler; the
facade's element routines (host code, evaluated in the reference's operation
order) are bit-identical to the oracle's restatements, which the golden
vectors written by the reference pin (tests/test_oracle.py).
GPU: the sample driver (pansfem2_amd/samples/simp_oc.cpp: mesh, Assembling,
CSR, ScalingCG, reactions, DensityFilter, OC with a host closure, all
through the facade) reproduces the reference's 50-iteration history at
120 x 40 (golden fixture simp_120x40_oc)."""
import os
import subprocess

import numpy as np
import pytest

from tests import golden_io as gi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAMPLE = os.path.join(ROOT, "pansfem2_amd", "samples", "simp_oc")


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("facade") / "facade_host_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "facade_host_check.cpp"), "-o", exe], check=True)
    return exe


def _run(exe, lines):
    out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    return np.array([float.fromhex(v) for v in out.stdout.split()])


def test_element_matrices_bit_identical_to_oracle(host_check):
    from oracle import pyoracle as po
    rng = np.random.default_rng(7)
    cases, want = [], []
    for k in range(6):
        X = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float) * rng.uniform(0.5, 3.0)
        if k >= 2:
            X = X + rng.uniform(-0.2, 0.2, X.shape)
        E, nu = rng.uniform(1e-3, 2.0), 0.3
        cases.append("q4 %r %r 1.0 " % (float(E), nu) + " ".join(repr(float(v)) for v in X.reshape(-1)))
        want.append(po.ke_q4(X, E, nu, 1.0).reshape(-1))
    for k in range(3):
        X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
        X = X * rng.uniform(0.5, 2.0) + (rng.uniform(-0.1, 0.1, X.shape) if k else 0.0)
        E, nu = rng.uniform(1e-3, 2.0), 0.3
        cases.append("h8 %r %r " % (float(E), nu) + " ".join(repr(float(v)) for v in X.reshape(-1)))
        want.append(po.ke_h8(X, E, nu).reshape(-1))
    got = _run(host_check, cases)
    assert got.size == sum(w.size for w in want)
    assert np.array_equal(got, np.concatenate(want))


def test_sample_driver_builds():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "pansfem2_amd", "samples")], check=True)
    assert os.access(SAMPLE, os.X_OK)


@pytest.mark.gpu
def test_sample_driver_against_reference(tmp_path):
    case = "simp_120x40_oc"
    g = gi.load(case)
    prefix = str(tmp_path / "c1")
    out = subprocess.run([SAMPLE, "120", "40", "50", prefix], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    hist = np.array([[float(v) for v in line.split()] for line in out.stdout.splitlines() if line.strip()])
    assert hist.shape == (50, 4)
    f = hist[:, 1]
    assert np.max(np.abs(f - g["hist_f"]) / g["hist_f"]) < 1e-9
    # g averages the filtered design, which may differ in the last bits
    assert np.max(np.abs(hist[:, 2] - g["hist_g"])) < 1e-9
    s = np.loadtxt(prefix + "_s.txt")
    rho = np.loadtxt(prefix + "_rho.txt")
    assert np.max(np.abs(s - g["s_final"])) < 1e-6
    assert np.max(np.abs(rho - g["rho_final"])) < 1e-6
    # the last iteration's VTK file: the reference sample's layout, and its
    # u and r (printed to 6 significant digits) give that iteration's compliance
    vtk = read_vtk(prefix + "_result49.vtk")
    assert vtk["points"].shape == (121 * 41, 3) and vtk["cells"].shape == (4800, 5)
    assert np.all(vtk["cells"][:, 0] == 4) and np.all(vtk["types"] == 9)
    f49 = float(np.sum(vtk["u"] * vtk["r"]))
    assert abs(f49 - g["hist_f"][49]) / g["hist_f"][49] < 1e-4
    assert vtk["s"].shape == (4800,) and np.all((vtk["s"] >= 0) & (vtk["s"] <= 1))


def read_vtk(path):
    """Legacy-ASCII VTK of ExportToVTK.h's layout: POINTS, CELLS, CELL_TYPES,
    POINT_DATA vectors u and r, CELL_DATA scalars s."""
    tok = open(path).read().split()
    out, i = {}, 0
    while i < len(tok):
        t = tok[i]
        if t == "POINTS":
            n = int(tok[i + 1])
            out["points"] = np.array(tok[i + 3:i + 3 + 3 * n], float).reshape(n, 3)
            i += 3 + 3 * n
        elif t == "CELLS":
            n, m = int(tok[i + 1]), int(tok[i + 2])
            out["cells"] = np.array(tok[i + 3:i + 3 + m], int).reshape(n, m // n)
            i += 3 + m
        elif t == "CELL_TYPES":
            n = int(tok[i + 1])
            out["types"] = np.array(tok[i + 2:i + 2 + n], int)
            i += 2 + n
        elif t in ("VECTORS", "SCALARS"):
            name, j = tok[i + 1], i + 3
            if t == "SCALARS":
                assert tok[j] == "LOOKUP_TABLE"
                j += 2
            k = j
            while k < len(tok) and tok[k][0] in "-+.0123456789":
                k += 1
            vals = np.array(tok[j:k], float)
            out[name] = vals.reshape(-1, 3) if t == "VECTORS" else vals
            i = k
        else:
            i += 1
    return out


@pytest.mark.gpu
def test_sample_driver_heaviside_against_reference(tmp_path):
    """The same driver with HeavisideFilter and its beta continuation."""
    case = "simp_120x40_ocH"
    g = gi.load(case)
    prefix = str(tmp_path / "c1h")
    out = subprocess.run([SAMPLE, "120", "40", "50", prefix, "H"], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    hist = np.array([[float(v) for v in line.split()] for line in out.stdout.splitlines() if line.strip()])
    assert hist.shape == (50, 4)
    assert np.max(np.abs(hist[:, 1] - g["hist_f"]) / g["hist_f"]) < 1e-9
    assert np.max(np.abs(np.loadtxt(prefix + "_s.txt") - g["s_final"])) < 1e-6
    assert np.max(np.abs(np.loadtxt(prefix + "_rho.txt") - g["rho_final"])) < 1e-6


def test_vtk_export_byte_identical_to_reference(tmp_path):
    """The facade's VTK functions write the same bytes as the reference's
    exporter (src/PrePost/Export/ExportToVTK.h) for the fixture's fields."""
    exe = str(tmp_path / "facade_vtk_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "facade_vtk_check.cpp"), "-o", exe], check=True)
    out = str(tmp_path / "out.vtk")
    subprocess.run([exe, out], check=True)
    with open(out, "rb") as a, open(os.path.join(ROOT, "tests", "golden", "vtk_4x2.vtk"), "rb") as b:
        assert a.read() == b.read()
