"""CPU checks of the C ABI: the library builds for gfx950, loads without a
GPU, and exports every entry point include/pansfem2_amd.h declares (no compute
calls are made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pansfem2_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pf_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    for required in ["pf_create", "pf_destroy", "pf_grid2d", "pf_set_dirichlet", "pf_assemble_csr", "pf_apply_K",
                     "pf_pcg", "pf_compliance_sens", "pf_filter", "pf_oc_update", "pf_mma_update", "pf_simp_run"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    import pansfem2_amd as pa
    L = pa.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    assert b"gfx950" in L.pf_build_info()


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import pansfem2_amd as pa
    with pytest.raises(pa.PfError):
        pa.Context(0)


def test_code_object_targets_gfx950():
    so = os.path.join(ROOT, "pansfem2_amd", "libpansfem2_amd.so")
    data = open(so, "rb").read()
    assert b"gfx950" in data
