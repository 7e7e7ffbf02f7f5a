"""The C ABI library loads and exports every symbol include/*.h declares
(no compute: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sci-solver_peridynamic_amd", "libpdsolver.so")


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pd_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("pd_ctx_create", "pd_step", "pd_build_bonds", "pd_spmv", "pd_solve"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail(f"library not built: {LIB}")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_reported_not_crashed():
    import numpy as np
    import sys
    sys.path.insert(0, os.path.join(ROOT, "sci-solver_peridynamic_amd"))
    import pdsolver as pds
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = True
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(pds.PDError) as e:
        pds.Context(10, [(0.03, 0.005, 100.0, 1e-3)])
    assert e.value.code in (pds.PD_ERR_NO_DEVICE, pds.PD_ERR_HIP)
