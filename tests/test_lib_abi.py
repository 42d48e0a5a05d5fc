"""The C-ABI library loads and exports every function declared in include/wbmi.h (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wbmi.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wbmi_[a-z0-9_]+)\s*\(", src)))


def test_header_lists_entry_points():
    names = declared_functions()
    assert "wbmi_static_scan" in names and "wbmi_eigh" in names and len(names) >= 15


def test_library_exports_all_symbols():
    from wannierberri_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libwbmi.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name
    assert set(declared_functions()) == set(_lib.SYMBOLS)
    assert L.wbmi_version() == 1


def test_ctx_create_rejects_bad_device_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present: covered by the gpu tests")
    from wannierberri_amd import _lib
    h = ctypes.c_void_p()
    rc = _lib.lib().wbmi_ctx_create(0, None, ctypes.byref(h))
    assert rc != _lib.WBMI_OK


def test_product_path_refuses_cpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import numpy as np
    import wannierberri_amd as wb
    from wannierberri_amd import models
    s = models.haldane()
    g = wb.Grid(system=s, NK=[4, 4, 1], NKFFT=[4, 4, 1])
    with pytest.raises(Exception):
        wb.run(s, g, {"ahc": wb.calculators.static.AHC(Efermi=np.linspace(-1, 1, 3))}, save_files=False)
