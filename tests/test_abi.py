"""The C-ABI library loads without a GPU and exports every entry point that
include/alien_hip.h declares (no compute call is made here)."""
import ctypes
import pathlib
import re

import alien_amd as aa
from alien_amd import _capi

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "alien_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(alien_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = aa.lib()
    names = declared_functions()
    assert len(names) >= 40
    for name in names:
        assert hasattr(L, name), name


def test_binding_covers_header():
    assert sorted(_capi.EXPORTED) == declared_functions()


def test_version_and_error_channel():
    L = aa.lib()
    assert b"gfx950" in L.alien_hip_version()
    # host-side argument validation returns EINVAL with a message, without a GPU
    rc = L.alien_hip_host_dcol(-1, None, None, None)
    assert rc == aa._capi.EINVAL
    assert b"bad arguments" in L.alien_hip_last_error()


def test_solve_rejects_bad_csr_before_any_gpu_work():
    import numpy as np
    ro = np.array([0, 2, 3], np.int32)
    cols = np.array([1, 0, 1], np.int32)  # row 0 lacks its diagonal
    vals = np.ones(3)
    p = _capi.SolveParams(aa.CG, aa.PREC_DIAG, aa.LAYOUT_CSR, aa.MODE_REFERENCE, 1e-8, 10, 0.5, 3, 10, 0, 1e-4,
                          0, 10, 4)
    r = _capi.SolveResult()
    b = np.ones(2)
    x = np.zeros(2)
    # a NULL context is rejected before the CSR is even read
    rc = aa.lib().alien_hip_solve_csr(None, ctypes.byref(p), 2, _capi.iptr(ro), _capi.iptr(cols), _capi.dptr(vals),
                                      _capi.dptr(b), _capi.dptr(x), ctypes.byref(r))
    assert rc == aa._capi.EINVAL
