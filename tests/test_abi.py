"""CPU-side checks of the C ABI: the built library loads and exports every
function declared in include/mnpbem_amd.h, and the ctypes table covers them.
No device compute is called here."""

import os

import numpy as np
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mnpbem_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(mnpbem_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "mnpbem_getrf_batched" in names and "mnpbem_bem_solve_batched" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from mnpbem_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libmnpbem_amd.so not built")
    lib = _native.load()
    for name in declared():
        assert hasattr(lib, name), name


def test_ctypes_table_matches_header():
    from mnpbem_amd import _native
    names = set(declared()) - {"mnpbem_last_error"}
    assert names == set(_native.SIGNATURES)


def test_invalid_arguments_are_rejected_without_device_work():
    """Argument validation happens on the host before any launch."""
    from mnpbem_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libmnpbem_amd.so not built")
    lib = _native.load()
    rc = lib.mnpbem_getrf_batched(_native.DTYPE_C128, 10, None, 5, 100, None, 10, None, 1, None)
    assert rc == 1
    assert b"getrf" in lib.mnpbem_last_error()
    rc = lib.mnpbem_refine_coeffs(9, 0, *([None] * 11), 28, None, 144, None, 0, None, None, None, None)
    assert rc == 1
    # panels beyond the register-resident limits are refused up front
    with pytest.raises(RuntimeError, match="limited to 20480"):
        _native.call("mnpbem_getrf_batched", _native.DTYPE_C128, 30000, None, 30000, 0, None, 30000, None, 1, None)
    with pytest.raises(RuntimeError, match="limited to 32768"):
        _native.call("mnpbem_getrf_batched", _native.DTYPE_C64, 40000, None, 40000, 0, None, 40000, None, 1, None)
    # ACA rank bound and H-matrix column limit
    with pytest.raises(RuntimeError, match="kmax"):
        _native.call("mnpbem_hmat_aca", 1, None, None, None, None, 1, 1, None, None, None, 600, 1e-6, None, None)
    with pytest.raises(RuntimeError, match="ncol"):
        _native.call("mnpbem_hmat_matvec", 2, 10, 1, 1, 1, *([None] * 7), 17, None, 17, None, 17, 0, None, None,
                     None)
    with pytest.raises(RuntimeError, match="order"):
        _native.call("mnpbem_refine_coeffs", 9, 0, *([None] * 11), 28, None, 144, None, 0, None, None, None, None)
    # recompression rank limits and tolerance
    with pytest.raises(RuntimeError, match="above 128"):
        _native.call("mnpbem_hmat_truncate", 2, 1, None, None, None, None, None, 129, 1e-6, None, None, None, None,
                     None)
    with pytest.raises(RuntimeError, match="XZ workspace"):
        _native.call("mnpbem_hmat_truncate", 2, 1, None, None, None, None, None, 80, 1e-6, None, None, None, None,
                     None)
    with pytest.raises(RuntimeError, match="htol"):
        _native.call("mnpbem_hmat_truncate", 2, 1, None, None, None, None, None, 8, 1.5, None, None, None, None,
                     None)
    with pytest.raises(RuntimeError, match="mode"):
        _native.call("mnpbem_hmat_diag", 2, 1, None, None, None, None, None, 3, None)


def test_gemm_tuning_table_roundtrip():
    """The bf16x3 tile table is host state: set / get round-trip, bad tables
    are refused, and the class of a shape follows x <= 64 * 2^c."""
    from mnpbem_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("libmnpbem_amd.so not built")
    from mnpbem_amd import autotune as at
    saved = at.get_table()
    try:
        t = np.zeros(at.NCELL, dtype=np.uint8)
        t[at.cell_of(64, 3000, 64)] = 1
        t[at.cell_of(65, 129, 4096)] = 1
        at.set_table(t)
        assert np.array_equal(at.get_table(), t)
        assert at.cell_of(64, 64, 64) == 0 and at.cell_of(65, 64, 64) == 49 and at.cell_of(10 ** 6, 1, 1) == 6 * 49
        t[at.cell_of(300, 300, 128)] = 2
        at.set_table(t)
        assert np.array_equal(at.get_table(), t)
        with pytest.raises(RuntimeError, match="0, 1 or 2"):
            at.set_table(np.full(at.NCELL, 3, dtype=np.uint8))
        with pytest.raises(ValueError):
            at.set_table(np.zeros(10, dtype=np.uint8))
        with pytest.raises(RuntimeError, match="variant"):
            _native.call("mnpbem_cgemm_sub_variant", 3, 1, 1, 1, None, 1, 1, None, 1, 1, None, 1, 1, 1, None)
    finally:
        at.set_table(saved)
