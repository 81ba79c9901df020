"""CPU checks of the C-ABI library: it loads, exports every function that
include/mtl4g.h declares, and its pure-host entry points answer.  No compute
calls (there is no GPU here)."""
import ctypes
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from mtl4_amd import _lib
    lib = _lib.lib()
    names = _lib.declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_symbols_are_c_linkage():
    lib = ctypes.CDLL(os.path.join(ROOT, "mtl4_amd", "libmtl4g.so"))
    for n in ("mtl4g_csr_create", "mtl4g_spmv", "mtl4g_cg", "mtl4g_bicgstab", "mtl4g_pc_create",
              "mtl4g_ir_solve"):
        assert getattr(lib, n) is not None


def test_status_strings():
    from mtl4_amd import _lib
    lib = _lib.lib()
    assert lib.mtl4g_status_string(0) == b"ok"
    assert lib.mtl4g_status_string(_lib.ERR_BREAKDOWN) == b"unexpected orthogonality"
    assert lib.mtl4g_status_string(_lib.ERR_MISSING_DIAG) == b"missing diagonal"
    assert lib.mtl4g_status_string(12345) == b"unknown status"
    assert lib.mtl4g_version() >= 100


def test_exception_mapping():
    from mtl4_amd import _lib
    with pytest.raises(_lib.unexpected_orthogonality):
        _lib.check(_lib.ERR_BREAKDOWN, "x")
    with pytest.raises(_lib.missing_diagonal):
        _lib.check(_lib.ERR_MISSING_DIAG, "x")
    assert _lib.check(1) == 1


def test_null_handles_rejected():
    from mtl4_amd import _lib
    lib = _lib.lib()
    assert lib.mtl4g_csr_info(None, None, None, None) == _lib.ERR_INVALID
    assert lib.mtl4g_pc_levels(None, None, None) == _lib.ERR_INVALID
    assert lib.mtl4g_csr_create(None, 1, 1, 0, None, None, None, 0) == _lib.ERR_INVALID


def test_sweep_schedule_and_chain_tile_arguments():
    """The schedule and tile setters check their arguments on the host."""
    from mtl4_amd import _lib
    lib = _lib.lib()
    assert lib.mtl4g_set_sweep_schedule(7) == _lib.ERR_INVALID
    for tx, ty in ((0, 4), (8, 4), (128, 1), (16, 32), (-64, -1)):
        assert lib.mtl4g_set_chain_tile(tx, ty) == _lib.ERR_INVALID, (tx, ty)
    for tx, ty in ((64, 1), (8, 8), (1, 64), (0, 0)):
        assert lib.mtl4g_set_chain_tile(tx, ty) == 0, (tx, ty)
    assert lib.mtl4g_pc_chain_tile(None, None, None, None, None) == _lib.ERR_INVALID
    assert lib.mtl4g_pc_sweeps(None, None, None) == _lib.ERR_INVALID


def test_csr_prepare_arguments():
    """mtl4g_csr_prepare rejects a null matrix and an unknown format on the
    host."""
    from mtl4_amd import _lib
    lib = _lib.lib()
    assert lib.mtl4g_csr_prepare(None, _lib.F32, None) == _lib.ERR_INVALID


def test_product_timing_arguments():
    """The product-timing diagnostic checks its arguments on the host."""
    from mtl4_amd import _lib
    lib = _lib.lib()
    assert lib.mtl4g_set_product_timing(2) == _lib.ERR_INVALID
    assert lib.mtl4g_product_timing(None, None) == _lib.ERR_INVALID
    assert lib.mtl4g_set_product_timing(0) == 0
