"""The C-ABI library loads on a CPU-only host and exports every symbol
declared in include/*.h (no compute calls without a GPU)."""
import ctypes

from blade_amd import _ffi


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_ffi.LIB_PATH)
    names = _ffi.exported_symbols()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in _ffi.lib().bm_version()
