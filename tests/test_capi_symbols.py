"""CPU: the C-ABI library loads and exports every function include/shmoai.h declares."""
import ctypes

from shmoai_amd import capi


def test_header_symbols_exported():
    names = capi.header_functions()
    assert len(names) > 50
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(capi.header_functions()) == set(capi._SIGS)


def test_version_string():
    assert b"gfx950" in capi.load().shm_version()
