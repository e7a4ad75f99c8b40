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


def test_bias_vec_known_answer():
    """bias_vec (Batch_encode_encrypt.cuh:45-57): out[j * num_X + i] = 1 for j < lengths[i]."""
    import shmoai_amd as shm
    got = shm.bias_vec([2, 0, 3], 3, 4)
    want = [0] * 12
    for i, ln in enumerate([2, 0, 3]):
        for j in range(ln):
            want[j * 3 + i] = 1
    assert list(got) == want
    assert [k for k, v in enumerate(got) if v] == [0, 2, 3, 5, 8]
    import pytest
    with pytest.raises(shm.capi.ShmInvalidArgument):
        shm.bias_vec([5, 0, 3], 3, 4)  # a length beyond num_row


def test_cpp_facade_header_compiles(tmp_path):
    """include/shmoai.hpp compiles as C++17 with every facade method instantiated by
    the facade test program (syntax and types only; no GPU)."""
    import os
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    assert gxx, "g++ is needed for the facade check"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(root, "include"), os.path.join(root, "tests", "cpp", "facade_c1.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_tuning_knobs_validate():
    """The process-wide knobs (no GPU work): defaults, accepted values, and refusals."""
    import pytest

    import shmoai_amd as shm
    from shmoai_amd.capi import ShmInvalidArgument
    assert shm.get_tuning("ntt_tile_threads") == 512
    assert shm.get_tuning("matmul_overlap") == 0
    assert shm.get_tuning("ks_batch_slots_log2") == -1
    try:
        shm.set_tuning("ntt_tile_threads", 256)
        assert shm.get_tuning("ntt_tile_threads") == 256
    finally:
        shm.set_tuning("ntt_tile_threads", 512)
    for name, value in (("ntt_tile_threads", 384), ("matmul_overlap", 2), ("ks_batch_slots_log2", 5),
                        ("ks_multi_slots_log2", -2), ("no_such_knob", 0)):
        with pytest.raises(ShmInvalidArgument):
            shm.set_tuning(name, value)
    with pytest.raises(ShmInvalidArgument):
        shm.get_tuning("no_such_knob")
    assert shm.get_tuning("ntt_tile_threads") == 512
