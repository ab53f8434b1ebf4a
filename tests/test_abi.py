"""CPU checks of the drop-in boundary: liblslab.so loads and exports every
symbol the headers in include/ declare (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

from lslab_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lslab_\w+)\s*\(", src, re.M)))


@pytest.fixture(scope="module")
def lib(lslab):
    return _lib.load()


def test_every_declared_symbol_is_exported(lib):
    names = _declared("lslab_c.h") + _declared("lslab_workload.h")
    assert len(names) >= 23
    for n in names:
        assert hasattr(lib, n), n
    assert set(_declared("lslab_c.h")) == set(_lib.EXPORTS)


def test_error_strings(lib):
    for code in range(0, 7):
        assert _lib.error_string(code) not in ("", "unknown status")
    assert _lib.error_string(99) == "unknown status"


def test_invalid_arguments_rejected_without_device_work(lib):
    h = ctypes.c_void_p()
    # widths other than 4/8 and n_log2 > 30 are refused before any allocation
    assert lib.lslab_map_create(ctypes.byref(h), 2, 8, 10, 10, None) == _lib.ERR_INVALID
    assert lib.lslab_map_create(ctypes.byref(h), 8, 8, 31, 10, None) == _lib.ERR_INVALID
    assert lib.lslab_set_create(ctypes.byref(h), 16, 10, 10, None) == _lib.ERR_INVALID
    bad = _lib.Tuning(12, 0, 256, 0, 1 << 20)
    assert lib.lslab_map_create(ctypes.byref(h), 8, 8, 10, 10, ctypes.byref(bad)) == _lib.ERR_INVALID
    bad = _lib.Tuning(4, 1, 256, 0, 1 << 20, 3)  # combining scope outside WAVE/BLOCK
    assert lib.lslab_map_create(ctypes.byref(h), 8, 8, 10, 10, ctypes.byref(bad)) == _lib.ERR_INVALID
    bad = _lib.Tuning(4, 1, 256, 0, 1 << 20, 2, 3)  # hash outside IDENTITY/MIX
    assert lib.lslab_map_create(ctypes.byref(h), 8, 8, 10, 10, ctypes.byref(bad)) == _lib.ERR_INVALID
    for spin in (5, 3600001):  # spin bound outside 10 ms .. 1 h
        bad = _lib.Tuning(4, 1, 256, 0, 1 << 20, 2, 0, 0, spin)
        assert lib.lslab_map_create(ctypes.byref(h), 8, 8, 10, 10, ctypes.byref(bad)) == _lib.ERR_INVALID
    bad = _lib.Tuning(4, 1, 256, 0, 1 << 20, 2, 0, 0, 0, 3)  # slot heads outside AUTO/ON/OFF
    assert lib.lslab_map_create(ctypes.byref(h), 8, 8, 10, 10, ctypes.byref(bad)) == _lib.ERR_INVALID
    assert lib.lslab_destroy(None) == _lib.ERR_INVALID
    assert lib.lslab_map_get(None, None, None, 0, None) == _lib.ERR_INVALID


def test_cpp_headers_compile():
    """The header-only C++ API compiles on its own (host-only translation unit)."""
    import subprocess
    import tempfile
    src = ('#include <cstddef>\n#include <lslab/map.h>\n#include <lslab/set.h>\n'
           '#include <lslab/wide.h>\n'
           'struct K20 { unsigned w[5]; };\n'
           'template class lslab::wide_map<K20, unsigned long long>;\n'
           'template class lslab::wide_set<K20>;\n'
           'int main(){ static_assert(sizeof(lslab::found_value<unsigned long long>)==16,"");'
           ' static_assert(sizeof(lslab::found_value<unsigned>)==8,"");'
           ' static_assert(sizeof(lslab::key_value<unsigned,unsigned long long>)==16,"");'
           f' static_assert(sizeof(lslab_tuning)=={ctypes.sizeof(_lib.Tuning)},"");'
           f' static_assert(offsetof(lslab_tuning, slot_heads)=={_lib.Tuning.slot_heads.offset},"");'
           f' static_assert(offsetof(lslab_tuning, runs)=={_lib.Tuning.runs.offset},"");'
           ' return 0; }\n')
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "t.cpp")
        open(f, "w").write(src)
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__",
                            "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", f],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_slab_bytes_closed_form(lib):
    """lslab_slab_bytes: (2^n + pool) slabs of the layout's size (a u32 set's
    128-B key line plus its 4-B next link); 0 for an invalid shape.  Pure
    arithmetic, no device call."""
    f = lib.lslab_slab_bytes
    assert f(1, 8, 8, 10, 100) == (1024 + 100) * 512
    assert f(1, 4, 4, 3, 0) == 8 * 256
    assert f(1, 4, 8, 3, 5) == 13 * 384
    assert f(1, 8, 4, 3, 5) == 13 * 384
    assert f(0, 8, 0, 4, 1) == 17 * 256
    assert f(0, 4, 0, 4, 1) == 17 * 132
    assert f(1, 2, 8, 4, 1) == 0
    assert f(0, 16, 0, 4, 1) == 0
    assert f(1, 8, 8, 31, 1) == 0
    assert f(1, 8, 8, 30, 1 << 32) == 0


def test_kvcg_and_info_reject_null_table(lib):
    """The KVCG and table-info entry points validate their handle before any
    device work."""
    import ctypes
    h = ctypes.c_void_p()
    assert lib.lslab_kvcg_create(ctypes.byref(h), None, 16, 0, 1) == 1
    assert lib.lslab_kvcg_create(None, None, 16, 0, 1) == 1
    assert lib.lslab_kvcg_destroy(None) == 1
    assert lib.lslab_kvcg_mixed(None, None, None, None, None, 0, None) == 1
    assert lib.lslab_kvcg_flush(None, None) == 1
    i = ctypes.c_int()
    assert lib.lslab_table_info(None, ctypes.byref(i), None, None, None) == 1
